// Small fully connected network (FP32) trained and evaluated on the GPU by
// the kernels of rcp_mlp.hip: the phi_J / schedule surrogate of the batched
// warm start.  Layers: n_in -> n_hidden (x n_layers, ReLU) -> n_out (linear).
// Parameters are one flat array, layer by layer: W (rows = outputs,
// row-major over inputs) then b.  Loss: mean squared error over the batch
// and the outputs.  Every reduction runs in a fixed order, so training is
// reproducible bit for bit on one device.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include <hip/hip_runtime_api.h>

namespace rcp {

class Mlp {
public:
    Mlp(int n_in, int n_hidden, int n_layers, int n_out, uint64_t seed);
    ~Mlp();
    Mlp(const Mlp&) = delete;
    Mlp& operator=(const Mlp&) = delete;

    size_t num_params() const { return n_params_; }
    int n_in() const { return n_in_; }
    int n_out() const { return n_out_; }
    void get_params(float* p) const;
    void set_params(const float* p);
    // host arrays: X[B * n_in] -> Y[B * n_out]
    void predict(const float* X, size_t B, float* Y);
    // host arrays: loss and the gradient of every parameter at X, T
    void gradients(const float* X, const float* T, size_t B, float* loss, float* grad);
    // Adam (beta 0.9 / 0.999, eps 1e-8) over minibatches of a shuffled
    // order (seeded); losses[e] = mean minibatch loss of epoch e
    void train(const float* X, const float* T, size_t n, size_t batch, size_t epochs, float lr,
               uint64_t seed, float* losses);

private:
    int n_in_, n_hidden_, n_layers_, n_out_;
    size_t n_params_ = 0;
    std::vector<size_t> w_off_, b_off_;    // per layer
    std::vector<int> fan_in_, fan_out_;
    hipStream_t s_ = nullptr;
    float *p_ = nullptr, *g_ = nullptr, *m_ = nullptr, *v_ = nullptr;
    // activations of the last forward (layer 0 = input) and their grads
    std::vector<float*> act_;
    float *dA_ = nullptr, *dB_ = nullptr;  // ping-pong gradient buffers
    float *part_ = nullptr;                // split-K partial sums
    float *loss_ = nullptr, *h_loss_ = nullptr;
    size_t cap_B_ = 0;
    uint64_t adam_t_ = 0;

    void ensure_batch(size_t B);
    void forward_dev(const float* X, size_t B);
    void backward_dev(const float* T, size_t B);
    void adam_dev(float lr);
};

}  // namespace rcp

// Launch interface between the host engine (compiled with g++) and the HIP
// kernels (compiled with hipcc for gfx950).  Everything here is plain C++
// over raw device pointers; no HIP types other than hipStream_t.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime_api.h>

namespace rcp {

constexpr int kMaxDim = 8;

// Scalar slots produced by one fused step (sums over the particles /
// pairs each block owns; the last block folds all block partials in block
// order, so the totals are deterministic run to run).
enum StepScalar : int {
    S_U = 0,      // sum of pair overlaps (j > i) + wall / curved overlap terms
    S_LC,         // sum of r_ij over contacts
    S_DK,         // sum r (dist - r) terms (pairs + walls + curved), without mu
    S_FM,         // sum of F_mag over contacts
    S_SUMD,       // sum of the (active) diameters
    S_SMAG,       // sum |F_i| over |F_i|^2 > 1e-16
    S_CNT,        // number of contacts (pairs + walls)
    S_MCNT,       // number of particles with |F_i|^2 > 1e-16
    S_MAXOV,      // max pair overlap
    S_MAXDISP2,   // max |x_new - x_anchor|^2 (minimum image)
    S_NUM
};

struct Geometry {
    int N = 0;
    int ndim = 0;
    double box[kMaxDim] = {};
    double inv_box[kMaxDim] = {};
    int8_t walls[kMaxDim] = {};
    int curved_T = 0;   // -walls[0] when walls[0] < 0, else 0
};

struct StepArgs {
    Geometry g;
    const double* x_in = nullptr;    // N*ndim, positions before the step
    double* x_out = nullptr;         // N*ndim, positions after ADAM + wrap
    double* m = nullptr;             // N*ndim
    double* v = nullptr;             // N*ndim
    const double* D0 = nullptr;      // N
    const int8_t* fixed = nullptr;   // N or null
    const double* x_anchor = nullptr;// N*ndim positions at the last list build
    const uint32_t* row_start = nullptr; // N+1
    const uint32_t* nbr = nullptr;   // full rows, ascending partner index
    double kappa = 1.0;              // D_i = D0_i * kappa (fixed: D0_i)
    int use_rk = 0;                  // cascade dkappa weights: rk_i = 0 if fixed, else D_i/2
    const double* rk = nullptr;      // explicit rk_i (overrides the computed weights)
    double alpha = 0.0, inv_b1 = 1.0, inv_b2 = 1.0;
    int do_adam = 1;                 // 0: forces only (x_out untouched)
    double* F_out = nullptr;         // optional N*ndim
    double* mag_out = nullptr;       // optional N: sqrt(|F|^2) or -1 (exact mode)
    double* contrib = nullptr;       // optional 4 doubles per row slot (exact mode)
    double* partials = nullptr;      // gridDim * S_NUM
    unsigned int* counter = nullptr; // last-block ticket
    double* result = nullptr;        // S_NUM totals
};

// One fused step: forces + walls + reductions + (optional) ADAM/wrap.
int launch_step(const StepArgs& a, hipStream_t s);
// Sequential (reference 1-thread order) scalar reduction for exact mode.
int launch_exact_reduce(const StepArgs& a, hipStream_t s);

// ---------------------------------------------------------------- lists
struct CellGrid {
    int nc[kMaxDim] = {};     // cells per axis (1 = axis not split)
    double cell_scale[kMaxDim] = {};  // nc / L
    long long total = 1;
};

struct ListArgs {
    Geometry g;
    const double* x = nullptr;
    const double* D0 = nullptr;
    const int8_t* fixed = nullptr;
    double kappa = 1.0;
    double t_shell = 0.0;          // list cutoff rc = (Di+Dj)*0.5 + t_shell
    int half = 0;                  // 1: keep only j > i (pair-set API)
    CellGrid grid;
    uint32_t* cell_of = nullptr;   // N
    uint32_t* slot = nullptr;      // N
    uint32_t* cell_count = nullptr;// total+1
    uint32_t* cell_start = nullptr;// total+1
    uint32_t* items = nullptr;     // N
    uint32_t* row_count = nullptr; // N+1
    uint32_t* row_start = nullptr; // N+1
    uint32_t* nbr = nullptr;       // capacity nbr_cap
    unsigned long long nbr_cap = 0;
};

int launch_cells(ListArgs& a, void* scan_tmp, size_t scan_tmp_bytes, hipStream_t s);
int launch_rows_count(ListArgs& a, hipStream_t s);
int launch_rows_fill(ListArgs& a, hipStream_t s);

// exclusive scan of n u32 values (out may alias in); out[n] = total
size_t scan_tmp_bytes(size_t n);
int launch_scan_u32(const uint32_t* in, uint32_t* out, size_t n, void* tmp, hipStream_t s);

// ---------------------------------------------------------------- reorder
struct SortTmp {
    unsigned long long* keys_alt = nullptr;
    uint32_t* vals_alt = nullptr;
    uint32_t* hist = nullptr;      // 256 * nblocks + 1
    void* scan_tmp = nullptr;
};
size_t sort_hist_entries(size_t n);
// keys: Morton-21 code (ndim == 3) or the ordered bits of x[i*ndim] (other
// ndim); vals = identity; then a stable LSD radix sort.
int launch_reorder_keys(const double* x, int N, int ndim, const double* box,
                        unsigned long long* keys, uint32_t* vals, hipStream_t s);
int launch_radix_sort(unsigned long long* keys, uint32_t* vals, size_t n, int key_bits,
                      SortTmp& t, hipStream_t s);
// dst[k*w + d] = src[perm[k]*w + d]
int launch_gather_f64(const double* src, double* dst, const uint32_t* perm, int N, int w,
                      hipStream_t s);
int launch_gather_u64(const unsigned long long* src, unsigned long long* dst,
                      const uint32_t* perm, int N, hipStream_t s);
int launch_gather_i8(const int8_t* src, int8_t* dst, const uint32_t* perm, int N,
                     hipStream_t s);

// ---------------------------------------------------------------- small
int launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
int launch_iota_u64(unsigned long long* p, size_t n, hipStream_t s);
int launch_scale_last_dim(double* x, int N, int ndim, double f, hipStream_t s);
// D[i] = fixed[i] ? D0[i] : D0[i] * kappa
int launch_diameters(const double* D0, const int8_t* fixed, double kappa, double* D, int N,
                     hipStream_t s);
// Standalone per-DOF ADAM + position + wrap over supplied forces.
int launch_adam_pass(const Geometry& g, const double* F, double* m, double* v, const double* x,
                     double* xo, double alpha, double inv_b1, double inv_b2, hipStream_t s);
// Mean-force numerator of supplied forces into a.result[S_SMAG], [S_MCNT]
// (a.partials / a.counter / a.result / a.mag_out used); exact = particle order.
int launch_mean_force(const StepArgs& a, const double* F, int exact, hipStream_t s);
// rk[i] = D_i / 2 with D_i = fixed[i] ? D0[i] : D0[i]*kappa (all particles active)
int launch_rk_active(const double* D0, const int8_t* fixed, double kappa, double* rk, int N,
                     hipStream_t s);
// x_anchor = x
int launch_copy_f64(const double* src, double* dst, size_t n, hipStream_t s);

}  // namespace rcp

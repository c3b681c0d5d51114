// HIP kernels of the MI355X (gfx950) packing engine.
//
// Numerics follow the reference engine's 1-thread evaluation order
// (KD-physics/RCPGenerator python_code/python/src/core/rcp_generator.cpp,
// cited as RG:<line>): each particle sums its contact forces over partners in
// ascending index order, which is the order in which the reference's
// row-ordered half-list loop touches F_i at one OpenMP thread (RG:4086-4139,
// F_i += for rows i, F_j -= for rows j < i).  Built with -ffp-contract=off so
// no multiply-add is fused; sqrt and division are the correctly rounded
// IEEE operations.  FP64 throughout, no MFMA (nothing here is a contraction).
#include <hip/hip_runtime.h>

#include "rcp_kernels.h"

namespace rcp {

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr double kBeta1 = 0.9;
constexpr double kBeta2 = 0.999;
constexpr double kEps = 1e-8;

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = kWave / 2; o > 0; o >>= 1) {
        double w = __shfl_xor(v, o, kWave);
        v = (w > v) ? w : v;
    }
    return v;
}

// Read lane k's double (k uniform across the wave).
__device__ __forceinline__ double readlane_f64(double v, int k)
{
    unsigned long long b = __double_as_longlong(v);
    unsigned lo = __builtin_amdgcn_readlane((unsigned)(b & 0xffffffffull), k);
    unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), k);
    return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ int readlane_i32(int v, int k)
{
    return __builtin_amdgcn_readlane(v, k);
}

// One DOF of the fused ADAM + position + wrap pass (RG:6507-6518).
__device__ __forceinline__ double adam_dof(double Fk, double& m, double& v, double x, double L,
                                           double alpha, double inv_b1, double inv_b2)
{
    const double mk = kBeta1 * m - (1.0 - kBeta1) * Fk;
    const double vk = kBeta2 * v + (1.0 - kBeta2) * Fk * Fk;
    m = mk;
    v = vk;
    const double mh = mk * inv_b1;
    const double vh = vk * inv_b2;
    double xv = x - alpha * mh / (sqrt(vh) + kEps);
    if (xv >= L) xv -= L;
    else if (xv < 0) xv += L;
    return xv;
}

__device__ __forceinline__ double mic(double delta, double inv_box, double box)
{
    // delta -= nearbyint(delta * inv_box) * box   (RG:4097-4101)
    return delta - rint(delta * inv_box) * box;
}

// Block-wide deterministic reduction of S_NUM scalars (sums except the two
// max slots); thread 0 of the block receives the result in out[].
__device__ void block_reduce(double (&v)[S_NUM], double* out)
{
    __shared__ double sh[kBlock / kWave][S_NUM];
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < S_NUM; ++k)
        v[k] = (k == S_MAXOV || k == S_MAXDISP2) ? wave_max(v[k]) : wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < S_NUM; ++k) sh[w][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < S_NUM; ++k) {
            double acc = sh[0][k];
            for (int q = 1; q < (int)(blockDim.x / kWave); ++q) {
                if (k == S_MAXOV || k == S_MAXDISP2) acc = (sh[q][k] > acc) ? sh[q][k] : acc;
                else acc += sh[q][k];
            }
            out[k] = acc;
        }
    }
    __syncthreads();
}

// The last block to finish folds all block partials (block order) into
// a.result and re-arms the ticket.
__device__ void finish_grid(const StepArgs& a, double (&blk)[S_NUM])
{
    __shared__ int is_last;
    if (threadIdx.x == 0) {
        for (int k = 0; k < S_NUM; ++k) a.partials[(size_t)blockIdx.x * S_NUM + k] = blk[k];
        unsigned int ticket = __hip_atomic_fetch_add(a.counter, 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        is_last = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double v[S_NUM];
    for (int k = 0; k < S_NUM; ++k) v[k] = 0.0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
        const double* p = a.partials + (size_t)b * S_NUM;
#pragma unroll
        for (int k = 0; k < S_NUM; ++k) {
            double t = __builtin_nontemporal_load(p + k);
            if (k == S_MAXOV || k == S_MAXDISP2) v[k] = (t > v[k]) ? t : v[k];
            else v[k] += t;
        }
    }
    double tot[S_NUM];
    block_reduce(v, tot);
    if (threadIdx.x == 0) {
        for (int k = 0; k < S_NUM; ++k) a.result[k] = tot[k];
        __hip_atomic_store(a.counter, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// -------------------------------------------------------------- fused step
//
// Per particle i (one thread):
//   contact forces over the ascending full row (RG:4086-4139 contact law),
//   flat walls and the curved container (RG:4198-4257), |F_i| for the mean
//   force (RG:769-803), then the fused per-DOF ADAM + position + wrap
//   (RG:6499-6520) into x_out, and the squared minimum-image displacement
//   from the list anchor.
template <int ND>
__global__ __launch_bounds__(kBlock) void k_step(StepArgs a)
{
    const int N = a.g.N;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[S_NUM];
#pragma unroll
    for (int k = 0; k < S_NUM; ++k) acc[k] = 0.0;

    if (i < N) {
        double xi[ND], F[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) { xi[d] = a.x_in[(size_t)i * ND + d]; F[d] = 0.0; }
        const bool fi = a.fixed && a.fixed[i];
        const double D0i = a.D0[i];
        const double Di = fi ? D0i : D0i * a.kappa;
        const double rki = a.rk ? a.rk[i] : (fi ? 0.0 : Di * 0.5);
        const uint32_t e0 = a.row_start[i], e1 = a.row_start[i + 1];
        for (uint32_t e = e0; e < e1; ++e) {
            const uint32_t j = a.nbr[e];
            const double* xj = a.x_in + (size_t)j * ND;
            const bool fj = a.fixed && a.fixed[j];
            const double D0j = a.D0[j];
            const double Dj = fj ? D0j : D0j * a.kappa;
            const double r = 0.5 * (Di + Dj);
            double dl[ND];
            double d2 = 0.0;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                double delta = xj[d] - xi[d];
                delta = mic(delta, a.g.inv_box[d], a.g.box[d]);
                dl[d] = delta;
                d2 += delta * delta;
            }
            bool valid = false;
            double ov = 0.0, dkt = 0.0, fm = 0.0;
            if (d2 < r * r) {
                const double dist = sqrt(d2);
                fm = -(r / dist - 1.0);
#pragma unroll
                for (int d = 0; d < ND; ++d) F[d] += fm * dl[d];
                if (j > (uint32_t)i) {
                    valid = true;
                    ov = 1.0 - dist / r;
                    acc[S_U] += ov;
                    acc[S_CNT] += 1.0;
                    acc[S_LC] += r;
                    acc[S_MAXOV] = (ov > acc[S_MAXOV]) ? ov : acc[S_MAXOV];
                    const double rkj = a.rk ? a.rk[j] : (fj ? 0.0 : Dj * 0.5);
                    dkt = (a.use_rk ? (rki + rkj) : r) * (dist - r);
                    acc[S_DK] = acc[S_DK] + dkt;
                    acc[S_FM] = acc[S_FM] + fm;
                }
            }
            if (a.contrib) {
                double* c = a.contrib + (size_t)e * 4;
                c[0] = ov;
                c[1] = valid ? r : 0.0;
                c[2] = dkt;
                c[3] = fm;
            }
        }
        // Flat walls (RG:4200-4220); walls[d] == 1 only.
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            if (a.g.walls[d] == 1) {
                const double rw = Di / 2;
                for (int wall = 0; wall < 2; ++wall) {
                    const double dxw = a.g.box[d] * (double)wall - xi[d];
                    const double dist = fabs(dxw);
                    if (dist < rw) {
                        const double fmw = -2.0 * (rw / dist - 1.0);
                        acc[S_U] += 2.0 * (1.0 - dist / rw);
                        acc[S_CNT] += 1.0;
                        F[d] += fmw * dxw;
                        acc[S_DK] = acc[S_DK] + (2.0 * (a.use_rk ? rki : rw)) * (dist - rw);
                        acc[S_FM] = acc[S_FM] + fmw;
                    }
                }
            }
        }
        // Curved container over the first T axes (RG:4222-4257).
        if (a.g.curved_T > 0) {
            const int T = a.g.curved_T;
            const double rc = Di / 2;
            const double R = a.g.box[0] / 2;
            double dij = 0;
#pragma unroll
            for (int M = 0; M < ND; ++M)
                if (M < T) dij = dij + (xi[M] - R) * (xi[M] - R);
            dij = sqrt(dij);
            double delta = R - dij;
            if (delta < rc) {
                delta = dij;
                dij = 0;
                double dx[ND];
#pragma unroll
                for (int M = 0; M < ND; ++M) {
                    dx[M] = 0.0;
                    if (M < T) {
                        dx[M] = fabs(R / delta - 1) * (xi[M] - R);
                        dij = dij + dx[M] * dx[M];
                    }
                }
                dij = sqrt(dij);
                const double fmc = -2.0 * fabs(rc / dij - 1.0);
                acc[S_U] = acc[S_U] + (1 - dij / rc);
                acc[S_CNT] += 1.0;
#pragma unroll
                for (int M = 0; M < ND; ++M)
                    if (M < T) F[M] = F[M] + fmc * dx[M];
                acc[S_DK] = acc[S_DK] + (2.0 * (a.use_rk ? rki : rc)) * (dij - rc);
                acc[S_FM] = acc[S_FM] + fmc;
            }
        }
        // Mean-force contribution (RG:784-796).
        double f2 = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d) f2 += F[d] * F[d];
        double mag = -1.0;
        if (f2 > 1.0E-16) {
            mag = sqrt(f2);
            acc[S_SMAG] += mag;
            acc[S_MCNT] += 1.0;
        }
        if (a.mag_out) a.mag_out[i] = mag;
        acc[S_SUMD] = (a.use_rk && !a.rk && fi) ? 0.0 : Di;
        if (a.F_out) {
#pragma unroll
            for (int d = 0; d < ND; ++d) a.F_out[(size_t)i * ND + d] = F[d];
        }
        if (a.do_adam) {
            double d2a = 0.0;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const size_t idx = (size_t)i * ND + d;
                double mm = a.m[idx], vv = a.v[idx];
                const double xv = adam_dof(F[d], mm, vv, xi[d], a.g.box[d], a.alpha, a.inv_b1,
                                           a.inv_b2);
                a.m[idx] = mm;
                a.v[idx] = vv;
                a.x_out[idx] = xv;
                double dd = xv - a.x_anchor[idx];
                dd = mic(dd, a.g.inv_box[d], a.g.box[d]);
                d2a += dd * dd;
            }
            acc[S_MAXDISP2] = d2a;
        }
    }
    double blk[S_NUM];
    block_reduce(acc, blk);
    finish_grid(a, blk);
}

// -------------------------------------------------------- exact reduction
//
// One wavefront re-sums the scalar terms in the reference's 1-thread order:
// pair terms row by row (rows ascending, partners ascending, j > i only;
// RG:4106-4138), then walls particle by particle (RG:4198-4257), then
// sum(D) (RG:4262-4265) and the mean-force numerator (RG:784-796).  Terms
// are computed in parallel (k_step stored the pair terms per row slot) and
// folded by lane 0..63 readlane in order, so every addition rounds exactly
// as the reference's sequential loop does.
template <int ND>
__global__ __launch_bounds__(kWave) void k_exact_reduce(StepArgs a)
{
    const int lane = threadIdx.x;
    const int N = a.g.N;
    double U = 0.0, Lc = 0.0, dk = 0.0, Fm = 0.0, maxov = 0.0, cnt = 0.0;
    const uint32_t total = a.row_start[N];
    for (uint32_t base = 0; base < total; base += kWave) {
        const uint32_t e = base + lane;
        double ov = 0.0, r = 0.0, dkt = 0.0, fm = 0.0;
        if (e < total) {
            const double* c = a.contrib + (size_t)e * 4;
            ov = c[0]; r = c[1]; dkt = c[2]; fm = c[3];
        }
        const int lim = (total - base < (uint32_t)kWave) ? (int)(total - base) : kWave;
        for (int k = 0; k < lim; ++k) {
            const double rk = readlane_f64(r, k);
            if (rk != 0.0) {
                const double ok = readlane_f64(ov, k);
                U += ok;
                cnt += 1.0;
                Lc += rk;
                if (ok > maxov) maxov = ok;
                dk = dk + readlane_f64(dkt, k);
                Fm = Fm + readlane_f64(fm, k);
            }
        }
    }
    // walls and curved container, in particle order
    constexpr int NT = 2 * ND + 1;
    for (int base = 0; base < N; base += kWave) {
        const int i = base + lane;
        double tu[NT], td[NT], tf[NT];
        int tv[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) { tu[t] = td[t] = tf[t] = 0.0; tv[t] = 0; }
        if (i < N) {
            double xi[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d) xi[d] = a.x_in[(size_t)i * ND + d];
            const bool fi = a.fixed && a.fixed[i];
            const double Di = fi ? a.D0[i] : a.D0[i] * a.kappa;
            const double rki = a.rk ? a.rk[i] : (fi ? 0.0 : Di * 0.5);
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                if (a.g.walls[d] == 1) {
                    const double rw = Di / 2;
#pragma unroll
                    for (int wall = 0; wall < 2; ++wall) {
                        const double dxw = a.g.box[d] * (double)wall - xi[d];
                        const double dist = fabs(dxw);
                        if (dist < rw) {
                            const int t = 2 * d + wall;
                            tv[t] = 1;
                            tf[t] = -2.0 * (rw / dist - 1.0);
                            tu[t] = 2.0 * (1.0 - dist / rw);
                            td[t] = (2.0 * (a.use_rk ? rki : rw)) * (dist - rw);
                        }
                    }
                }
            }
            if (a.g.curved_T > 0) {
                const int T = a.g.curved_T;
                const double rc = Di / 2;
                const double R = a.g.box[0] / 2;
                double dij = 0;
#pragma unroll
                for (int M = 0; M < ND; ++M)
                    if (M < T) dij = dij + (xi[M] - R) * (xi[M] - R);
                dij = sqrt(dij);
                double delta = R - dij;
                if (delta < rc) {
                    delta = dij;
                    dij = 0;
#pragma unroll
                    for (int M = 0; M < ND; ++M) {
                        if (M < T) {
                            const double dxm = fabs(R / delta - 1) * (xi[M] - R);
                            dij = dij + dxm * dxm;
                        }
                    }
                    dij = sqrt(dij);
                    tv[NT - 1] = 1;
                    tf[NT - 1] = -2.0 * fabs(rc / dij - 1.0);
                    tu[NT - 1] = (1 - dij / rc);
                    td[NT - 1] = (2.0 * (a.use_rk ? rki : rc)) * (dij - rc);
                }
            }
        }
        const int lim = (N - base < kWave) ? (N - base) : kWave;
        for (int k = 0; k < lim; ++k) {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                if (readlane_i32(tv[t], k)) {
                    U += readlane_f64(tu[t], k);
                    cnt += 1.0;
                    dk = dk + readlane_f64(td[t], k);
                    Fm = Fm + readlane_f64(tf[t], k);
                }
            }
        }
    }
    // sum(D) or the active-only sum, then the mean-force numerator
    double sumD = 0.0, smag = 0.0, mcnt = 0.0;
    for (int base = 0; base < N; base += kWave) {
        const int i = base + lane;
        double Di = 0.0, mag = -1.0;
        int act = 0;
        if (i < N) {
            const bool fi = a.fixed && a.fixed[i];
            Di = fi ? a.D0[i] : a.D0[i] * a.kappa;
            act = !(a.use_rk && !a.rk && fi);
            mag = a.mag_out[i];
        }
        const int lim = (N - base < kWave) ? (N - base) : kWave;
        for (int k = 0; k < lim; ++k) {
            if (readlane_i32(act, k)) sumD += readlane_f64(Di, k);
            const double mk = readlane_f64(mag, k);
            if (mk >= 0.0) { smag += mk; mcnt += 1.0; }
        }
    }
    if (lane == 0) {
        a.result[S_U] = U;
        a.result[S_LC] = Lc;
        a.result[S_DK] = dk;
        a.result[S_FM] = Fm;
        a.result[S_SUMD] = sumD;
        a.result[S_SMAG] = smag;
        a.result[S_CNT] = cnt;
        a.result[S_MCNT] = mcnt;
        a.result[S_MAXOV] = maxov;
    }
}

// ------------------------------------------------------------ cell lists

template <int ND>
__global__ __launch_bounds__(kBlock) void k_cell_assign(ListArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.g.N) return;
    long long c = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const double L = a.g.box[d];
        double xw = a.x[(size_t)i * ND + d];
        xw = xw - L * floor(xw / L);
        int k = (int)floor(xw * a.grid.cell_scale[d]);
        if (k < 0) k = 0;
        if (k >= a.grid.nc[d]) k = a.grid.nc[d] - 1;
        c += (long long)k * stride;
        stride *= a.grid.nc[d];
    }
    a.cell_of[i] = (uint32_t)c;
    a.slot[i] = atomicAdd(&a.cell_count[c], 1u);
}

__global__ __launch_bounds__(kBlock) void k_cell_scatter(ListArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.g.N) return;
    a.items[a.cell_start[a.cell_of[i]] + a.slot[i]] = (uint32_t)i;
}

// Candidate rows.  Predicate: include j iff !(d2 >= rc*rc) with
// rc = (Di + Dj) * 0.5 + t_shell (the kd-tree test, RG:3276-3297), MIC on
// every axis.  FILL=false counts, FILL=true writes and sorts the row.
template <int ND, bool FILL>
__global__ __launch_bounds__(kBlock) void k_rows(ListArgs a)
{
    const int N = a.g.N;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double xi[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) xi[d] = a.x[(size_t)i * ND + d];
    const bool fi = a.fixed && a.fixed[i];
    const double Di = fi ? a.D0[i] : a.D0[i] * a.kappa;
    int ci[ND], span[ND];
    {
        uint32_t c = a.cell_of[i];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            ci[d] = (int)(c % (uint32_t)a.grid.nc[d]);
            c /= (uint32_t)a.grid.nc[d];
            span[d] = a.grid.nc[d] >= 3 ? 3 : 1;
        }
    }
    int nst = 1;
#pragma unroll
    for (int d = 0; d < ND; ++d) nst *= span[d];
    uint32_t count = 0;
    const uint32_t base = FILL ? a.row_start[i] : 0;
    for (int s = 0; s < nst; ++s) {
        int t = s;
        long long cc = 0, stride = 1;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            int o = 0;
            if (span[d] == 3) { o = t % 3 - 1; t /= 3; }
            int k = ci[d] + o;
            if (k < 0) k += a.grid.nc[d];
            if (k >= a.grid.nc[d]) k -= a.grid.nc[d];
            cc += (long long)k * stride;
            stride *= a.grid.nc[d];
        }
        const uint32_t q0 = a.cell_start[cc], q1 = a.cell_start[cc + 1];
        for (uint32_t q = q0; q < q1; ++q) {
            const uint32_t j = a.items[q];
            if (j == (uint32_t)i) continue;
            if (a.half && j < (uint32_t)i) continue;
            const bool fj = a.fixed && a.fixed[j];
            const double Dj = fj ? a.D0[j] : a.D0[j] * a.kappa;
            const double rc = (Di + Dj) * 0.5 + a.t_shell;
            double d2 = 0.0;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                double dd = a.x[(size_t)j * ND + d] - xi[d];
                dd = mic(dd, a.g.inv_box[d], a.g.box[d]);
                d2 += dd * dd;
            }
            if (d2 >= rc * rc) continue;
            if (FILL) a.nbr[base + count] = j;
            ++count;
        }
    }
    if (!FILL) {
        a.row_count[i] = count;
    } else {
        uint32_t* row = a.nbr + base;
        for (uint32_t p = 1; p < count; ++p) {
            const uint32_t key = row[p];
            uint32_t q = p;
            while (q > 0 && row[q - 1] > key) { row[q] = row[q - 1]; --q; }
            row[q] = key;
        }
    }
}

// ------------------------------------------------------------------ scan

constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;

__global__ __launch_bounds__(kBlock) void k_scan_tile(const uint32_t* in, uint32_t* out,
                                                       size_t n, uint32_t* bsum)
{
    __shared__ uint32_t wsum[kBlock / kWave];
    const size_t b0 = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t tsum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (b0 + k < n) ? in[b0 + k] : 0u;
        tsum += v[k];
    }
    // inclusive wave scan of thread sums
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    uint32_t incl = tsum;
    for (int o = 1; o < kWave; o <<= 1) {
        uint32_t y = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += y;
    }
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (int q = 0; q < w; ++q) woff += wsum[q];
    uint32_t run = woff + incl - tsum;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (b0 + k < n) out[b0 + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kBlock - 1) {
        uint32_t tot = 0;
        for (int q = 0; q < kBlock / kWave; ++q) tot += wsum[q];
        bsum[blockIdx.x] = tot;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* out, size_t n,
                                                      const uint32_t* boff)
{
    const size_t b0 = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    const uint32_t add = boff[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (b0 + k < n) out[b0 + k] += add;
}

__global__ void k_set_total(uint32_t* out, size_t n, const uint32_t* src)
{
    out[n] = src[0];
}

// ----------------------------------------------------------- radix sort

__device__ __forceinline__ unsigned long long ordered_bits(double x)
{
    // order-preserving map of a double to u64; -0.0 and +0.0 compare equal
    // in the reference comparator, so both map to the +0.0 key.
    if (x == 0.0) x = 0.0;
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}

__device__ __forceinline__ unsigned long long morton_split21(unsigned long long v)
{
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x001F00000000FFFFull;
    v = (v | (v << 16)) & 0x001F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ __launch_bounds__(kBlock) void k_reorder_keys(const double* x, int N, int ndim,
                                                          double b0, double b1, double b2,
                                                          unsigned long long* keys,
                                                          uint32_t* vals)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    unsigned long long key;
    if (ndim == 3) {
        // RG:1127-1150: clamp to [0, box - 1e-12], quantise to 21 bits
        const double box[3] = {b0, b1, b2};
        unsigned long long code = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double scale = ((double)(1u << 21) - 1.0) / box[d];
            const double hi = box[d] - 1e-12, v = x[(size_t)i * 3 + d];
            const double m = (v < hi) ? v : hi;
            const double c = (0.0 < m) ? m : 0.0;
            const unsigned q = (unsigned)(c * scale);
            code |= morton_split21(q) << d;
        }
        key = code;
    } else {
        key = ordered_bits(x[(size_t)i * ndim]);
    }
    keys[i] = key;
    vals[i] = (uint32_t)i;
}

constexpr int kSortItems = 4;
constexpr int kSortTile = kBlock * kSortItems;

__global__ __launch_bounds__(kBlock) void k_radix_hist(const unsigned long long* keys, size_t n,
                                                        int shift, uint32_t* hist, int nblocks)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t t0 = (size_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const size_t e = t0 + (size_t)k * kBlock + threadIdx.x;
        if (e < n) atomicAdd(&h[(keys[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: elements of the tile are ranked in tile order (striped
// rounds of 256, wave ballots for the in-wave rank).
__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const unsigned long long* keys, const uint32_t* vals, unsigned long long* keys_out,
    uint32_t* vals_out, size_t n, int shift, const uint32_t* hist_scan, int nblocks)
{
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kBlock / kWave][256];
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    run[threadIdx.x] = hist_scan[(size_t)threadIdx.x * nblocks + blockIdx.x];
    const size_t t0 = (size_t)blockIdx.x * kSortTile;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int k = 0; k < kSortItems; ++k) {
#pragma unroll
        for (int q = 0; q < kBlock / kWave; ++q) wcnt[q][threadIdx.x] = 0;
        __syncthreads();
        const size_t e = t0 + (size_t)k * kBlock + threadIdx.x;
        const bool ok = e < n;
        unsigned long long key = ok ? keys[e] : 0ull;
        const uint32_t dig = ok ? (uint32_t)((key >> shift) & 255u) : 256u;
        unsigned long long mask = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const bool bit = (dig >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            mask &= bit ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(mask & lt);
        if (ok && rank == 0) wcnt[w][dig] = (uint32_t)__popcll(mask);
        __syncthreads();
        if (ok) {
            uint32_t off = run[dig] + rank;
            for (int q = 0; q < w; ++q) off += wcnt[q][dig];
            keys_out[off] = key;
            vals_out[off] = vals[e];
        }
        __syncthreads();
        {
            uint32_t add = 0;
#pragma unroll
            for (int q = 0; q < kBlock / kWave; ++q) add += wcnt[q][threadIdx.x];
            run[threadIdx.x] += add;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------- standalone passes

__global__ __launch_bounds__(kBlock) void k_adam_pass(Geometry g, const double* F, double* m,
                                                      double* v, const double* x, double* xo,
                                                      double alpha, double inv_b1, double inv_b2)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)g.N * g.ndim) return;
    const int d = (int)(t % g.ndim);
    double mm = m[t], vv = v[t];
    xo[t] = adam_dof(F[t], mm, vv, x[t], g.box[d], alpha, inv_b1, inv_b2);
    m[t] = mm;
    v[t] = vv;
}

// Mean-force numerator (RG:784-796): block sums folded by the last block.
__global__ __launch_bounds__(kBlock) void k_mean_force(StepArgs a, const double* F)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[S_NUM];
#pragma unroll
    for (int k = 0; k < S_NUM; ++k) acc[k] = 0.0;
    if (i < a.g.N) {
        double f2 = 0.0;
        for (int d = 0; d < a.g.ndim; ++d) {
            const double f = F[(size_t)i * a.g.ndim + d];
            f2 += f * f;
        }
        double mag = -1.0;
        if (f2 > 1.0E-16) {
            mag = sqrt(f2);
            acc[S_SMAG] = mag;
            acc[S_MCNT] = 1.0;
        }
        if (a.mag_out) a.mag_out[i] = mag;
    }
    double blk[S_NUM];
    block_reduce(acc, blk);
    finish_grid(a, blk);
}

// The same numerator folded sequentially in particle order (one wave).
__global__ __launch_bounds__(kWave) void k_mean_force_exact(StepArgs a)
{
    const int lane = threadIdx.x;
    double smag = 0.0, mcnt = 0.0;
    for (int base = 0; base < a.g.N; base += kWave) {
        const int i = base + lane;
        const double mag = (i < a.g.N) ? a.mag_out[i] : -1.0;
        const int lim = (a.g.N - base < kWave) ? (a.g.N - base) : kWave;
        for (int k = 0; k < lim; ++k) {
            const double mk = readlane_f64(mag, k);
            if (mk >= 0.0) { smag += mk; mcnt += 1.0; }
        }
    }
    if (lane == 0) {
        a.result[S_SMAG] = smag;
        a.result[S_MCNT] = mcnt;
    }
}

// ----------------------------------------------------------------- small

__global__ void k_gather_f64(const double* src, double* dst, const uint32_t* perm, int N, int w)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * w) return;
    const size_t k = t / w, d = t % w;
    dst[t] = src[(size_t)perm[k] * w + d];
}
__global__ void k_gather_u64(const unsigned long long* src, unsigned long long* dst,
                             const uint32_t* perm, int N)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N) dst[k] = src[perm[k]];
}
__global__ void k_gather_i8(const int8_t* src, int8_t* dst, const uint32_t* perm, int N)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N) dst[k] = src[perm[k]];
}
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}
__global__ void k_iota_u64(unsigned long long* p, size_t n)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = k;
}
__global__ void k_scale_last(double* x, int N, int ndim, double f)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N) x[(size_t)k * ndim + ndim - 1] = x[(size_t)k * ndim + ndim - 1] * f;
}
__global__ void k_diameters(const double* D0, const int8_t* fixed, double kappa, double* D, int N)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N) D[k] = (fixed && fixed[k]) ? D0[k] : D0[k] * kappa;
}
__global__ void k_rk_active(const double* D0, const int8_t* fixed, double kappa, double* rk,
                            int N)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N) rk[k] = ((fixed && fixed[k]) ? D0[k] : D0[k] * kappa) * 0.5;
}
__global__ void k_copy_f64(const double* src, double* dst, size_t n)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

inline unsigned grid_for(size_t n, int block = kBlock)
{
    return (unsigned)((n + block - 1) / block);
}

inline int check_launch()
{
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace

// ============================================================== launchers

#define RCP_DISPATCH_ND(nd, KERNEL, ...)                                               \
    switch (nd) {                                                                      \
        case 1: KERNEL<1> __VA_ARGS__; break;                                          \
        case 2: KERNEL<2> __VA_ARGS__; break;                                          \
        case 3: KERNEL<3> __VA_ARGS__; break;                                          \
        case 4: KERNEL<4> __VA_ARGS__; break;                                          \
        case 5: KERNEL<5> __VA_ARGS__; break;                                          \
        case 6: KERNEL<6> __VA_ARGS__; break;                                          \
        case 7: KERNEL<7> __VA_ARGS__; break;                                          \
        case 8: KERNEL<8> __VA_ARGS__; break;                                          \
        default: return 2;                                                             \
    }

int launch_step(const StepArgs& a, hipStream_t s)
{
    if (a.g.N <= 0) return 0;
    const unsigned g = grid_for(a.g.N);
    RCP_DISPATCH_ND(a.g.ndim, k_step, <<<g, kBlock, 0, s>>>(a));
    return check_launch();
}

int launch_exact_reduce(const StepArgs& a, hipStream_t s)
{
    if (a.g.N <= 0) return 0;
    RCP_DISPATCH_ND(a.g.ndim, k_exact_reduce, <<<1, kWave, 0, s>>>(a));
    return check_launch();
}

int launch_cells(ListArgs& a, void* scan_tmp, size_t, hipStream_t s)
{
    const size_t total = (size_t)a.grid.total;
    k_fill_u32<<<grid_for(total + 1), kBlock, 0, s>>>(a.cell_count, 0u, total + 1);
    const unsigned g = grid_for(a.g.N);
    RCP_DISPATCH_ND(a.g.ndim, k_cell_assign, <<<g, kBlock, 0, s>>>(a));
    if (launch_scan_u32(a.cell_count, a.cell_start, total, scan_tmp, s)) return 1;
    k_cell_scatter<<<g, kBlock, 0, s>>>(a);
    return check_launch();
}

int launch_rows_count(ListArgs& a, hipStream_t s)
{
    const unsigned g = grid_for(a.g.N);
    switch (a.g.ndim) {
        case 1: k_rows<1, false><<<g, kBlock, 0, s>>>(a); break;
        case 2: k_rows<2, false><<<g, kBlock, 0, s>>>(a); break;
        case 3: k_rows<3, false><<<g, kBlock, 0, s>>>(a); break;
        case 4: k_rows<4, false><<<g, kBlock, 0, s>>>(a); break;
        case 5: k_rows<5, false><<<g, kBlock, 0, s>>>(a); break;
        case 6: k_rows<6, false><<<g, kBlock, 0, s>>>(a); break;
        case 7: k_rows<7, false><<<g, kBlock, 0, s>>>(a); break;
        case 8: k_rows<8, false><<<g, kBlock, 0, s>>>(a); break;
        default: return 2;
    }
    return check_launch();
}

int launch_rows_fill(ListArgs& a, hipStream_t s)
{
    const unsigned g = grid_for(a.g.N);
    switch (a.g.ndim) {
        case 1: k_rows<1, true><<<g, kBlock, 0, s>>>(a); break;
        case 2: k_rows<2, true><<<g, kBlock, 0, s>>>(a); break;
        case 3: k_rows<3, true><<<g, kBlock, 0, s>>>(a); break;
        case 4: k_rows<4, true><<<g, kBlock, 0, s>>>(a); break;
        case 5: k_rows<5, true><<<g, kBlock, 0, s>>>(a); break;
        case 6: k_rows<6, true><<<g, kBlock, 0, s>>>(a); break;
        case 7: k_rows<7, true><<<g, kBlock, 0, s>>>(a); break;
        case 8: k_rows<8, true><<<g, kBlock, 0, s>>>(a); break;
        default: return 2;
    }
    return check_launch();
}

size_t scan_tmp_bytes(size_t n)
{
    // per level: block sums (nb) + their scan (nb + 1), recursively
    size_t bytes = 0;
    size_t m = n;
    while (true) {
        const size_t nb = (m + kScanTile - 1) / kScanTile;
        bytes += (2 * nb + 2) * sizeof(uint32_t);
        if (nb <= 1) break;
        m = nb;
    }
    return bytes + 256;
}

int launch_scan_u32(const uint32_t* in, uint32_t* out, size_t n, void* tmp, hipStream_t s)
{
    if (n == 0) {
        k_fill_u32<<<1, 1, 0, s>>>(out, 0u, 1);
        return check_launch();
    }
    const size_t nb = (n + kScanTile - 1) / kScanTile;
    uint32_t* bsum = static_cast<uint32_t*>(tmp);
    uint32_t* bscan = bsum + nb + 1;
    k_scan_tile<<<(unsigned)nb, kBlock, 0, s>>>(in, out, n, bsum);
    if (nb == 1) {
        k_set_total<<<1, 1, 0, s>>>(out, n, bsum);
        return check_launch();
    }
    if (launch_scan_u32(bsum, bscan, nb, bscan + nb + 1, s)) return 1;
    k_scan_add<<<(unsigned)nb, kBlock, 0, s>>>(out, n, bscan);
    k_set_total<<<1, 1, 0, s>>>(out, n, bscan + nb);
    return check_launch();
}

size_t sort_hist_entries(size_t n)
{
    const size_t nb = (n + kSortTile - 1) / kSortTile;
    return 256 * nb + 1;
}

int launch_reorder_keys(const double* x, int N, int ndim, const double* box,
                        unsigned long long* keys, uint32_t* vals, hipStream_t s)
{
    if (N <= 0) return 0;
    const double b0 = box[0], b1 = ndim > 1 ? box[1] : 1.0, b2 = ndim > 2 ? box[2] : 1.0;
    k_reorder_keys<<<grid_for(N), kBlock, 0, s>>>(x, N, ndim, b0, b1, b2, keys, vals);
    return check_launch();
}

int launch_radix_sort(unsigned long long* keys, uint32_t* vals, size_t n, int key_bits,
                      SortTmp& t, hipStream_t s)
{
    if (n <= 1) return 0;
    const int nb = (int)((n + kSortTile - 1) / kSortTile);
    unsigned long long* kin = keys;
    unsigned long long* kout = t.keys_alt;
    uint32_t* vin = vals;
    uint32_t* vout = t.vals_alt;
    int passes = (key_bits + 7) / 8;
    if (passes % 2) ++passes;  // even pass count: result lands back in keys/vals
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        k_radix_hist<<<nb, kBlock, 0, s>>>(kin, n, shift, t.hist, nb);
        if (launch_scan_u32(t.hist, t.hist, (size_t)256 * nb, t.scan_tmp, s)) return 1;
        k_radix_scatter<<<nb, kBlock, 0, s>>>(kin, vin, kout, vout, n, shift, t.hist, nb);
        unsigned long long* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    return check_launch();
}

int launch_gather_f64(const double* src, double* dst, const uint32_t* perm, int N, int w,
                      hipStream_t s)
{
    if (N <= 0) return 0;
    k_gather_f64<<<grid_for((size_t)N * w), kBlock, 0, s>>>(src, dst, perm, N, w);
    return check_launch();
}
int launch_gather_u64(const unsigned long long* src, unsigned long long* dst,
                      const uint32_t* perm, int N, hipStream_t s)
{
    if (N <= 0) return 0;
    k_gather_u64<<<grid_for(N), kBlock, 0, s>>>(src, dst, perm, N);
    return check_launch();
}
int launch_gather_i8(const int8_t* src, int8_t* dst, const uint32_t* perm, int N,
                     hipStream_t s)
{
    if (N <= 0) return 0;
    k_gather_i8<<<grid_for(N), kBlock, 0, s>>>(src, dst, perm, N);
    return check_launch();
}
int launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s)
{
    if (n == 0) return 0;
    k_fill_u32<<<grid_for(n), kBlock, 0, s>>>(p, v, n);
    return check_launch();
}
int launch_iota_u64(unsigned long long* p, size_t n, hipStream_t s)
{
    if (n == 0) return 0;
    k_iota_u64<<<grid_for(n), kBlock, 0, s>>>(p, n);
    return check_launch();
}
int launch_scale_last_dim(double* x, int N, int ndim, double f, hipStream_t s)
{
    if (N <= 0) return 0;
    k_scale_last<<<grid_for(N), kBlock, 0, s>>>(x, N, ndim, f);
    return check_launch();
}
int launch_diameters(const double* D0, const int8_t* fixed, double kappa, double* D, int N,
                     hipStream_t s)
{
    if (N <= 0) return 0;
    k_diameters<<<grid_for(N), kBlock, 0, s>>>(D0, fixed, kappa, D, N);
    return check_launch();
}
int launch_adam_pass(const Geometry& g, const double* F, double* m, double* v, const double* x,
                     double* xo, double alpha, double inv_b1, double inv_b2, hipStream_t s)
{
    if (g.N <= 0) return 0;
    k_adam_pass<<<grid_for((size_t)g.N * g.ndim), kBlock, 0, s>>>(g, F, m, v, x, xo, alpha,
                                                                  inv_b1, inv_b2);
    return check_launch();
}
int launch_mean_force(const StepArgs& a, const double* F, int exact, hipStream_t s)
{
    if (a.g.N <= 0) return 0;
    k_mean_force<<<grid_for(a.g.N), kBlock, 0, s>>>(a, F);
    if (exact) k_mean_force_exact<<<1, kWave, 0, s>>>(a);
    return check_launch();
}
int launch_rk_active(const double* D0, const int8_t* fixed, double kappa, double* rk, int N,
                     hipStream_t s)
{
    if (N <= 0) return 0;
    k_rk_active<<<grid_for(N), kBlock, 0, s>>>(D0, fixed, kappa, rk, N);
    return check_launch();
}
int launch_copy_f64(const double* src, double* dst, size_t n, hipStream_t s)
{
    if (n == 0) return 0;
    k_copy_f64<<<grid_for(n), kBlock, 0, s>>>(src, dst, n);
    return check_launch();
}

}  // namespace rcp

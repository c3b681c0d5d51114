// C-ABI shim over the reference engine's kernel functions -- TEST
// INFRASTRUCTURE ONLY.  Compiled together with the reference sources (which
// stay under the read-only reference checkout; oracle/Makefile points at
// them) into oracle/_ref/libref_shim.so, so the oracle's restatement can be
// checked against the reference's own arithmetic.  Nothing in the product
// links this file.
//
// Reference functions wrapped (python_code/python/...):
//   get_forces_nd_3       src/core/rcp_generator.cpp:3597, include/.../rcp_generator.hpp:207
//   get_pairs_kdtree      src/core/rcp_generator.cpp:2758 (external linkage, not in the header)
//   get_pairs_nd_3        src/core/rcp_generator.cpp:1198, include/.../rcp_generator.hpp:185
//   compute_mean_force    src/core/rcp_generator.cpp:769
//   scale_diameters_nd    src/core/rcp_generator.cpp:1162
#include "rcpgenerator/rcp_generator.hpp"

#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

#include <omp.h>

namespace rcpgenerator {
// Defined in rcp_generator.cpp with external linkage; declared here because
// the reference header does not export it.
void get_pairs_kdtree(std::size_t N, std::size_t Ndim, const double* x,
                      const std::vector<double>& D, const std::vector<double>& box,
                      const std::vector<std::int8_t>& walls,
                      const std::vector<std::uint32_t>& refresh,
                      const std::uint32_t* max_neighbors_per_particle,
                      const std::uint64_t* pair_offsets, std::uint32_t* pairs_data,
                      std::uint32_t* out_max_observed, std::uint32_t* out_min_observed,
                      double* packed_d0, const double* D0_data);
}

namespace {
// Reference row layout: slot[0] = count, slot[1..count] = partners.
void csr_to_slots(std::size_t N, const std::uint64_t* off, const std::uint32_t* nbr,
                  std::vector<std::uint64_t>& pair_offsets,
                  std::vector<std::uint32_t>& pairs)
{
    pair_offsets.assign(N + 1, 0);
    std::uint64_t total = 0;
    for (std::size_t i = 0; i < N; ++i) {
        pair_offsets[i] = total;
        total += (off[i + 1] - off[i]) + 1;
    }
    pair_offsets[N] = total;
    pairs.assign(total, 0);
    for (std::size_t i = 0; i < N; ++i) {
        std::uint32_t* s = pairs.data() + pair_offsets[i];
        s[0] = static_cast<std::uint32_t>(off[i + 1] - off[i]);
        for (std::uint64_t e = off[i]; e < off[i + 1]; ++e) s[1 + e - off[i]] = nbr[e];
    }
}
}  // namespace

extern "C" {

void ref_set_threads(int n) { omp_set_num_threads(n); }

// out[0..6] = U, max_min_dist, Lc, Fmean, dkappa, (unused), (unused)
int ref_forces(std::size_t N, std::size_t ndim, const double* x, const double* D,
               const double* box, const std::int8_t* walls, double mu,
               const std::uint64_t* off, const std::uint32_t* nbr,
               const double* rk, double sumD_active, double* F, double* out)
{
    try {
        std::vector<std::uint64_t> po;
        std::vector<std::uint32_t> pd;
        csr_to_slots(N, off, nbr, po, pd);
        std::vector<double> Dv(D, D + N), bv(box, box + ndim);
        std::vector<std::int8_t> wv(walls, walls + ndim);
        std::vector<std::vector<double>> min_dist;
        std::vector<std::size_t> z(N, 0);
        double U = 0, mmd = 0, Lc = 0, Fm = 0, dk = 0;
        rcpgenerator::get_forces_nd_3(pd.data(), po.data(), x, N, ndim, Dv, bv, wv, F, U,
                                      min_dist, mmd, Lc, Fm, mu, dk, z, nullptr, 1.0,
                                      rk, sumD_active);
        out[0] = U; out[1] = mmd; out[2] = Lc; out[3] = Fm; out[4] = dk;
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// Full-refresh pair build through the reference backends.  backend 0 =
// kd-tree, 1 = sort+walk.  caps[i] = per-row capacity; rows come back as
// CSR (off[N+1], nbr[cap_total]).  Returns -1 on overflow/error, else the
// number of pairs.
long long ref_pairs(int backend, std::size_t N, std::size_t ndim, const double* x,
                    const double* D, const double* box, std::uint32_t cap_per_row,
                    std::uint64_t* off, std::uint32_t* nbr, std::size_t nbr_cap)
{
    try {
        std::vector<double> Dv(D, D + N), bv(box, box + ndim);
        std::vector<std::int8_t> wv(ndim, 0);
        std::vector<std::uint32_t> refresh(N, 1u), caps(N, cap_per_row);
        std::vector<std::uint64_t> po(N + 1);
        for (std::size_t i = 0; i <= N; ++i) po[i] = static_cast<std::uint64_t>(i) * (cap_per_row + 1);
        std::vector<std::uint32_t> pd(po[N], 0);
        for (std::size_t i = 0; i < N; ++i) pd[po[i]] = 0;
        std::uint32_t mx = 0, mn = 0;
        if (backend == 0)
            rcpgenerator::get_pairs_kdtree(N, ndim, x, Dv, bv, wv, refresh, caps.data(),
                                           po.data(), pd.data(), &mx, &mn, nullptr, nullptr);
        else
            rcpgenerator::get_pairs_nd_3(N, ndim, x, Dv, bv, wv, refresh, caps.data(),
                                         po.data(), pd.data(), &mx, &mn, nullptr, nullptr);
        std::size_t used = 0;
        for (std::size_t i = 0; i < N; ++i) {
            off[i] = used;
            const std::uint32_t* s = pd.data() + po[i];
            for (std::uint32_t k = 1; k <= s[0]; ++k) {
                if (used < nbr_cap) nbr[used] = s[k];
                ++used;
            }
        }
        off[N] = used;
        return static_cast<long long>(used);
    } catch (const std::exception&) {
        return -1;
    }
}

double ref_mean_force(const double* F, std::size_t N, std::size_t ndim)
{
    std::vector<std::size_t> z(N, 0);
    return rcpgenerator::compute_mean_force(F, N, ndim, 0.0, z);
}

double ref_scale_diameters(const double* D, std::size_t N, double phi_target,
                           const double* box, std::size_t ndim, int fix_height,
                           double* out)
{
    std::vector<double> Dv(D, D + N), bv(box, box + ndim);
    auto r = rcpgenerator::scale_diameters_nd(Dv, phi_target, bv, ndim, fix_height != 0);
    std::memcpy(out, r.first.data(), N * sizeof(double));
    return r.second;
}

}  // extern "C"

// Host initializer: diameter distributions, phi scaling and random
// sequential addition, the behaviour of rcpgenerator::initialize_particles
// (KD-physics/RCPGenerator python_code/python/src/core/initialize_particles.cpp,
// IP:<line> below).  It runs once per packing on the host CPU, before any
// device work, and is not part of the per-step hot path.
//
// Reproducing the reference's outputs requires the same random streams in
// the same order: std::mt19937 for diameters (IP:435), std::mt19937_64 for
// placement (IP:238), and the libstdc++ distributions the reference uses.
// The placement test here is a plain cell grid sized to the largest
// diameter (every already-placed particle that could overlap a trial lies
// in the trial's 3^d stencil), which accepts or rejects exactly the trials
// the reference's halo-registered fine grid does, because both test every
// possible contact with the same inequality.
#include "rcp_amd.h"
#include "rcp_engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

namespace rcp {
namespace {

struct Dist {
    std::string type = "mono";
    double d = 1.0, mu = 1.0, sigma = 0.2, mu1 = 1.0, sigma1 = 0.2, mu2 = 2.0, sigma2 = 0.3,
           p = 0.5, d1 = 1.0, d2 = 2.0, d_min = 0.5, d_max = 1.5, exponent = -2.5, scale = 1.0,
           shape = 2.0;
    std::vector<double> custom;
};

// Stretch the realised extremes onto [lo, hi] (IP:56-66, 82-92, 103-113).
void stretch(std::vector<double>& D, double lo, double hi)
{
    const double mn = *std::min_element(D.begin(), D.end());
    const double mx = *std::max_element(D.begin(), D.end());
    if (mx > mn + 1e-12) {
        for (double& v : D) v = lo + (v - mn) * (hi - lo) / (mx - mn);
    } else {
        std::fill(D.begin(), D.end(), lo);
    }
}

std::vector<double> draw_diameters(std::size_t N, const Dist& ds, std::mt19937& rng)
{
    // IP:17-130, one branch per distribution, same draw order
    std::vector<double> D(N);
    const std::string& t = ds.type;
    if (t == "mono") {
        std::fill(D.begin(), D.end(), ds.d);
    } else if (t == "gaussian") {
        std::normal_distribution<> g(ds.mu, ds.sigma);
        for (double& v : D) v = std::abs(g(rng));
    } else if (t == "bigaussian" || t == "bidisperse") {
        const std::size_t n1 = static_cast<std::size_t>(std::round(ds.p * N));
        if (t == "bigaussian") {
            std::normal_distribution<> g1(ds.mu1, ds.sigma1), g2(ds.mu2, ds.sigma2);
            for (std::size_t i = 0; i < n1; ++i) D[i] = std::abs(g1(rng));
            for (std::size_t i = n1; i < N; ++i) D[i] = std::abs(g2(rng));
        } else {
            for (std::size_t i = 0; i < N; ++i) D[i] = i < n1 ? ds.d1 : ds.d2;
        }
        std::shuffle(D.begin(), D.end(), rng);
    } else if (t == "lognormal") {
        std::lognormal_distribution<> g(ds.mu, ds.sigma);
        for (double& v : D) v = std::clamp(g(rng), 0.0025, 30.0);
    } else if (t == "flat") {
        std::uniform_real_distribution<> u(ds.d_min, ds.d_max);
        for (double& v : D) v = u(rng);
        stretch(D, ds.d_min, ds.d_max);
    } else if (t == "powerlaw") {
        std::uniform_real_distribution<> u(0.0, 1.0);
        const double a = ds.exponent + 1.0;
        for (double& v : D) {
            const double q = u(rng);
            if (std::abs(a) < 1e-12) {
                v = ds.d_min * std::exp(q * std::log(ds.d_max / ds.d_min));
            } else {
                const double top = std::pow(ds.d_max, a) - std::pow(ds.d_min, a);
                v = std::pow(top * q + std::pow(ds.d_min, a), 1.0 / a);
            }
        }
        stretch(D, ds.d_min, ds.d_max);
    } else if (t == "exponential") {
        const double lambda = -std::log(0.5) / (ds.d_max - ds.d_min);
        std::uniform_real_distribution<> u(0.0, 1.0);
        for (double& v : D) {
            const double q = u(rng);
            v = ds.d_min - (1.0 / lambda) * std::log(1.0 - q);
            if (v > ds.d_max) v = ds.d_max;
        }
        stretch(D, ds.d_min, ds.d_max);
    } else if (t == "weibull") {
        std::weibull_distribution<> w(ds.shape, ds.scale);
        for (double& v : D) v = w(rng);
    } else if (t == "custom") {
        if (ds.custom.size() != N) throw RtError("custom distribution needs N entries");
        D = ds.custom;
    } else {
        // The reference terminates the process here (IP:124-127); raising
        // keeps the caller alive.
        throw RtError("Unsupported distribution: " + t);
    }
    return D;
}

// IP:132-150 (the initializer's own form: unit volume times r^d per term)
double init_scale(std::vector<double>& D, double phi_target, const std::vector<double>& box,
                  std::size_t nd)
{
    double V_box = 1.0;
    for (double L : box) V_box *= L;
    double sumVol = 0.0;
    for (double Di : D) {
        const double r = Di * 0.5;
        const double vol_unit = std::pow(M_PI, nd / 2.0) / std::tgamma(nd / 2.0 + 1.0);
        sumVol += vol_unit * std::pow(r, double(nd));
    }
    const double factor = std::pow(phi_target * V_box / sumVol, 1.0 / double(nd));
    for (double& v : D) v *= factor;
    return factor;
}

// Random sequential addition, largest first (IP:184-343 behaviour).
void place(std::vector<double>& x, const std::vector<double>& D, const std::vector<double>& box,
           const std::vector<int8_t>& walls, std::uint64_t seed)
{
    const std::size_t N = D.size(), nd = box.size();
    if (N == 0) return;
    double Dmax = 0.0;
    for (double v : D) Dmax = std::max(Dmax, v);
    if (Dmax <= 0.0) Dmax = 1e-12;
    // grid with cells no smaller than Dmax (contact distance <= Dmax)
    std::vector<long> nc(nd);
    std::vector<double> h(nd);
    std::size_t total = 1;
    for (std::size_t d = 0; d < nd; ++d) {
        long n = (long)std::floor(box[d] / (Dmax * (1.0 + 1e-9)));
        if (n < 3) n = 1;
        nc[d] = n;
        h[d] = box[d] / double(n);
        total *= (std::size_t)n;
    }
    while (total > 4 * N + 64) {
        std::size_t dm = 0;
        for (std::size_t d = 1; d < nd; ++d)
            if (nc[d] > nc[dm]) dm = d;
        long n = nc[dm] / 2;
        if (n < 3) n = 1;
        total = total / (std::size_t)nc[dm] * (std::size_t)n;
        nc[dm] = n;
        h[dm] = box[dm] / double(n);
    }
    std::vector<std::vector<std::uint32_t>> cells(total);

    const bool spherical = walls[0] < 0;
    const std::size_t Ksph = spherical ? static_cast<std::size_t>(-walls[0]) : 0;
    const double Rsph = box[0] * 0.5;
    std::mt19937_64 rng{seed ? seed : 0x9E3779B97F4A7C15ULL};
    std::uniform_real_distribution<double> ud(0.0, 1.0);
    const std::size_t kMaxAttempts = std::size_t{1} << 22;

    std::vector<double> xt(nd);
    std::vector<long> ci(nd);
    for (std::size_t c = 0; c < N; ++c) {
        const double Dc = D[c];
        std::size_t attempts = 0;
        while (true) {
            if (++attempts > kMaxAttempts)
                throw RtError(
                    "initializer: could not place a non-overlapping particle "
                    "(phi_init too high for random sequential addition; "
                    "lower the initial packing fraction)");
            for (std::size_t d = 0; d < nd; ++d)
                xt[d] = walls[d] == 0 ? ud(rng) * box[d] : Dc * 0.5 + ud(rng) * (box[d] - Dc);
            if (spherical) {
                double s = 0.0;
                for (std::size_t d = 0; d < Ksph; ++d) {
                    const double q = xt[d] - Rsph;
                    s += q * q;
                }
                if (s > (Rsph - Dc * 0.5) * (Rsph - Dc * 0.5)) continue;
            }
            for (std::size_t d = 0; d < nd; ++d) {
                long k = (long)std::floor(xt[d] / h[d]);
                ci[d] = std::min(std::max(k, 0L), nc[d] - 1);
            }
            // every placed particle within Dmax of the trial is in the stencil
            bool ok = true;
            std::size_t nst = 1;
            for (std::size_t d = 0; d < nd; ++d) nst *= (nc[d] >= 3 ? 3 : 1);
            for (std::size_t s = 0; s < nst && ok; ++s) {
                std::size_t q = s, cell = 0, stride = 1;
                for (std::size_t d = 0; d < nd; ++d) {
                    long o = 0;
                    if (nc[d] >= 3) { o = (long)(q % 3) - 1; q /= 3; }
                    long k = ci[d] + o;
                    if (k < 0) k += nc[d];
                    if (k >= nc[d]) k -= nc[d];
                    cell += (std::size_t)k * stride;
                    stride *= (std::size_t)nc[d];
                }
                for (std::uint32_t j : cells[cell]) {
                    double s2 = 0.0;
                    for (std::size_t d = 0; d < nd; ++d) {
                        double dx = xt[d] - x[(std::size_t)j * nd + d];
                        if (walls[d] == 0) dx -= box[d] * std::nearbyint(dx / box[d]);
                        s2 += dx * dx;
                    }
                    const double r = 0.5 * (D[j] + Dc);
                    if (s2 <= r * r) { ok = false; break; }
                }
            }
            if (!ok) continue;
            std::size_t cell = 0, stride = 1;
            for (std::size_t d = 0; d < nd; ++d) {
                cell += (std::size_t)ci[d] * stride;
                stride *= (std::size_t)nc[d];
            }
            cells[cell].push_back((std::uint32_t)c);
            for (std::size_t d = 0; d < nd; ++d) x[c * nd + d] = xt[d];
            break;
        }
    }
}

}  // namespace
}  // namespace rcp

extern "C" {

void rcp_init_config_init(rcp_init_config* c)
{
    std::memset(c, 0, sizeof(*c));
    rcp::Dist d;
    c->phi = 0.05;
    c->dist_type = "mono";
    c->d = d.d; c->mu = d.mu; c->sigma = d.sigma; c->mu1 = d.mu1; c->sigma1 = d.sigma1;
    c->mu2 = d.mu2; c->sigma2 = d.sigma2; c->p = d.p; c->d1 = d.d1; c->d2 = d.d2;
    c->d_min = d.d_min; c->d_max = d.d_max; c->exponent = d.exponent; c->scale = d.scale;
    c->shape = d.shape;
}

int rcp_initialize_particles(const rcp_init_config* c, double* positions, double* diameters,
                             double* box_out, int8_t* walls_out, double* diameter_scale_factor,
                             double* phi_modifier_out)
{
    try {
        rcp::g_last_error.clear();
        const std::size_t N = c->N, nd = c->ndim;
        // IP:399-474
        if (nd < 2) throw rcp::RtError("--Ndim must be >=2");
        std::vector<int8_t> walls(c->walls, c->walls + c->nwalls);
        if (walls.size() != nd) walls.resize(nd, 0);
        std::vector<double> box(c->box, c->box + c->nbox);
        if (box.empty()) box.assign(nd, 1.0);
        if (c->phi <= 0 || c->phi >= 1 || N == 0) throw rcp::RtError("Missing or invalid phi/N");
        if (box.size() != nd) throw rcp::ArgError("box length must equal Ndim");
        const double Llast = box[0];
        double h = 1;
        if (c->fix_height) {
            h = box[nd - 1];
            box[nd - 1] = box[0];
        }
        double phi_modifier = 1.0;
        if (walls[0] < 0) phi_modifier = rcp::sphere_volume(1 / 2., static_cast<std::size_t>(-walls[0]));

        rcp::Dist ds;
        ds.type = c->dist_type ? c->dist_type : "mono";
        ds.d = c->d; ds.mu = c->mu; ds.sigma = c->sigma; ds.mu1 = c->mu1; ds.sigma1 = c->sigma1;
        ds.mu2 = c->mu2; ds.sigma2 = c->sigma2; ds.p = c->p; ds.d1 = c->d1; ds.d2 = c->d2;
        ds.d_min = c->d_min; ds.d_max = c->d_max; ds.exponent = c->exponent; ds.scale = c->scale;
        ds.shape = c->shape;
        if (c->custom) ds.custom.assign(c->custom, c->custom + c->ncustom);

        std::mt19937 rng{static_cast<std::mt19937::result_type>(c->seed)};
        std::vector<double> D = rcp::draw_diameters(N, ds, rng);
        const double factor = rcp::init_scale(D, c->phi * phi_modifier, box, nd);

        // descending order with the same comparator as IP:441-449 (ties are
        // ordered by std::sort itself, so the same library call is used)
        std::vector<std::size_t> idx(N);
        std::iota(idx.begin(), idx.end(), 0);
        std::sort(idx.begin(), idx.end(), [&](std::size_t a, std::size_t b) { return D[a] > D[b]; });
        std::vector<double> Ds(N);
        for (std::size_t i = 0; i < N; ++i) Ds[i] = D[idx[i]];

        if (c->fix_height) {
            const double s = std::pow(h * Ds[0] / Llast, 1.0 / (nd - 1));
            const double eps = std::pow(s, nd);
            for (double& v : Ds) v *= s;
            box[nd - 1] *= eps;
        }
        std::vector<double> x(N * nd, 0.0);
        rcp::place(x, Ds, box, walls, c->seed);

        std::memcpy(positions, x.data(), x.size() * sizeof(double));
        std::memcpy(diameters, Ds.data(), N * sizeof(double));
        std::memcpy(box_out, box.data(), nd * sizeof(double));
        std::memcpy(walls_out, walls.data(), nd);
        if (diameter_scale_factor) *diameter_scale_factor = factor;
        if (phi_modifier_out) *phi_modifier_out = phi_modifier;
        return RCP_OK;
    } catch (const rcp::ArgError& e) {
        rcp::g_last_error = e.what();
        return RCP_E_ARG;
    } catch (const std::bad_alloc&) {
        rcp::g_last_error = "out of memory";
        return RCP_E_NOMEM;
    } catch (const std::exception& e) {
        rcp::g_last_error = e.what();
        return RCP_E_RUNTIME;
    }
}

}  // extern "C"

// Host side of the MI355X packing engine: the run driver and schedule
// controller of rcpgenerator::run_packing_observed, restated for a device-
// resident state, plus the C ABI declared in include/rcp_amd.h.
//
// Compiled with g++ (the compiler the reference is built with) so the
// scalar controller arithmetic and libm calls round exactly as the
// reference's do.  Reference citations: KD-physics/RCPGenerator,
// python_code/python/src/core/rcp_generator.cpp (RG:<line>).
#include "rcp_amd.h"
#include "rcp_engine.h"
#include "rcp_kernels.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace rcp {

thread_local std::string g_last_error;

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw HipFailure(std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------ DevArena

void* dev_alloc(size_t bytes)
{
    void* p = nullptr;
    if (bytes == 0) bytes = 8;
    hip_check(hipMalloc(&p, bytes), "hipMalloc");
    return p;
}

void dev_free(void* p)
{
    if (p) (void)hipFree(p);
}

double sphere_volume(double r, std::size_t ndim)
{
    // rcp_generator.hpp:165-168
    return std::pow(M_PI, ndim / 2.0) * std::pow(r, ndim) / std::tgamma(ndim / 2.0 + 1.0);
}

double scale_diameters(const std::vector<double>& D, double phi_target,
                       const std::vector<double>& box, std::size_t ndim, bool fix_height,
                       std::vector<double>& out)
{
    // RG:1162-1196, same summation order
    double V_box = 1.0;
    for (double L : box) V_box *= L;
    double sumVol = 0.0;
    for (double Di : D) sumVol += sphere_volume(Di * 0.5, ndim);
    const double factor = fix_height
        ? std::pow(phi_target * V_box / sumVol, 1.0 / double(ndim - 1))
        : std::pow(phi_target * V_box / sumVol, 1.0 / double(ndim));
    out.resize(D.size());
    for (std::size_t i = 0; i < D.size(); ++i) out[i] = D[i] * factor;
    return factor;
}

double reference_shell(std::vector<double> D)
{
    // t_shell = min(max(median*1.02, Dmin*2.25), Dmax) * 0.5 (RG:2846),
    // median = element N/2 of the ascending order
    const std::size_t n = D.size();
    std::nth_element(D.begin(), D.begin() + n / 2, D.end());
    const double median = D[n / 2];
    const double Dmin = *std::min_element(D.begin(), D.end());
    const double Dmax = *std::max_element(D.begin(), D.end());
    return std::min(std::max(median * 1.02, Dmin * 2.25), Dmax) * 0.5;
}

// ------------------------------------------------------------- Engine

Engine::Engine(const rcp_run_config& c)
{
    if (c.N == 0) throw ArgError("rcp: N must be >= 1");
    if (c.ndim < 1 || c.ndim > (size_t)kMaxDim)
        throw ArgError("rcp: Ndim must be in 1..8");
    if (!c.positions || !c.diameters) throw ArgError("rcp: positions/diameters missing");
    if (c.N > 0x7fffffffull) throw ArgError("rcp: N too large");
    N_ = c.N;
    nd_ = c.ndim;
    opt_ = c;
    opt_.positions = nullptr;
    opt_.diameters = nullptr;
    opt_.fixed = nullptr;
    opt_.box = nullptr;
    opt_.walls = nullptr;

    std::vector<double> x(c.positions, c.positions + N_ * nd_);
    std::vector<double> D(c.diameters, c.diameters + N_);
    std::vector<int8_t> fixed;
    if (c.fixed) {
        fixed.assign(c.fixed, c.fixed + N_);
        has_fixed_ = std::any_of(fixed.begin(), fixed.end(), [](int8_t f) { return f != 0; });
    }
    fixed_nonempty_ = !fixed.empty();

    // box / walls (RG:4437-4447)
    if (c.nbox == 0) box_.assign(nd_, 1.0);
    else if (c.nbox != nd_) throw RtError("Error: --box length mismatch");
    else box_.assign(c.box, c.box + nd_);
    if (c.nwalls == 0) walls_.assign(nd_, 0);
    else if (c.nwalls != nd_) throw RtError("Error: --walls expects Ndim entries");
    else walls_.assign(c.walls, c.walls + nd_);

    // curved container normalisation (RG:4449-4458)
    phi_ = 0.025;
    phi_modifier_ = 1;
    if (walls_[0] < 0) {
        if (walls_[0] == -1) walls_[0] = -2;
        for (std::size_t k = 1; k < static_cast<std::size_t>(-walls_[0]) && k < nd_; ++k) {
            walls_[k] = -1;
            box_[k] = box_[0];
        }
        phi_modifier_ = sphere_volume(1 / 2., static_cast<std::size_t>(-walls_[0]));
        curved_T_ = -walls_[0];
        if ((std::size_t)curved_T_ > nd_)
            throw ArgError("rcp: curved container rank -walls[0] exceeds Ndim");
    }
    if (c.fix_height) box_[nd_ - 1] = box_[nd_ - 1] * D[0];

    // start-phi rescale (RG:4460-4493)
    if (c.has_initial_phi_target) phi_ = c.initial_phi_target;
    else if (nd_ == 2) phi_ = 0.575;
    else if (nd_ == 3) phi_ = 0.33;
    else if (nd_ == 4) phi_ = 0.15;
    else if (nd_ == 5) phi_ = 0.10;
    else if (nd_ > 5) phi_ = 0.10 / std::pow(2, (nd_ - 4));
    if (!c.fix_diameter) {
        std::vector<double> Ds;
        const double factor = scale_diameters(D, phi_ * phi_modifier_, box_, nd_,
                                              c.fix_height != 0, Ds);
        if (c.fix_height) {
            box_[nd_ - 1] = box_[nd_ - 1] * factor;
            for (std::size_t i = 0; i < N_; ++i)
                x[i * nd_ + nd_ - 1] = x[i * nd_ + nd_ - 1] * factor;
        }
        D = std::move(Ds);
    }
    double coeff = std::pow(M_PI, nd_ / 2.0) / std::tgamma(nd_ / 2.0 + 1.0);
    double current_volume = 0.0;
    for (std::size_t i = 0; i < N_; ++i)
        current_volume += coeff * std::pow(D[i] / 2.0, nd_);
    double box_volume = std::accumulate(box_.begin(), box_.end(), 1.0, std::multiplies<double>());
    phi_ = current_volume / box_volume / phi_modifier_;

    // step cap (RG:4713-4737), speed preset "quick"
    step_cap_ = c.max_steps > 0 ? c.max_steps : [&]() {
        double rung1 = 2.0 * 4000.0 + 9.0 * std::sqrt(double(N_) * double(nd_)) - 2000.0;
        if (rung1 < 1500.0) rung1 = 1500.0;
        const std::size_t formula = static_cast<std::size_t>(rung1 * double(2 + 3));
        return std::max<std::size_t>(60000, formula);
    }();
    const std::size_t hist = step_cap_ + 1;
    phi_h_.assign(hist, 0.0);
    F_h_.assign(hist, 0.0);
    U_h_.assign(hist, 0.0);
    ov_h_.assign(hist, 0.0);
    mu_h_.assign(hist, 0.0);
    al_h_.assign(hist, 0.0);
    fl_h_.assign(hist, 0);

    // schedule constants (RG:4950-4965, 4781-4787)
    alpha_ = alpha_max_ = 0.005;
    if (nd_ == 3) alpha_ = alpha_max_ = 0.0045;
    if (nd_ == 4) alpha_ = alpha_max_ = 0.0035;
    if (nd_ >= 5) alpha_ = alpha_max_ = 0.0025;
    mu_ = c.has_mu ? c.mu : 2.5e-4;
    runtime_fix_ = c.fix_diameter != 0;
    rung1_deadline_ = [&]() {
        double dl = 2.0 * 4000.0 + 9.0 * std::sqrt(double(N_) * double(nd_)) - 2000.0;
        if (dl < 1500.0) dl = 1500.0;
        return static_cast<std::size_t>(dl);
    }();

    // D0 caches before the initial reorder (RG:4808-4826)
    recompute_d0_cache(D, fixed);

    // ---- device state
    geo_.N = (int)N_;
    geo_.ndim = (int)nd_;
    for (std::size_t d = 0; d < nd_; ++d) {
        geo_.box[d] = box_[d];
        geo_.inv_box[d] = 1.0 / box_[d];
        geo_.walls[d] = walls_[d];
    }
    geo_.curved_T = curved_T_;
    hip_check(hipStreamCreate(&stream_), "hipStreamCreate");
    if (opt_.time_kernels)
        for (auto& ev : ev_) hip_check(hipEventCreate(&ev), "hipEventCreate");
    alloc_device();
    hip_check(hipMemcpyAsync(x_[cur_], x.data(), N_ * nd_ * sizeof(double),
                             hipMemcpyHostToDevice, stream_), "upload x");
    hip_check(hipMemcpyAsync(D0_, D.data(), N_ * sizeof(double), hipMemcpyHostToDevice,
                             stream_), "upload D0");
    if (fixed_nonempty_)
        hip_check(hipMemcpyAsync(fixed_, fixed.data(), N_, hipMemcpyHostToDevice, stream_),
                  "upload fixed");
    hip_check(hipMemsetAsync(m_, 0, N_ * nd_ * sizeof(double), stream_), "zero m");
    hip_check(hipMemsetAsync(v_, 0, N_ * nd_ * sizeof(double), stream_), "zero v");
    if (launch_iota_u64(origin_, N_, stream_)) throw HipFailure("iota launch");

    // initial reorder (RG:4990-4993) and first build + forces (RG:5011-5013)
    if (opt_.reorder_interval > 0) reorder();
    kappa_D_ = 1.0;
    build_lists();
    const Scalars s0 = step_device(/*do_adam=*/false, /*alpha*/ 0.0, 1.0, 1.0, /*use_rk*/ 0);
    finalize_scalars(s0, mu_, /*use_active*/ false);
}

Engine::~Engine()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : allocs_) dev_free(p);
    if (h_result_) (void)hipHostFree(h_result_);
    for (auto& ev : ev_)
        if (ev) (void)hipEventDestroy(ev);
    if (stream_) (void)hipStreamDestroy(stream_);
}

template <class T>
T* Engine::dalloc(size_t n)
{
    T* p = static_cast<T*>(dev_alloc(n * sizeof(T)));
    allocs_.push_back(p);
    return p;
}

void Engine::dfree(void* p)
{
    auto it = std::find(allocs_.begin(), allocs_.end(), p);
    if (it != allocs_.end()) {
        dev_free(p);
        allocs_.erase(it);
    }
}

void Engine::alloc_device()
{
    const size_t Nd = N_ * nd_;
    x_[0] = dalloc<double>(Nd);
    x_[1] = dalloc<double>(Nd);
    m_ = dalloc<double>(Nd);
    v_ = dalloc<double>(Nd);
    tmp_f64_ = dalloc<double>(Nd);
    x_anchor_ = dalloc<double>(Nd);
    D0_ = dalloc<double>(N_);
    origin_ = dalloc<unsigned long long>(N_);
    origin_tmp_ = dalloc<unsigned long long>(N_);
    if (fixed_nonempty_) {
        fixed_ = dalloc<int8_t>(N_);
        fixed_tmp_ = dalloc<int8_t>(N_);
    }
    cell_of_ = dalloc<uint32_t>(N_);
    slot_ = dalloc<uint32_t>(N_);
    items_ = dalloc<uint32_t>(N_);
    row_count_ = dalloc<uint32_t>(N_ + 1);
    row_start_ = dalloc<uint32_t>(N_ + 1);
    keys_ = dalloc<unsigned long long>(N_);
    keys_alt_ = dalloc<unsigned long long>(N_);
    vals_ = dalloc<uint32_t>(N_);
    vals_alt_ = dalloc<uint32_t>(N_);
    hist_ = dalloc<uint32_t>(sort_hist_entries(N_));
    const size_t nblocks = (N_ + 255) / 256;
    partials_ = dalloc<double>(nblocks * S_NUM + S_NUM);
    result_ = dalloc<double>(S_NUM);
    counter_ = dalloc<unsigned int>(1);
    hip_check(hipMemsetAsync(counter_, 0, sizeof(unsigned int), stream_), "zero counter");
    if (opt_.exact_reductions) {
        F_ = dalloc<double>(Nd);
        mag_ = dalloc<double>(N_);
    }
    hip_check(hipHostMalloc((void**)&h_result_, S_NUM * sizeof(double), hipHostMallocDefault),
              "hipHostMalloc");
    ensure_scan_tmp(sort_hist_entries(N_) + 8);
}

void Engine::ensure_scan_tmp(size_t n)
{
    const size_t bytes = scan_tmp_bytes(n);
    if (bytes <= scan_tmp_bytes_) return;
    if (scan_tmp_) dfree(scan_tmp_);
    scan_tmp_ = dalloc<char>(bytes);
    scan_tmp_bytes_ = bytes;
}

void Engine::recompute_d0_cache(const std::vector<double>& D0, const std::vector<int8_t>& fixed)
{
    // Dmin / S_D0 caches (RG:4808-4826); max per class for list sizing
    D0_min_ = D0[0];
    S_D0_ = 0.0;
    D0max_free_ = 0.0;
    D0max_fixed_ = 0.0;
    for (std::size_t i = 0; i < N_; ++i) {
        if (D0[i] < D0_min_) D0_min_ = D0[i];
        const double r = D0[i] * 0.5;
        double rN = 1.0;
        for (std::size_t k = 0; k < nd_; ++k) rN *= r;
        S_D0_ += rN;
        const bool f = !fixed.empty() && fixed[i];
        if (f) D0max_fixed_ = std::max(D0max_fixed_, D0[i]);
        else D0max_free_ = std::max(D0max_free_, D0[i]);
    }
    shell0_ = reference_shell(D0);
}

void Engine::download_d0(std::vector<double>& D0, std::vector<int8_t>& fixed)
{
    D0.resize(N_);
    hip_check(hipMemcpyAsync(D0.data(), D0_, N_ * sizeof(double), hipMemcpyDeviceToHost,
                             stream_), "download D0");
    fixed.clear();
    if (fixed_nonempty_) {
        fixed.resize(N_);
        hip_check(hipMemcpyAsync(fixed.data(), fixed_, N_, hipMemcpyDeviceToHost, stream_),
                  "download fixed");
    }
    hip_check(hipStreamSynchronize(stream_), "sync");
}

// --------------------------------------------------------------- reorder

void Engine::reorder()
{
    // RG:4850-4948: Morton-21 (Ndim 3) or x-column order, stable by index;
    // every per-particle array is permuted, then the lists are rebuilt.
    double* x = x_[cur_];
    if (launch_reorder_keys(x, (int)N_, (int)nd_, geo_.box, keys_, vals_, stream_))
        throw HipFailure("reorder keys launch");
    SortTmp t;
    t.keys_alt = keys_alt_;
    t.vals_alt = vals_alt_;
    t.hist = hist_;
    t.scan_tmp = scan_tmp_;
    if (launch_radix_sort(keys_, vals_, N_, 64, t, stream_)) throw HipFailure("radix sort");
    const uint32_t* perm = vals_;
    auto permute = [&](double*& a, int w) {
        if (launch_gather_f64(a, tmp_f64_, perm, (int)N_, w, stream_)) throw HipFailure("gather");
        std::swap(a, tmp_f64_);
    };
    permute(x_[cur_], (int)nd_);
    permute(m_, (int)nd_);
    permute(v_, (int)nd_);
    // D0 is N wide; reuse the N*ndim scratch
    if (launch_gather_f64(D0_, tmp_f64_, perm, (int)N_, 1, stream_)) throw HipFailure("gather D0");
    hip_check(hipMemcpyAsync(D0_, tmp_f64_, N_ * sizeof(double), hipMemcpyDeviceToDevice, stream_),
              "copy D0");
    if (launch_gather_u64(origin_, origin_tmp_, perm, (int)N_, stream_)) throw HipFailure("gather origin");
    std::swap(origin_, origin_tmp_);
    if (fixed_nonempty_) {
        if (launch_gather_i8(fixed_, fixed_tmp_, perm, (int)N_, stream_)) throw HipFailure("gather fixed");
        std::swap(fixed_, fixed_tmp_);
    }
    ++reorders_;
    need_build_ = true;
}

// ------------------------------------------------------------ neighbour lists

void Engine::build_lists()
{
    const double kD = kappa_D_;
    const double Dmax = std::max(D0max_free_ * kD, D0max_fixed_);
    t_list_ = shell0_ * kD;
    if (!(t_list_ > 0.0)) t_list_ = 1e-3 * std::max(Dmax, 1e-300);
    ListArgs la;
    la.g = geo_;
    la.x = x_[cur_];
    la.D0 = D0_;
    la.fixed = fixed_;
    la.kappa = kD;
    la.t_shell = t_list_;
    la.half = 0;
    configure_grid(Dmax + t_list_, la.grid);
    if (opt_.time_kernels) hip_check(hipEventRecord(ev_[2], stream_), "event");
    build_rows(la);
    if (opt_.time_kernels) {
        hip_check(hipEventRecord(ev_[3], stream_), "event");
        hip_check(hipEventSynchronize(ev_[3]), "event sync");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, ev_[2], ev_[3]), "event time");
        t_build_ms_ += ms;
        ++t_build_n_;
    }
    hip_check(hipMemcpyAsync(x_anchor_, x_[cur_], N_ * nd_ * sizeof(double),
                             hipMemcpyDeviceToDevice, stream_), "anchor copy");
    kappa_b_ = kD;
    fh_drift_ = 0.0;
    need_build_ = false;
    ++list_builds_;
}

void Engine::configure_grid(double reach, CellGrid& grid)
{
    const double edge = reach * (1.0 + 1e-6);
    long long total = 1;
    for (std::size_t d = 0; d < nd_; ++d) {
        long long n = (edge > 0.0) ? (long long)std::floor(box_[d] / edge) : 1;
        if (n < 3) n = 1;
        if (n > (1 << 20)) n = 1 << 20;
        grid.nc[d] = (int)n;
        total *= n;
    }
    // keep the table near the particle count (bigger cells stay correct)
    const long long cap = std::max<long long>(64, 2 * (long long)N_);
    while (total > cap) {
        std::size_t dmax = 0;
        for (std::size_t d = 1; d < nd_; ++d)
            if (grid.nc[d] > grid.nc[dmax]) dmax = d;
        int n = grid.nc[dmax] / 2;
        if (n < 3) n = 1;
        total = total / grid.nc[dmax] * n;
        grid.nc[dmax] = n;
    }
    for (std::size_t d = 0; d < nd_; ++d) grid.cell_scale[d] = grid.nc[d] / box_[d];
    grid.total = total;
}

void Engine::build_rows(ListArgs& la)
{
    const size_t total = (size_t)la.grid.total;
    if (total + 1 > cell_cap_) {
        if (cell_count_) { dfree(cell_count_); dfree(cell_start_); }
        cell_cap_ = total + 1;
        cell_count_ = dalloc<uint32_t>(cell_cap_);
        cell_start_ = dalloc<uint32_t>(cell_cap_);
    }
    ensure_scan_tmp(std::max(total, N_) + 8);
    la.cell_of = cell_of_;
    la.slot = slot_;
    la.cell_count = cell_count_;
    la.cell_start = cell_start_;
    la.items = items_;
    la.row_count = row_count_;
    la.row_start = row_start_;
    if (launch_cells(la, scan_tmp_, scan_tmp_bytes_, stream_)) throw HipFailure("cells launch");
    if (launch_rows_count(la, stream_)) throw HipFailure("rows count launch");
    if (launch_scan_u32(row_count_, row_start_, N_, scan_tmp_, stream_))
        throw HipFailure("rows scan launch");
    uint32_t nnz = 0;
    hip_check(hipMemcpyAsync(&nnz, row_start_ + N_, sizeof(uint32_t), hipMemcpyDeviceToHost,
                             stream_), "nnz download");
    hip_check(hipStreamSynchronize(stream_), "sync nnz");
    if (nnz + 1ull > nbr_cap_) {
        if (nbr_) dfree(nbr_);
        if (contrib_) dfree(contrib_);
        nbr_cap_ = (size_t)(nnz * 1.25) + 1024;
        nbr_ = dalloc<uint32_t>(nbr_cap_);
        if (opt_.exact_reductions) contrib_ = dalloc<double>(nbr_cap_ * 4);
    }
    la.nbr = nbr_;
    la.nbr_cap = nbr_cap_;
    if (launch_rows_fill(la, stream_)) throw HipFailure("rows fill launch");
    nnz_ = nnz;
    list_entries_ += nnz;
}

// ------------------------------------------------------------ device step

Engine::Scalars Engine::step_device(bool do_adam, double alpha, double inv_b1, double inv_b2,
                                    int use_rk, const double* rk)
{
    StepArgs a;
    a.g = geo_;
    a.x_in = x_[cur_];
    a.x_out = x_[1 - cur_];
    a.m = m_;
    a.v = v_;
    a.D0 = D0_;
    a.fixed = fixed_;
    a.x_anchor = x_anchor_;
    a.row_start = row_start_;
    a.nbr = nbr_;
    a.kappa = kappa_D_;
    a.use_rk = use_rk;
    a.rk = rk;
    a.alpha = alpha;
    a.inv_b1 = inv_b1;
    a.inv_b2 = inv_b2;
    a.do_adam = do_adam ? 1 : 0;
    a.partials = partials_;
    a.counter = counter_;
    a.result = result_;
    if (opt_.exact_reductions) {
        a.F_out = F_;
        a.mag_out = mag_;
        a.contrib = contrib_;
    }
    if (opt_.time_kernels) hip_check(hipEventRecord(ev_[0], stream_), "event");
    if (launch_step(a, stream_)) throw HipFailure("step launch");
    if (opt_.time_kernels) hip_check(hipEventRecord(ev_[1], stream_), "event");
    if (opt_.exact_reductions && launch_exact_reduce(a, stream_))
        throw HipFailure("exact reduce launch");
    hip_check(hipMemcpyAsync(h_result_, result_, S_NUM * sizeof(double), hipMemcpyDeviceToHost,
                             stream_), "scalars download");
    hip_check(hipStreamSynchronize(stream_), "step sync");
    if (opt_.time_kernels) {
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "event time");
        t_step_ms_ += ms;
        ++t_step_n_;
    }
    Scalars s;
    for (int k = 0; k < S_NUM; ++k) s.v[k] = h_result_[k];
    return s;
}

void Engine::finalize_scalars(const Scalars& s, double mu_eff, bool use_active)
{
    // RG:4191-4268 finalisation; count starts at 1
    const double count = 1.0 + s.v[S_CNT];
    U_ = std::pow(s.v[S_U] / count, 2);
    Lc_ = s.v[S_LC] / count;
    (void)use_active;
    dkappa_ = s.v[S_DK] + mu_eff * s.v[S_SUMD];
    Fmean_ = -s.v[S_FM] / count;
    max_min_dist_ = s.v[S_MAXOV];
    // compute_mean_force (RG:769-803): mean |F_i| / sqrt(Ndim)
    const double mean_mag = s.v[S_SMAG] / double((uint32_t)s.v[S_MCNT]);
    mean_force_ = mean_mag / std::sqrt(double(nd_));
    maxdisp2_ = s.v[S_MAXDISP2];
}

// ------------------------------------------------------------- schedule

static double mean_inclusive(const std::vector<double>& v, std::size_t start, std::size_t end)
{
    // RG:739-749: inclusive end, unsigned index arithmetic
    double sum = 0.0;
    for (std::size_t k = start; k <= end; ++k) sum += v[k];
    return sum / double(end - start + 1);
}

static double phi_range(const std::vector<double>& ph, std::size_t step)
{
    auto mm = std::minmax_element(ph.begin() + (step - 250), ph.begin() + (step - 1));
    return *mm.second - *mm.first;
}

bool Engine::schedule_top(std::size_t step)
{
    // mu decay ticks (RG:5529-5546)
    if (mu_decay_active_ && step >= mu_decay_next_tick_ && step > 250) {
        if (phi_range(phi_h_, step) < 5e-3) {
            mu_ *= 0.85;
            if (mu_ <= mu_decay_target_) {
                mu_ = mu_decay_target_;
                mu_decay_active_ = false;
            }
            mu_decay_next_tick_ = step + 250;
        }
    }
    // rung 0 -> oscillation start (RG:5555-5580)
    if (osc_ == Osc::Inactive && mu_flag_ == 0 && !mu_decay_active_ && mu_change_ > 0 &&
        step > mu_change_) {
        osc_mu_0_ = mu_;
        osc_mu_m1_ = mu_ / 10.0;
        osc_t_relax_ = step - mu_change_;
        if (osc_t_relax_ < 250) osc_t_relax_ = 250;
        osc_start_ = step;
        osc_period_ = static_cast<std::size_t>(0.5 * mu_change_);
        if (osc_period_ < 1000) osc_period_ = 1000;
        osc_ = Osc::Active;
        mu_flag_ = -1;
        mu_change_ = step;
        alpha_ /= 2.0;
    }
    // oscillation / final descent / convergence gate (RG:5588-5655).  The
    // gate reads F_history[step-1], the slot this step has not written yet.
    if (osc_ == Osc::Active) {
        const double t_into = static_cast<double>(step - osc_start_);
        const double cp = t_into / static_cast<double>(osc_period_);
        if (cp >= 2.0) {
            mu_ = osc_mu_m1_;
            osc_ = Osc::Final;
            osc_final_start_ = step;
            mu_decay_next_tick_ = step + 250;
        } else {
            const double ph = cp - std::floor(cp);
            const double sin_val = -std::sin(2.0 * 3.14159265358979323846 * ph);
            const double mu_ref = (osc_mu_0_ - osc_mu_m1_) *
                std::exp(-t_into / static_cast<double>(osc_t_relax_)) + osc_mu_m1_;
            mu_ = mu_ref * std::pow(10.0, sin_val);
            if (cp >= 2.0 - 0.25) {
                const double F_now = F_h_[step - 1];
                if (F_now < 5e-3 && max_min_dist_ < 5e-4) {
                    if (++osc_consec_ >= 3) osc_ = Osc::Done;
                } else {
                    osc_consec_ = 0;
                }
            }
        }
    } else if (osc_ == Osc::Final && step >= mu_decay_next_tick_ && step > 250) {
        if (phi_range(phi_h_, step) < 5e-3) {
            const double F_now = F_h_[step - 1];
            if (F_now < 5e-3 && max_min_dist_ < 5e-4) {
                if (++osc_consec_ >= 3) osc_ = Osc::Done;
            } else {
                osc_consec_ = 0;
                mu_ *= 0.85;
            }
            if (step - osc_final_start_ > 50000) osc_ = Osc::Done;
            mu_decay_next_tick_ = step + 250;
        }
    }
    return osc_ != Osc::Done;
}

void Engine::schedule_alpha(std::size_t step)
{
    // learning-rate rules (RG:5665-5721); phi is the previous step's value
    const std::size_t back = std::max<std::size_t>(step - 250, 1);
    if ((step - last_alpha_update_ > 500) && ((phi_ - phi_h_[back]) < 5e-4) && (step > 2500) &&
        step > mu_change_ + 500) {
        double trend1 = mean_inclusive(F_h_, step - 450, step - 250) /
                        mean_inclusive(F_h_, step - 75, step - 1);
        const double trend2 = phi_ - phi_h_[back];
        if ((mu_flag_ < 1 && trend1 < 0.85) || trend2 < -5E-4) {
            last_alpha_update_ = step;
            alpha_ /= 1.5;
        } else if (mu_flag_ < 1 && step - last_alpha_update_ > 1500) {
            trend1 = mean_inclusive(F_h_, step - 200, step - 1) /
                     mean_inclusive(F_h_, step - 3000, step - 2500);
            if (trend1 > 0.5) {
                last_alpha_update_ = step;
                alpha_ *= 1.15;
            }
        }
        if (mu_flag_ == 1) alpha_ = std::max(alpha_, alpha_max_ / 10.0);
    }
    alpha_ = std::min(alpha_ * 1.00005, alpha_max_);
    if (step % 500 == 0) alpha_ = std::min(alpha_max_, 1.1 * alpha_);
    if (mu_flag_ == -1) {
        const double cap = alpha_max_ * 0.01;
        if (alpha_ > cap) alpha_ = cap;
    }
}

double Engine::current_volume_now()
{
    const double coeff = std::pow(M_PI, nd_ / 2.0) / std::tgamma(nd_ / 2.0 + 1.0);
    if (has_fixed_) {
        // mixed scaling: sum over the current diameters (RG:5761-5766)
        std::vector<double> D0;
        std::vector<int8_t> fx;
        download_d0(D0, fx);
        double cv = 0.0;
        for (std::size_t i = 0; i < N_; ++i) {
            const double Di = fx[i] ? D0[i] : D0[i] * kappa_D_;
            cv += coeff * std::pow(Di / 2.0, nd_);
        }
        return cv;
    }
    double k2n = 1.0;
    if (!runtime_fix_)
        for (std::size_t k = 0; k < nd_; ++k) k2n *= kappa_;
    return coeff * S_D0_ * k2n;
}

void Engine::lock_target_phi(double box_volume)
{
    // RG:5783-5832: rescale D to the target, freeze diameters
    std::vector<double> D0;
    std::vector<int8_t> fx;
    download_d0(D0, fx);
    std::vector<double> D(N_);
    for (std::size_t i = 0; i < N_; ++i)
        D[i] = (!fx.empty() && fx[i]) ? D0[i] : D0[i] * kappa_D_;
    std::vector<double> Ds;
    scale_diameters(D, opt_.target_phi * phi_modifier_, box_, nd_, opt_.fix_height != 0, Ds);
    hip_check(hipMemcpyAsync(D0_, Ds.data(), N_ * sizeof(double), hipMemcpyHostToDevice, stream_),
              "upload D0");
    kappa_ = 1.0;
    kappa_D_ = 1.0;
    mu_ = 0.0;
    runtime_fix_ = true;
    target_locked_ = true;
    // D0 = D: every particle now carries its own diameter in D0; the
    // cascade mask keeps selecting D0 for frozen particles, which equals D
    recompute_d0_cache(Ds, fx);
    const double coeff = std::pow(M_PI, nd_ / 2.0) / std::tgamma(nd_ / 2.0 + 1.0);
    const double cv = coeff * S_D0_;
    phi_ = cv / box_volume;
    need_build_ = true;
    hip_check(hipStreamSynchronize(stream_), "sync");
}

bool Engine::step_once()
{
    const std::size_t step = hist_filled_ + 1;
    steps_var_ = step;
    t_ += 1;
    if (!schedule_top(step)) {
        done_ = true;
        return false;
    }
    // periodic spatial reorder (RG:5661-5663)
    if (opt_.reorder_interval > 0 && step > 1 && (step % opt_.reorder_interval == 0)) reorder();
    schedule_alpha(step);

    // growth (RG:5723-5834): D = D0 * kappa evaluated inside the kernels
    if (!runtime_fix_) kappa_D_ = kappa_;
    const double current_volume = current_volume_now();
    const double box_volume =
        std::accumulate(box_.begin(), box_.end(), 1.0, std::multiplies<double>());
    if (opt_.fix_height && !runtime_fix_) {
        const double L_old = box_[nd_ - 1];
        box_[nd_ - 1] = box_[nd_ - 1] * kappa_;
        geo_.box[nd_ - 1] = box_[nd_ - 1];
        geo_.inv_box[nd_ - 1] = 1.0 / box_[nd_ - 1];
        if (launch_scale_last_dim(x_[cur_], (int)N_, (int)nd_, kappa_, stream_))
            throw HipFailure("scale launch");
        fh_drift_ += std::fabs(box_[nd_ - 1] - L_old);
    }
    phi_ = current_volume / box_volume;
    if (opt_.has_target_phi && !target_locked_ && !runtime_fix_) {
        if (phi_ / phi_modifier_ >= opt_.target_phi) lock_target_phi(box_volume);
    }
    phi_h_[step - 1] = phi_ / phi_modifier_;

    // rung 1 -> 0 (RG:5840-5931)
    bool release_now = false;
    if (step > 500 && mu_flag_ == 1) {
        if (step > static_cast<std::size_t>(static_cast<int>(rung1_deadline_))) {
            mu_decay_target_ = mu_ / 10.0;
            mu_decay_active_ = true;
            mu_decay_next_tick_ = step + 250;
            alpha_ /= 2.0;
            mu_flag_ = 0;
            mu_change_ = step;
            release_now = has_fixed_ && !cascade_released_;
        }
    }

    // neighbour-list maintenance: rebuild when the skin could be crossed
    {
        const double growth = std::max(0.0, D0max_free_ * (kappa_D_ - kappa_b_));
        const double reach = 2.0 * std::sqrt(maxdisp2_) + growth + 2.0 * fh_drift_;
        if (need_build_ || !(reach < t_list_ * (1.0 - 1e-6))) build_lists();
    }

    // forces (RG:6211-6237) + fused ADAM (RG:6409-6521)
    const double mu_eff = (mu_flag_ == 1)
        ? mu_ * (1.0 + (5.0 - 1.0) * std::exp(-static_cast<double>(step) / 350.0)) : mu_;
    const double inv_b1 = 1.0 / (1.0 - std::pow(0.9, t_));
    const double inv_b2 = 1.0 / (1.0 - std::pow(0.999, t_));
    // In the release step the reference clears the mask before the force call
    // while D still holds the frozen sizes: all particles count as active.
    const double* rk = nullptr;
    if (release_now) {
        if (!rk_tmp_) rk_tmp_ = dalloc<double>(N_);
        if (launch_rk_active(D0_, fixed_, kappa_D_, rk_tmp_, (int)N_, stream_))
            throw HipFailure("rk launch");
        rk = rk_tmp_;
    }
    const Scalars s = step_device(true, alpha_, inv_b1, inv_b2, has_fixed_ ? 1 : 0, rk);
    finalize_scalars(s, mu_eff, has_fixed_);
    force_magnitude_ = mean_force_ / Fmean_;
    F_h_[step - 1] = force_magnitude_;
    ov_h_[step - 1] = max_min_dist_;
    mu_h_[step - 1] = mu_;
    al_h_[step - 1] = alpha_;
    fl_h_[step - 1] = static_cast<int8_t>(mu_flag_);

    // kappa ADAM (RG:6354-6358) and step (RG:6405-6407)
    m_kappa_ = 0.9 * m_kappa_ - (1 - 0.9) * dkappa_;
    v_kappa_ = 0.999 * v_kappa_ + (1 - 0.999) * (dkappa_ * dkappa_);
    const double m_hat = m_kappa_ / (1.0 - std::pow(0.9, t_));
    const double v_hat = v_kappa_ / (1.0 - std::pow(0.999, t_));
    if (!runtime_fix_) kappa_ = kappa_ - alpha_ * m_hat / (std::sqrt(v_hat) + 1e-8);

    cur_ = 1 - cur_;
    if (release_now) release_cascade();
    U_h_[step - 1] = U_;
    hist_filled_ = step;
    return true;
}

void Engine::release_cascade()
{
    // RG:5910-5921: D0 of frozen particles re-anchored so D = D0*kappa stays
    // continuous, then the mask is cleared
    std::vector<double> D0;
    std::vector<int8_t> fx;
    download_d0(D0, fx);
    for (std::size_t i = 0; i < N_; ++i) {
        if (fx[i]) D0[i] = D0[i] / kappa_D_;
        fx[i] = 0;
    }
    hip_check(hipMemcpyAsync(D0_, D0.data(), N_ * sizeof(double), hipMemcpyHostToDevice, stream_),
              "upload D0");
    hip_check(hipMemcpyAsync(fixed_, fx.data(), N_, hipMemcpyHostToDevice, stream_),
              "upload fixed");
    D0max_free_ = std::max(D0max_free_, D0max_fixed_ / kappa_D_);
    D0max_fixed_ = 0.0;
    shell0_ = reference_shell(D0);
    cascade_released_ = true;
    need_build_ = true;
    hip_check(hipStreamSynchronize(stream_), "sync");
}

void Engine::current_diameters(std::vector<double>& D)
{
    std::vector<double> D0;
    std::vector<int8_t> fx;
    download_d0(D0, fx);
    D.resize(N_);
    for (std::size_t i = 0; i < N_; ++i)
        D[i] = (!fx.empty() && fx[i]) ? D0[i] : D0[i] * kappa_D_;
}

void Engine::origin(std::vector<uint64_t>& o)
{
    o.resize(N_);
    hip_check(hipMemcpyAsync(o.data(), origin_, N_ * sizeof(uint64_t), hipMemcpyDeviceToHost,
                             stream_), "download origin");
    hip_check(hipStreamSynchronize(stream_), "sync");
}

void Engine::histories(double* phi, double* F, double* U, double* ov, double* mu, double* al,
                       int8_t* fl) const
{
    const std::size_t n = hist_filled_;
    auto cp = [n](double* dst, const std::vector<double>& src) {
        if (dst && n) std::memcpy(dst, src.data(), n * sizeof(double));
    };
    cp(phi, phi_h_);
    cp(F, F_h_);
    cp(U, U_h_);
    cp(ov, ov_h_);
    cp(mu, mu_h_);
    cp(al, al_h_);
    if (fl && n) std::memcpy(fl, fl_h_.data(), n);
}

void Engine::current_positions(std::vector<double>& x)
{
    x.resize(N_ * nd_);
    hip_check(hipMemcpyAsync(x.data(), x_[cur_], N_ * nd_ * sizeof(double),
                             hipMemcpyDeviceToHost, stream_), "download x");
    hip_check(hipStreamSynchronize(stream_), "sync");
}

int Engine::record(std::size_t step, bool capture, bool final_sample, rcp_observer_fn fn,
                   void* user)
{
    TraceSample t;
    t.step = step;
    t.phi = final_sample ? phi_ : phi_ / phi_modifier_;
    t.force = force_magnitude_;
    t.energy = U_;
    t.max_min_dist = max_min_dist_;
    if (capture || final_sample) current_diameters(t.diameters);
    if (capture) current_positions(t.positions);
    t.has_positions = capture;
    const bool keep = capture || final_sample;
    if (keep) trace_.push_back(t);
    if (fn) {
        rcp_observation o;
        o.step = step;
        o.phi = t.phi;
        o.force_magnitude = t.force;
        o.energy = t.energy;
        o.max_min_dist = t.max_min_dist;
        const double* pos = capture ? (keep ? trace_.back().positions.data() : nullptr) : nullptr;
        if (fn(user, &o, pos)) return 1;
    }
    return 0;
}

int Engine::run(std::size_t nsteps, std::size_t progress_interval, bool capture,
                std::size_t trajectory_interval, rcp_observer_fn fn, void* user)
{
    if (!initial_recorded_) {
        initial_recorded_ = true;
        const bool capture_initial = capture && trajectory_interval > 0;
        const bool report_initial = fn && progress_interval > 0;
        if (capture_initial || report_initial) {
            // initial sample: the trace keeps it only when capturing
            TraceSample t;
            t.step = 0;
            t.phi = phi_ / phi_modifier_;
            t.force = force_magnitude_;
            t.energy = U_;
            t.max_min_dist = max_min_dist_;
            if (capture_initial) {
                current_diameters(t.diameters);
                current_positions(t.positions);
                t.has_positions = true;
                trace_.push_back(t);
            }
            if (fn) {
                rcp_observation o{0, t.phi, t.force, t.energy, t.max_min_dist};
                if (fn(user, &o, capture_initial ? trace_.back().positions.data() : nullptr))
                    return RCP_E_CALLBACK;
            }
        }
    }
    std::size_t done_here = 0;
    while (!done_ && hist_filled_ < step_cap_ && (nsteps == 0 || done_here < nsteps)) {
        if (!step_once()) break;
        ++done_here;
        const std::size_t step = hist_filled_;
        const bool report = fn && progress_interval > 0 && (step % progress_interval == 0);
        const bool cap = capture && trajectory_interval > 0 && (step % trajectory_interval == 0);
        if (report || cap) {
            if (record(step, cap, false, fn, user)) return RCP_E_CALLBACK;
        }
    }
    if (hist_filled_ >= step_cap_) done_ = true;
    return RCP_OK;
}

int Engine::finish(bool capture, std::size_t trajectory_interval, std::size_t progress_interval,
                   rcp_observer_fn fn, void* user)
{
    // RG:6737-6763: a final sample unless the last step already produced one
    const std::size_t steps = steps_var_;
    const bool fac = capture && trajectory_interval > 0 && steps > 0 &&
                     (steps % trajectory_interval == 0);
    const bool far = fn && progress_interval > 0 && steps > 0 && (steps % progress_interval == 0);
    if ((capture && !fac) || (fn && !far)) {
        if (record(steps, capture, true, fn, user)) return RCP_E_CALLBACK;
    }
    return RCP_OK;
}

}  // namespace rcp

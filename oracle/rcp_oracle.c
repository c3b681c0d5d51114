/*
 * CPU oracle for the Packing.pack() hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded restatement of the reference engine
 * (KD-physics/RCPGenerator, python_code/python/src/core/rcp_generator.cpp =
 * RG below).  It follows the reference's 1-thread evaluation order op for op
 * so that, compiled without FMA contraction (-ffp-contract=off), its results
 * are bit-identical to the reference built with its portable flags.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   orc_forces_list / orc_forces   bit-exact vs the reference's own
 *        get_forces_nd_3 (oracle/_ref, built from the reference sources) on
 *        random 2/3/4/5-D configurations with periodic, flat-wall and curved
 *        containers; closed forms for one pair / one flat-wall / one curved
 *        contact; Newton-III (sum of pair forces = 0); F = -grad E by finite
 *        differences of E = sum (r - d)^2/2 ... (see the test file).
 *   orc_mean_force          brute force + known answer on a 2-particle case.
 *   orc_pairs_full          bit-exact vs the reference's get_pairs_kdtree
 *        (ndim <= 3) and get_pairs_nd_3 (ndim > 3) on full refresh.
 *   orc_reorder_perm        equality with the permutation the reference
 *        applies (particle_origin of a 0-step run) + sortedness of keys.
 *   orc_adam_positions      first-step known answer dx = alpha*F/(|F|+eps).
 *   orc_run_packing         bit-exact full runs vs the reference run_packing
 *        (4-particle acceptance case, 16-disk smoke case, 64-disk hard-wall
 *        case, curved containers, 4-D/5-D, options max_steps / mu /
 *        target_phi / fix_diameter / initial_phi_target / reorder_interval).
 */
#include "rcp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define ORC_MAXD 16

double orc_sphere_volume(double r, size_t ndim)
{
    /* rcp_generator.hpp:165-168 */
    return pow(M_PI, ndim / 2.0) * pow(r, (double)ndim) / tgamma(ndim / 2.0 + 1.0);
}

double orc_scale_diameters(const double* D, size_t N, double phi_target,
                           const double* box, size_t ndim, int fix_height,
                           double* out)
{
    /* RG:1162-1196 */
    double V_box = 1.0;
    for (size_t d = 0; d < ndim; ++d) V_box *= box[d];
    double sumVol = 0.0;
    for (size_t i = 0; i < N; ++i) sumVol += orc_sphere_volume(D[i] * 0.5, ndim);
    double factor;
    if (fix_height)
        factor = pow(phi_target * V_box / sumVol, 1.0 / (double)(ndim - 1));
    else
        factor = pow(phi_target * V_box / sumVol, 1.0 / (double)ndim);
    for (size_t i = 0; i < N; ++i) out[i] = D[i] * factor;
    return factor;
}

/* ------------------------------------------------------------------ forces */

static void walls_and_finalize(size_t N, size_t ndim, const double* x,
                               const double* D, const double* box,
                               const int8_t* walls, double mu, const double* rk,
                               double sumD_active, double* F, orc_force_out* o)
{
    /* RG:4198-4268: the wall loop always runs (~no_walls is always true). */
    const double K = 1.0;
    const int circle = walls[0] < 0;
    double dx[ORC_MAXD];
    for (size_t i = 0; i < N; ++i) {
        for (size_t d = 0; d < ndim; ++d) {
            if (walls[d] == 1) {
                double r_ij = D[i] / 2;
                for (size_t wall = 0; wall < 2; ++wall) {
                    dx[d] = box[d] * (double)wall - x[i * ndim + d];
                    double dist = fabs(dx[d]);
                    if (dist < r_ij) {
                        double F_mag = -2.0 * K * (r_ij / dist - 1.0);
                        o->U += 2.0 * (1.0 - dist / r_ij);
                        ++o->count;
                        double fcomp = F_mag * dx[d];
                        F[i * ndim + d] += fcomp;
                        o->dkappa = o->dkappa + 2 * K * (rk ? rk[i] : r_ij) * (dist - r_ij);
                        o->Fmean = o->Fmean + F_mag;
                    }
                }
            }
        }
        if (circle) {
            double r_ij = D[i] / 2;
            double d_ij = 0;
            double R = box[0] / 2;
            size_t T = (size_t)(-walls[0]);
            for (size_t M = 0; M < T; ++M)
                d_ij = d_ij + (x[i * ndim + M] - R) * (x[i * ndim + M] - R);
            d_ij = sqrt(d_ij);
            double delta = R - d_ij;
            if (delta < r_ij) {
                delta = d_ij;
                d_ij = 0;
                for (size_t M = 0; M < T; ++M) {
                    dx[M] = fabs(R / delta - 1) * (x[i * ndim + M] - R);
                    d_ij = d_ij + dx[M] * dx[M];
                }
                d_ij = sqrt(d_ij);
                double F_mag = -2.0 * K * fabs(r_ij / d_ij - 1.0);
                o->U = o->U + (1 - d_ij / r_ij);
                ++o->count;
                for (size_t M = 0; M < T; ++M)
                    F[i * ndim + M] = F[i * ndim + M] + F_mag * dx[M];
                o->dkappa = o->dkappa + 2 * K * (rk ? rk[i] : r_ij) * (d_ij - r_ij);
                o->Fmean = o->Fmean + F_mag;
            }
        }
    }
    o->U_raw = o->U;
    o->Lc_raw = o->Lc;
    o->Fmean_raw = o->Fmean;
    o->U = pow(o->U / (double)o->count, 2);
    o->Lc = o->Lc / (double)o->count;
    double sumD = 0.0;
    for (size_t i = 0; i < N; ++i) sumD += D[i];
    o->sumD = sumD;
    o->dkappa_raw = o->dkappa;
    o->dkappa = o->dkappa + mu * (sumD_active >= 0.0 ? sumD_active : sumD);
    o->Fmean = -o->Fmean / (double)o->count;
}

void orc_forces_list(size_t N, size_t ndim, const double* x, const double* D,
                     const double* box, const int8_t* walls, double mu,
                     const uint64_t* row_off, const uint32_t* nbr,
                     const double* rk, double sumD_active,
                     double* F, orc_force_out* o)
{
    /* RG:3597-4189 at one thread: a per-thread buffer Ft accumulates in row
     * order, then F = 0 + (0 + Ft) (the reduce at RG:4174-4188). */
    const double K = 1.0;
    double* Ft = (double*)calloc(N * ndim + 1, sizeof(double));
    double inv_box[ORC_MAXD];
    for (size_t d = 0; d < ndim; ++d) inv_box[d] = 1.0 / box[d];
    double U_red = 0.0, Lc_red = 0.0, dk_red = 0.0, Fm_red = 0.0, max_red = 0.0;
    uint64_t cnt_red = 0;
    for (size_t i = 0; i < N; ++i) {
        const double* xi = x + i * ndim;
        const double Di = D[i];
        for (uint64_t e = row_off[i]; e < row_off[i + 1]; ++e) {
            size_t j = nbr[e];
            double r_ij = 0.5 * (Di + D[j]);
            const double* xj = x + j * ndim;
            double d2 = 0.0;
            double dxl[ORC_MAXD];
            for (size_t d = 0; d < ndim; ++d) {
                double delta = xj[d] - xi[d];
                delta -= nearbyint(delta * inv_box[d]) * box[d];
                dxl[d] = delta;
                d2 += delta * delta;
            }
            const double r_ij_sq = r_ij * r_ij;
            if (d2 < r_ij_sq) {
                double dist = sqrt(d2);
                double F_mag = -K * (r_ij / dist - 1.0);
                double overlap = 1.0 - dist / r_ij;
                U_red += overlap;
                ++cnt_red;
                Lc_red += r_ij;
                if (overlap > max_red) max_red = overlap;
                for (size_t d = 0; d < ndim; ++d) {
                    double fcomp = F_mag * dxl[d];
                    Ft[i * ndim + d] += fcomp;
                    Ft[j * ndim + d] -= fcomp;
                }
                dk_red = dk_red + K * (rk ? (rk[i] + rk[j]) : r_ij) * (dist - r_ij);
                Fm_red = Fm_red + F_mag;
            }
        }
    }
    for (size_t k = 0; k < N * ndim; ++k) {
        double acc = 0.0;
        acc += Ft[k];
        F[k] = 0.0;
        F[k] += acc;
    }
    free(Ft);
    memset(o, 0, sizeof(*o));
    o->U = U_red;
    o->Lc = Lc_red;
    o->dkappa = dk_red;
    o->Fmean = Fm_red;
    o->count = 1 + cnt_red;
    o->pair_contacts = cnt_red;
    o->max_min_dist = max_red;
    o->dkappa_raw = 0.0;
    walls_and_finalize(N, ndim, x, D, box, walls, mu, rk, sumD_active, F, o);
}

/* Cell search for every pair with d2 < r_ij^2 (the force kernel's own
 * contact test, evaluated with the same operations).  Rows hold j > i in
 * ascending order.  Cells are at least 1.000001 * Dmax wide, so every
 * contact lies in the 3^d stencil; an axis with fewer than 3 cells is
 * treated as a single cell. */
typedef struct { uint32_t cell, idx; } cell_item;

static int cmp_cell_item(const void* a, const void* b)
{
    const cell_item* p = (const cell_item*)a;
    const cell_item* q = (const cell_item*)b;
    if (p->cell != q->cell) return p->cell < q->cell ? -1 : 1;
    return p->idx < q->idx ? -1 : (p->idx > q->idx);
}

static int cmp_u32(const void* a, const void* b)
{
    uint32_t p = *(const uint32_t*)a, q = *(const uint32_t*)b;
    return p < q ? -1 : (p > q);
}

static size_t contact_rows(size_t N, size_t ndim, const double* x,
                           const double* D, const double* box,
                           uint64_t** row_off_out, uint32_t** nbr_out)
{
    double Dmax = 0.0;
    for (size_t i = 0; i < N; ++i) if (D[i] > Dmax) Dmax = D[i];
    double inv_box[ORC_MAXD];
    for (size_t d = 0; d < ndim; ++d) inv_box[d] = 1.0 / box[d];
    size_t nc[ORC_MAXD], total = 1;
    for (size_t d = 0; d < ndim; ++d) {
        double h = Dmax * 1.000001;
        size_t n = (h > 0.0) ? (size_t)floor(box[d] / h) : 1;
        if (n < 3) n = 1;
        if (n > 4096) n = 4096;
        nc[d] = n;
        total *= n;
        if (total > (size_t)1 << 26) { nc[d] = 1; total /= n; }
    }
    cell_item* items = (cell_item*)malloc((N + 1) * sizeof(cell_item));
    for (size_t i = 0; i < N; ++i) {
        size_t c = 0, stride = 1;
        for (size_t d = 0; d < ndim; ++d) {
            double L = box[d];
            double xw = x[i * ndim + d];
            xw = xw - L * floor(xw / L);
            long k = (long)floor(xw / L * (double)nc[d]);
            if (k < 0) k = 0;
            if (k >= (long)nc[d]) k = (long)nc[d] - 1;
            c += (size_t)k * stride;
            stride *= nc[d];
        }
        items[i].cell = (uint32_t)c;
        items[i].idx = (uint32_t)i;
    }
    qsort(items, N, sizeof(cell_item), cmp_cell_item);
    size_t* cstart = (size_t*)calloc(total + 1, sizeof(size_t));
    for (size_t k = 0; k < N; ++k) cstart[items[k].cell + 1]++;
    for (size_t c = 0; c < total; ++c) cstart[c + 1] += cstart[c];
    size_t* cell_of = (size_t*)malloc((N + 1) * sizeof(size_t));
    for (size_t k = 0; k < N; ++k) cell_of[items[k].idx] = items[k].cell;

    size_t cap = 16 * N + 16, used = 0;
    uint32_t* nbr = (uint32_t*)malloc(cap * sizeof(uint32_t));
    uint64_t* off = (uint64_t*)malloc((N + 1) * sizeof(uint64_t));
    size_t nstencil = 1;
    for (size_t d = 0; d < ndim; ++d) nstencil *= (nc[d] >= 3 ? 3 : 1);
    for (size_t i = 0; i < N; ++i) {
        off[i] = used;
        size_t ci[ORC_MAXD];
        size_t c = cell_of[i];
        for (size_t d = 0; d < ndim; ++d) { ci[d] = c % nc[d]; c /= nc[d]; }
        for (size_t s = 0; s < nstencil; ++s) {
            size_t t = s, cc = 0, stride = 1;
            for (size_t d = 0; d < ndim; ++d) {
                long o = 0;
                if (nc[d] >= 3) { o = (long)(t % 3) - 1; t /= 3; }
                long k = (long)ci[d] + o;
                if (k < 0) k += (long)nc[d];
                if (k >= (long)nc[d]) k -= (long)nc[d];
                cc += (size_t)k * stride;
                stride *= nc[d];
            }
            for (size_t q = cstart[cc]; q < cstart[cc + 1]; ++q) {
                size_t j = items[q].idx;
                if (j <= i) continue;
                double r_ij = 0.5 * (D[i] + D[j]);
                double d2 = 0.0;
                for (size_t d = 0; d < ndim; ++d) {
                    double delta = x[j * ndim + d] - x[i * ndim + d];
                    delta -= nearbyint(delta * inv_box[d]) * box[d];
                    d2 += delta * delta;
                }
                if (d2 < r_ij * r_ij) {
                    if (used == cap) {
                        cap *= 2;
                        nbr = (uint32_t*)realloc(nbr, cap * sizeof(uint32_t));
                    }
                    nbr[used++] = (uint32_t)j;
                }
            }
        }
        qsort(nbr + off[i], used - off[i], sizeof(uint32_t), cmp_u32);
    }
    off[N] = used;
    free(items);
    free(cstart);
    free(cell_of);
    *row_off_out = off;
    *nbr_out = nbr;
    return used;
}

void orc_forces(size_t N, size_t ndim, const double* x, const double* D,
                const double* box, const int8_t* walls, double mu,
                const double* rk, double sumD_active,
                double* F, orc_force_out* out)
{
    uint64_t* off = NULL;
    uint32_t* nbr = NULL;
    contact_rows(N, ndim, x, D, box, &off, &nbr);
    orc_forces_list(N, ndim, x, D, box, walls, mu, off, nbr, rk, sumD_active, F, out);
    free(off);
    free(nbr);
}

double orc_mean_force(const double* F, size_t N, size_t ndim)
{
    /* RG:769-803 */
    double sum_mag = 0.0;
    uint32_t count = 0;
    for (size_t i = 0; i < N; ++i) {
        double d2 = 0.0;
        for (size_t d = 0; d < ndim; ++d) {
            double Fid = F[i * ndim + d];
            d2 += Fid * Fid;
        }
        if (d2 > 1.0E-16) {
            sum_mag += sqrt(d2);
            ++count;
        }
    }
    double acc = 0.0;
    acc += sum_mag;
    double mean_mag = acc / (double)count;
    return mean_mag / sqrt((double)ndim);
}

/* ------------------------------------------------------------------- pairs */

static int cmp_double(const void* a, const void* b)
{
    double p = *(const double*)a, q = *(const double*)b;
    return p < q ? -1 : (p > q);
}

double orc_shell(const double* D, size_t N)
{
    /* RG:2846 (kd-tree) and RG:1215-1222 (sort+walk): median = the element at
     * position N/2 of the ascending order. */
    double* s = (double*)malloc((N + 1) * sizeof(double));
    memcpy(s, D, N * sizeof(double));
    qsort(s, N, sizeof(double), cmp_double);
    double median = s[N / 2];
    double Dmin = s[0], Dmax = s[N - 1];
    free(s);
    double a = median * 1.02, b = Dmin * 2.25;
    double mx = (a < b) ? b : a;
    double mn = (Dmax < mx) ? Dmax : mx;
    return mn * 0.5;
}

typedef struct { double x0; uint32_t idx; } walk_item;

static int cmp_walk(const void* a, const void* b)
{
    const walk_item* p = (const walk_item*)a;
    const walk_item* q = (const walk_item*)b;
    if (p->x0 < q->x0) return -1;
    if (p->x0 > q->x0) return 1;
    return p->idx < q->idx ? -1 : (p->idx > q->idx);
}

size_t orc_pairs_full(size_t N, size_t ndim, const double* x, const double* D,
                      const double* box, uint64_t* row_off, uint32_t* nbr,
                      size_t cap)
{
    const double t = orc_shell(D, N);
    double inv_box[ORC_MAXD];
    for (size_t d = 0; d < ndim; ++d) inv_box[d] = 1.0 / box[d];
    size_t used = 0;
    if (ndim <= 3) {
        /* kd-tree backend (RG:3276-3297): every pair passing the test once */
        for (size_t i = 0; i < N; ++i) {
            row_off[i] = used;
            for (size_t j = i + 1; j < N; ++j) {
                double d2 = 0.0;
                for (size_t d = 0; d < ndim; ++d) {
                    double dd = x[j * ndim + d] - x[i * ndim + d];
                    dd -= nearbyint(dd * inv_box[d]) * box[d];
                    d2 += dd * dd;
                }
                const double r_c = (D[i] + D[j]) * 0.5 + t;
                if (d2 >= r_c * r_c) continue;
                if (nbr && used < cap) nbr[used] = (uint32_t)j;
                ++used;
            }
        }
        row_off[N] = used;
        return used;
    }
    /* sort+walk backend (RG:1215-1459) at full refresh: from each i walk the
     * x-sorted order in both directions up to N/2 places, stopping at the
     * first |dx| > r_c,max; own-write partners j > i without a duplicate
     * check, then sort each row. */
    double Dmax = D[0];
    for (size_t i = 1; i < N; ++i) if (D[i] > Dmax) Dmax = D[i];
    walk_item* w = (walk_item*)malloc((N + 1) * sizeof(walk_item));
    size_t* loc = (size_t*)malloc((N + 1) * sizeof(size_t));
    for (size_t i = 0; i < N; ++i) { w[i].x0 = x[i * ndim]; w[i].idx = (uint32_t)i; }
    qsort(w, N, sizeof(walk_item), cmp_walk);
    for (size_t k = 0; k < N; ++k) loc[w[k].idx] = k;
    uint32_t* row = (uint32_t*)malloc((N + 1) * sizeof(uint32_t));
    for (size_t i = 0; i < N; ++i) {
        row_off[i] = used;
        size_t cnt = 0;
        const double r_c_max = (D[i] + 1.01 * Dmax) / 2 + t;
        const double Di_half_plus_t = D[i] * 0.5 + t;
        for (int dir = -1; dir <= 1; dir += 2) {
            long long delta = 0;
            for (size_t jdx = 1; jdx <= N / 2; ++jdx) {
                delta += dir;
                long long pos = (long long)loc[i] + delta;
                if (pos < 0) pos += (long long)N;
                else if (pos >= (long long)N) pos -= (long long)N;
                const size_t j = w[pos].idx;
                double dx = x[j * ndim] - x[i * ndim];
                dx -= nearbyint(dx * inv_box[0]) * box[0];
                if (!(fabs(dx) <= r_c_max)) break;
                const double r_c = Di_half_plus_t + (double)(float)(D[j] * 0.5);
                if (fabs(dx) < r_c) {
                    double d2 = dx * dx;
                    for (size_t d = 1; d < ndim; ++d) {
                        double dz = x[j * ndim + d] - x[i * ndim + d];
                        dz -= nearbyint(dz * inv_box[d]) * box[d];
                        d2 += dz * dz;
                    }
                    if (d2 >= r_c * r_c) continue;
                    if (j > i) row[cnt++] = (uint32_t)j;
                }
            }
        }
        qsort(row, cnt, sizeof(uint32_t), cmp_u32);
        for (size_t k = 0; k < cnt; ++k) {
            if (nbr && used < cap) nbr[used] = row[k];
            ++used;
        }
    }
    row_off[N] = used;
    free(w); free(loc); free(row);
    return used;
}

/* ----------------------------------------------------------------- reorder */

static uint64_t morton_split21(uint64_t v)
{
    /* interleave the low 21 bits with two zero bits between each */
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1ull) << (3 * b);
    return r;
}

typedef struct { uint64_t key; uint32_t idx; double xv; } perm_item;

static int cmp_morton(const void* a, const void* b)
{
    const perm_item* p = (const perm_item*)a;
    const perm_item* q = (const perm_item*)b;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    return p->idx < q->idx ? -1 : (p->idx > q->idx);
}

static int cmp_xcol(const void* a, const void* b)
{
    const perm_item* p = (const perm_item*)a;
    const perm_item* q = (const perm_item*)b;
    if (p->xv < q->xv) return -1;
    if (p->xv > q->xv) return 1;
    return p->idx < q->idx ? -1 : (p->idx > q->idx);
}

void orc_reorder_perm(const double* x, size_t N, size_t ndim,
                      const double* box, uint32_t* perm)
{
    perm_item* it = (perm_item*)malloc((N + 1) * sizeof(perm_item));
    if (ndim == 3) {
        /* RG:1127-1160 */
        double scale[3];
        for (int d = 0; d < 3; ++d) scale[d] = ((double)(1u << 21) - 1.0) / box[d];
        for (size_t i = 0; i < N; ++i) {
            uint64_t code = 0;
            for (int d = 0; d < 3; ++d) {
                double hi = box[d] - 1e-12, v = x[i * 3 + d];
                double m = (v < hi) ? v : hi;
                double c = (0.0 < m) ? m : 0.0;
                uint32_t q = (uint32_t)(c * scale[d]);
                code |= morton_split21(q) << d;
            }
            it[i].key = code;
            it[i].idx = (uint32_t)i;
        }
        qsort(it, N, sizeof(perm_item), cmp_morton);
    } else {
        /* RG:1084-1105 */
        for (size_t i = 0; i < N; ++i) {
            it[i].xv = x[i * ndim];
            it[i].idx = (uint32_t)i;
        }
        qsort(it, N, sizeof(perm_item), cmp_xcol);
    }
    for (size_t k = 0; k < N; ++k) perm[k] = it[k].idx;
    free(it);
}

/* -------------------------------------------------------------------- ADAM */

static const double BETA1 = 0.9, BETA2 = 0.999, EPSILON = 1e-8;

void orc_adam_positions(size_t N, size_t ndim, const double* F, double* m,
                        double* v, double* x, const double* box, double alpha,
                        double inv_b1, double inv_b2)
{
    /* RG:6499-6520 */
    for (size_t k = 0; k < N; ++k) {
        for (size_t kk = 0; kk < ndim; ++kk) {
            size_t idx = k * ndim + kk;
            double Fkd = F[idx];
            double mk = BETA1 * m[idx] - (1.0 - BETA1) * Fkd;
            double vk = BETA2 * v[idx] + (1.0 - BETA2) * Fkd * Fkd;
            m[idx] = mk;
            v[idx] = vk;
            double mhk = mk * inv_b1;
            double vhk = vk * inv_b2;
            double xv = x[idx] - alpha * mhk / (sqrt(vhk) + EPSILON);
            double L = box[kk];
            if (xv >= L) xv -= L;
            else if (xv < 0) xv += L;
            x[idx] = xv;
        }
    }
}

/* ------------------------------------------------------------------ driver */

static double rung1_deadline_d(size_t N, size_t ndim)
{
    /* RG:4722-4734 and RG:5873-5900, speed preset "quick" (Y=2, X=9) */
    double dl = (double)2 * 4000.0 + (double)9 * sqrt((double)N * (double)ndim) - 2000.0;
    if (dl < 1500.0) dl = 1500.0;
    return dl;
}

size_t orc_step_cap(size_t N, size_t ndim, size_t max_steps)
{
    if (max_steps > 0) return max_steps;
    double rung1 = rung1_deadline_d(N, ndim);
    size_t formula = (size_t)(rung1 * (double)(2 + 3));
    return formula > 60000 ? formula : 60000;
}

static double mean_range(const double* v, size_t start, size_t end)
{
    /* RG:739-749, `end` inclusive; unsigned arithmetic as in the reference */
    double sum = 0.0;
    for (size_t k = start; k <= end; ++k) sum += v[k];
    return sum / (double)(end - start + 1);
}

static double phi_range_250(const double* ph, size_t step)
{
    size_t lo = step - 250, hi = step - 1;   /* [lo, hi) */
    double mn = ph[lo], mx = ph[lo];
    for (size_t k = lo + 1; k < hi; ++k) {
        if (ph[k] < mn) mn = ph[k];
        if (!(ph[k] < mx)) mx = ph[k];
    }
    return mx - mn;
}

static void apply_perm_d(double* a, const uint32_t* perm, size_t N, size_t w, double* tmp)
{
    for (size_t k = 0; k < N; ++k)
        for (size_t d = 0; d < w; ++d) tmp[k * w + d] = a[(size_t)perm[k] * w + d];
    memcpy(a, tmp, N * w * sizeof(double));
}

enum { OSC_INACTIVE, OSC_ACTIVE, OSC_FINAL, OSC_DONE };

int orc_run_packing(const orc_run_in* in, orc_run_out* out)
{
    const size_t N = in->N, nd = in->ndim;
    out->error = 0;
    double box[ORC_MAXD];
    int8_t walls[ORC_MAXD];
    if (in->nbox == 0) { for (size_t d = 0; d < nd; ++d) box[d] = 1.0; }
    else if (in->nbox != nd) { out->error = 1; return 1; }
    else memcpy(box, in->box, nd * sizeof(double));
    if (in->nwalls == 0) memset(walls, 0, nd);
    else if (in->nwalls != nd) { out->error = 2; return 2; }
    else memcpy(walls, in->walls, nd);

    double* x = (double*)malloc(N * nd * sizeof(double) + 8);
    double* D = (double*)malloc(N * sizeof(double) + 8);
    double* D0 = (double*)malloc(N * sizeof(double) + 8);
    double* F = (double*)calloc(N * nd + 1, sizeof(double));
    double* m = (double*)calloc(N * nd + 1, sizeof(double));
    double* v = (double*)calloc(N * nd + 1, sizeof(double));
    double* tmp = (double*)malloc(N * nd * sizeof(double) + 8);
    double* rk = (double*)malloc(N * sizeof(double) + 8);
    uint32_t* perm = (uint32_t*)malloc(N * sizeof(uint32_t) + 8);
    uint64_t* origin = (uint64_t*)malloc(N * sizeof(uint64_t) + 8);
    int8_t* fixed = (int8_t*)calloc(N + 1, 1);
    memcpy(x, in->x, N * nd * sizeof(double));
    memcpy(D, in->D, N * sizeof(double));
    int has_fixed = 0;
    if (in->fixed) {
        memcpy(fixed, in->fixed, N);
        for (size_t i = 0; i < N; ++i) if (fixed[i]) has_fixed = 1;
    }
    int cascade_released = 0;

    /* RG:4449-4458 */
    double phi = 0.025, phi_modifier = 1;
    if (walls[0] < 0) {
        if (walls[0] == -1) walls[0] = -2;
        for (size_t k = 1; k < (size_t)(-walls[0]) && k < nd; ++k) {
            walls[k] = -1;
            box[k] = box[0];
        }
        phi_modifier = orc_sphere_volume(1 / 2., (size_t)(-walls[0]));
    }
    if (in->fix_height) box[nd - 1] = box[nd - 1] * D[0];
    if (in->has_initial_phi) phi = in->initial_phi;
    else if (nd == 2) phi = 0.575;
    else if (nd == 3) phi = 0.33;
    else if (nd == 4) phi = 0.15;
    else if (nd == 5) phi = 0.10;
    else if (nd > 5) phi = 0.10 / pow(2, (double)(nd - 4));
    if (!in->fix_diameter) {
        double factor = orc_scale_diameters(D, N, phi * phi_modifier, box, nd,
                                            in->fix_height, tmp);
        if (in->fix_height) {
            box[nd - 1] = box[nd - 1] * factor;
            for (size_t i = 0; i < N; ++i) x[i * nd + nd - 1] = x[i * nd + nd - 1] * factor;
        }
        memcpy(D, tmp, N * sizeof(double));
    }
    double coeff = pow(M_PI, nd / 2.0) / tgamma(nd / 2.0 + 1.0);
    double current_volume = 0.0;
    for (size_t i = 0; i < N; ++i) current_volume += coeff * pow(D[i] / 2.0, (double)nd);
    double box_volume = 1.0;
    for (size_t d = 0; d < nd; ++d) box_volume = box_volume * box[d];
    phi = current_volume / box_volume / phi_modifier;

    const size_t step_cap = orc_step_cap(N, nd, in->max_steps);
    const size_t hist_size = step_cap + 1;
    double* phi_h = (double*)calloc(hist_size, sizeof(double));
    double* F_h = (double*)calloc(hist_size, sizeof(double));
    double* U_h = (double*)calloc(hist_size, sizeof(double));
    double* ov_h = (double*)calloc(hist_size, sizeof(double));
    double* mu_h = (double*)calloc(hist_size, sizeof(double));
    double* al_h = (double*)calloc(hist_size, sizeof(double));
    int8_t* fl_h = (int8_t*)calloc(hist_size, 1);

    const double R = 1.00005;
    double t = 0;
    double alpha = 0.005, alpha_max = 0.005;
    if (nd == 3) { alpha = 0.0045; alpha_max = 0.0045; }
    if (nd == 4) { alpha = 0.0035; alpha_max = 0.0035; }
    if (nd >= 5) { alpha = 0.0025; alpha_max = 0.0025; }
    double mu = in->has_mu ? in->mu : 2.5e-4;
    int runtime_fix = in->fix_diameter;
    int target_locked = 0;
    int mu_flag = 1;
    double mu_change = 0.;
    double kappa = 1.0, dkappa = 1.0;
    double m_kappa = 0, v_kappa = 0, m_hat_kappa = 0, v_hat_kappa = 0;
    size_t LastAlphaUpdate = 0;
    double F_magnitude = 0.0;

    memcpy(D0, D, N * sizeof(double));
    /* RG:4808-4826: D0 caches, computed before the initial reorder */
    double D0_min = D0[0], S_D0 = 0.0;
    for (size_t i = 0; i < N; ++i) {
        if (D0[i] < D0_min) D0_min = D0[i];
        double r = D0[i] * 0.5, rN = 1.0;
        for (size_t k = 0; k < nd; ++k) rN *= r;
        S_D0 += rN;
    }
    for (size_t k = 0; k < N; ++k) origin[k] = k;

    const size_t RI = in->reorder_interval;
    #define ORC_REORDER() do {                                                   \
        orc_reorder_perm(x, N, nd, box, perm);                                   \
        for (size_t k = 0; k < N; ++k) ((uint64_t*)tmp)[k] = origin[perm[k]];   \
        memcpy(origin, tmp, N * sizeof(uint64_t));                               \
        apply_perm_d(x, perm, N, nd, tmp);                                       \
        apply_perm_d(D, perm, N, 1, tmp);                                        \
        apply_perm_d(D0, perm, N, 1, tmp);                                       \
        if (in->fixed) {                                                         \
            int8_t* tf = (int8_t*)tmp;                                           \
            for (size_t k = 0; k < N; ++k) tf[k] = fixed[perm[k]];               \
            memcpy(fixed, tf, N);                                                \
        }                                                                        \
        apply_perm_d(m, perm, N, nd, tmp);                                       \
        apply_perm_d(v, perm, N, nd, tmp);                                       \
    } while (0)

    if (RI > 0) ORC_REORDER();

    orc_force_out fo;
    orc_forces(N, nd, x, D, box, walls, mu, NULL, -1.0, F, &fo);
    double U = fo.U, max_min_dist = fo.max_min_dist, Fmean = fo.Fmean;
    dkappa = fo.dkappa;

    int osc_phase = OSC_INACTIVE;
    double osc_mu_0 = 0.0, osc_mu_minus1 = 0.0;
    size_t osc_start_step = 0, osc_period = 0, osc_t_relax = 0;
    size_t osc_final_start = 0, osc_consec = 0;
    int mu_decay_active = 0;
    double mu_decay_target = 0.0;
    size_t mu_decay_next_tick = 0;
    size_t steps = 0, hist_filled = 0;
    (void)steps;

    for (size_t step = 1; step <= step_cap; ++step) {
        steps = step;
        t += 1;
        /* RG:5529-5546 */
        if (mu_decay_active && step >= mu_decay_next_tick && step > 250) {
            if (phi_range_250(phi_h, step) < 5e-3) {
                mu *= 0.85;
                if (mu <= mu_decay_target) { mu = mu_decay_target; mu_decay_active = 0; }
                mu_decay_next_tick = step + 250;
            }
        }
        /* RG:5555-5580 */
        if (osc_phase == OSC_INACTIVE && mu_flag == 0 && !mu_decay_active &&
            mu_change > 0 && (double)step > mu_change) {
            osc_mu_0 = mu;
            osc_mu_minus1 = mu / 10.0;
            osc_t_relax = (size_t)((double)step - mu_change);
            if (osc_t_relax < 250) osc_t_relax = 250;
            osc_start_step = step;
            osc_period = (size_t)(0.5 * mu_change);
            if (osc_period < 1000) osc_period = 1000;
            osc_phase = OSC_ACTIVE;
            mu_flag = -1;
            mu_change = (double)step;
            alpha /= 2.0;
        }
        /* RG:5588-5655 */
        if (osc_phase == OSC_ACTIVE) {
            const double t_into = (double)(step - osc_start_step);
            const double cp = t_into / (double)osc_period;
            if (cp >= (double)2) {
                mu = osc_mu_minus1;
                osc_phase = OSC_FINAL;
                osc_final_start = step;
                mu_decay_next_tick = step + 250;
            } else {
                const double ph = cp - floor(cp);
                const double sin_val = -sin(2.0 * 3.14159265358979323846 * ph);
                const double mu_ref = (osc_mu_0 - osc_mu_minus1) *
                    exp(-t_into / (double)osc_t_relax) + osc_mu_minus1;
                mu = mu_ref * pow(10.0, sin_val);
                if (cp >= (double)2 - 0.25) {
                    const double F_now = F_h[step - 1];
                    if ((F_now < 5e-3) && (max_min_dist < 5e-4)) {
                        osc_consec += 1;
                        if (osc_consec >= 3) osc_phase = OSC_DONE;
                    } else {
                        osc_consec = 0;
                    }
                }
            }
        } else if (osc_phase == OSC_FINAL && step >= mu_decay_next_tick && step > 250) {
            if (phi_range_250(phi_h, step) < 5e-3) {
                const double F_now = F_h[step - 1];
                if ((F_now < 5e-3) && (max_min_dist < 5e-4)) {
                    osc_consec += 1;
                    if (osc_consec >= 3) osc_phase = OSC_DONE;
                } else {
                    osc_consec = 0;
                    mu *= 0.85;
                }
                if (step - osc_final_start > 50000) osc_phase = OSC_DONE;
                mu_decay_next_tick = step + 250;
            }
        }
        if (osc_phase == OSC_DONE) { steps = step; break; }

        /* RG:5661-5663 */
        if (RI > 0 && step > 1 && (step % RI == 0)) ORC_REORDER();

        /* RG:5665-5721 */
        if ((step - LastAlphaUpdate > 500) &&
            ((phi - phi_h[(step - 250) > 1 ? (step - 250) : 1]) < 5e-4) &&
            (step > 2500) && (double)step > mu_change + 500) {
            double trend1 = mean_range(F_h, step - 450, step - 250) /
                            mean_range(F_h, step - 75, step - 1);
            double trend2 = (phi - phi_h[(step - 250) > 1 ? (step - 250) : 1]);
            if ((mu_flag < 1 && trend1 < 0.85) || trend2 < -5E-4) {
                LastAlphaUpdate = step;
                alpha /= 1.5;
            } else if (mu_flag < 1 && step - LastAlphaUpdate > 1500) {
                trend1 = mean_range(F_h, step - 200, step - 1) /
                         mean_range(F_h, step - 3000, step - 2500);
                if (trend1 > 0.5) {
                    LastAlphaUpdate = step;
                    alpha *= 1.15;
                }
            }
            if (mu_flag == 1) {
                double lo = alpha_max / 10.0;
                alpha = (alpha < lo) ? lo : alpha;
            }
        }
        {
            double a = alpha * R;
            alpha = (alpha_max < a) ? alpha_max : a;
        }
        if (step % 500 == 0) {
            double a = 1.1 * alpha;
            alpha = (a < alpha_max) ? a : alpha_max;
        }
        if (mu_flag == -1) {
            const double cap = alpha_max * 0.01;
            if (alpha > cap) alpha = cap;
        }

        /* RG:5723-5743 */
        if (!runtime_fix) {
            if (!has_fixed) for (size_t i = 0; i < N; ++i) D[i] = D0[i] * kappa;
            else for (size_t i = 0; i < N; ++i) D[i] = fixed[i] ? D0[i] : D0[i] * kappa;
        }
        coeff = pow(M_PI, nd / 2.0) / tgamma(nd / 2.0 + 1.0);
        double k2n = 1.0;
        if (!runtime_fix) for (size_t k = 0; k < nd; ++k) k2n *= kappa;
        if (has_fixed) {
            current_volume = 0.0;
            for (size_t i = 0; i < N; ++i) current_volume += coeff * pow(D[i] / 2.0, (double)nd);
        } else {
            current_volume = coeff * S_D0 * k2n;
        }
        box_volume = 1.0;
        for (size_t d = 0; d < nd; ++d) box_volume = box_volume * box[d];
        if (in->fix_height && !runtime_fix) {
            box[nd - 1] = box[nd - 1] * kappa;
            for (size_t i = 0; i < N; ++i) x[i * nd + nd - 1] = x[i * nd + nd - 1] * kappa;
        }
        phi = current_volume / box_volume;

        /* RG:5783-5832 */
        if (in->has_target_phi && !target_locked && !runtime_fix) {
            const double user_phi = phi / phi_modifier;
            if (user_phi >= in->target_phi) {
                orc_scale_diameters(D, N, in->target_phi * phi_modifier, box, nd,
                                    in->fix_height, tmp);
                memcpy(D, tmp, N * sizeof(double));
                memcpy(D0, D, N * sizeof(double));
                kappa = 1.0;
                mu = 0.0;
                runtime_fix = 1;
                target_locked = 1;
                D0_min = D0[0];
                S_D0 = 0.0;
                for (size_t i = 0; i < N; ++i) {
                    if (D0[i] < D0_min) D0_min = D0[i];
                    double r = D0[i] * 0.5, rN = 1.0;
                    for (size_t k = 0; k < nd; ++k) rN *= r;
                    S_D0 += rN;
                }
                coeff = pow(M_PI, nd / 2.0) / tgamma(nd / 2.0 + 1.0);
                current_volume = coeff * S_D0;
                phi = current_volume / box_volume;
            }
        }
        phi_h[step - 1] = phi / phi_modifier;

        /* RG:5840-5931 */
        if (step > 500 && mu_flag == 1) {
            const size_t deadline = (size_t)rung1_deadline_d(N, nd);
            if (step > (size_t)(int)deadline) {
                mu_decay_target = mu / 10.0;
                mu_decay_active = 1;
                mu_decay_next_tick = step + 250;
                alpha /= 2.0;
                mu_flag = 0;
                mu_change = (double)step;
                if (has_fixed && !cascade_released) {
                    for (size_t ci = 0; ci < N; ++ci) {
                        if (fixed[ci]) D0[ci] = D[ci] / kappa;
                        fixed[ci] = 0;
                    }
                    cascade_released = 1;
                }
            }
        }

        /* RG:6213-6237 */
        const double mu_eff = (mu_flag == 1)
            ? mu * (1.0 + (5.0 - 1.0) * exp(-(double)step / 350.0)) : mu;
        const double* rkp = NULL;
        double sumD_active = -1.0;
        if (has_fixed) {
            sumD_active = 0.0;
            for (size_t i = 0; i < N; ++i) {
                if (fixed[i]) rk[i] = 0.0;
                else { rk[i] = D[i] * 0.5; sumD_active += D[i]; }
            }
            rkp = rk;
        }
        orc_forces(N, nd, x, D, box, walls, mu_eff, rkp, sumD_active, F, &fo);
        U = fo.U;
        max_min_dist = fo.max_min_dist;
        Fmean = fo.Fmean;
        dkappa = fo.dkappa;

        /* RG:6250-6260 */
        F_magnitude = orc_mean_force(F, N, nd) / Fmean;
        F_h[step - 1] = F_magnitude;
        ov_h[step - 1] = max_min_dist;
        mu_h[step - 1] = mu;
        al_h[step - 1] = alpha;
        fl_h[step - 1] = (int8_t)mu_flag;

        /* RG:6354-6358 */
        m_kappa = BETA1 * m_kappa - (1 - BETA1) * dkappa;
        v_kappa = BETA2 * v_kappa + (1 - BETA2) * (dkappa * dkappa);
        m_hat_kappa = m_kappa / (1.0 - pow(BETA1, t));
        v_hat_kappa = v_kappa / (1.0 - pow(BETA2, t));

        /* RG:6405-6407 */
        if (!runtime_fix)
            kappa = kappa - alpha * m_hat_kappa / (sqrt(v_hat_kappa) + EPSILON);

        /* RG:6409-6521 */
        const double inv_b1 = 1.0 / (1.0 - pow(BETA1, t));
        const double inv_b2 = 1.0 / (1.0 - pow(BETA2, t));
        orc_adam_positions(N, nd, F, m, v, x, box, alpha, inv_b1, inv_b2);

        U_h[step - 1] = U;
        hist_filled = step;
    }
    #undef ORC_REORDER

    memcpy(out->x, x, N * nd * sizeof(double));
    memcpy(out->D, D, N * sizeof(double));
    memcpy(out->origin, origin, N * sizeof(uint64_t));
    memcpy(out->box, box, nd * sizeof(double));
    memcpy(out->walls, walls, nd);
    size_t nh = hist_filled < out->hist_cap ? hist_filled : out->hist_cap;
    memcpy(out->phi_hist, phi_h, nh * sizeof(double));
    memcpy(out->F_hist, F_h, nh * sizeof(double));
    memcpy(out->U_hist, U_h, nh * sizeof(double));
    memcpy(out->maxov_hist, ov_h, nh * sizeof(double));
    memcpy(out->mu_hist, mu_h, nh * sizeof(double));
    memcpy(out->alpha_hist, al_h, nh * sizeof(double));
    memcpy(out->muflag_hist, fl_h, nh);
    out->steps = hist_filled;
    out->phi = phi;
    out->max_min_dist = max_min_dist;
    out->force_magnitude = F_magnitude;
    out->kappa = kappa;

    free(x); free(D); free(D0); free(F); free(m); free(v); free(tmp); free(rk);
    free(perm); free(origin); free(fixed);
    free(phi_h); free(F_h); free(U_h); free(ov_h); free(mu_h); free(al_h); free(fl_h);
    (void)Fmean; (void)U;
    return 0;
}

/*
 * CPU oracle for the Packing.pack() hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C, single-threaded restatement of the reference engine's numerics
 * (KD-physics/RCPGenerator, python_code/python/src/core/rcp_generator.cpp,
 * cited below as RG:<line>).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker; the
 * product (rcpgenerator_amd) never links or calls it.
 *
 * Every function follows the reference's 1-thread (OpenMP T=1) evaluation
 * order, so its outputs are bit-identical to the reference built without
 * FMA contraction (the portable wheel flags).  How each function is pinned
 * (golden vectors from the reference build, closed forms, invariants,
 * brute force) is listed in the header of rcp_oracle.c.
 */
#ifndef RCP_ORACLE_H
#define RCP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scalars produced by one force evaluation (RG:3597-4269). */
typedef struct {
    double U;            /* (sum overlaps / count)^2 (RG:4259) */
    double Lc;           /* sum r_ij / count */
    double Fmean;        /* -sum F_mag / count */
    double dkappa;       /* sum r(dist-r) + walls + mu * sum D */
    double max_min_dist; /* max pair overlap 1 - dist/r */
    double U_raw;        /* sum of overlaps before finalisation */
    double Lc_raw, Fmean_raw, dkappa_raw, sumD;
    uint64_t count;      /* 1 + #pair contacts + #wall contacts */
    uint64_t pair_contacts;
} orc_force_out;

/* pi^(d/2)/Gamma(d/2+1) * r^d  (rcp_generator.hpp:165-168). */
double orc_sphere_volume(double r, size_t ndim);

/* scale_diameters_nd (RG:1162-1196); writes the scaled diameters to out
 * and returns the factor. */
double orc_scale_diameters(const double* D, size_t N, double phi_target,
                           const double* box, size_t ndim, int fix_height,
                           double* out);

/* Force evaluation on an explicit half list (row i holds partners j > i in
 * ascending order), exactly get_forces_nd_3 at T=1 (RG:3597-4269).
 * rk may be NULL (standard path); sumD_active < 0 selects sum(D). */
void orc_forces_list(size_t N, size_t ndim, const double* x, const double* D,
                     const double* box, const int8_t* walls, double mu,
                     const uint64_t* row_off, const uint32_t* nbr,
                     const double* rk, double sumD_active,
                     double* F, orc_force_out* out);

/* Same, with the contact list found internally (cell search over all
 * pairs with d < (Di+Dj)/2, periodic minimum image on every axis). */
void orc_forces(size_t N, size_t ndim, const double* x, const double* D,
                const double* box, const int8_t* walls, double mu,
                const double* rk, double sumD_active,
                double* F, orc_force_out* out);

/* compute_mean_force (RG:769-803): mean |F_i| over |F_i|^2 > 1e-16,
 * divided by sqrt(ndim).  NaN when no particle qualifies. */
double orc_mean_force(const double* F, size_t N, size_t ndim);

/* Pair shell half-width t_shell (RG:1215-1223, RG:2846). */
double orc_shell(const double* D, size_t N);

/* Full-refresh neighbour set of the default backend: the kd-tree predicate
 * for ndim <= 3 (RG:3276-3297), the sort+walk predicate for ndim > 3
 * (RG:1305-1375).  Brute force over all pairs.  Returns the number of
 * pairs; fills row_off[N+1] and, if nbr != NULL, nbr[] (capacity cap).
 * Rows are ascending and hold partners j > i only. */
size_t orc_pairs_full(size_t N, size_t ndim, const double* x, const double* D,
                      const double* box, uint64_t* row_off, uint32_t* nbr,
                      size_t cap);

/* Spatial-reorder permutation (RG:1084-1160): Morton-21 for ndim == 3,
 * x-column sort otherwise; perm[k] = old index placed at slot k. */
void orc_reorder_perm(const double* x, size_t N, size_t ndim,
                      const double* box, uint32_t* perm);

/* One fused per-DOF ADAM + position + wrap pass (RG:6499-6520). */
void orc_adam_positions(size_t N, size_t ndim, const double* F, double* m,
                        double* v, double* x, const double* box, double alpha,
                        double inv_b1, double inv_b2);

/* Step cap (RG:4713-4737) for the default speed preset. */
size_t orc_step_cap(size_t N, size_t ndim, size_t max_steps);

/* Full run_packing_observed restatement (default environment). */
typedef struct {
    size_t N, ndim;
    const double* x;         /* N*ndim */
    const double* D;         /* N */
    const int8_t* fixed;     /* N or NULL */
    const double* box;       /* nbox entries (0 = default ones) */
    size_t nbox;
    const int8_t* walls;     /* nwalls entries (0 = all periodic) */
    size_t nwalls;
    int fix_height;
    size_t max_steps;        /* 0 = schedule-derived cap */
    int has_mu;              double mu;
    int fix_diameter;
    int has_target_phi;      double target_phi;
    size_t reorder_interval; /* reference default 2000 */
    int has_initial_phi;     double initial_phi;
} orc_run_in;

typedef struct {
    /* caller-allocated: x[N*ndim], D[N], origin[N], box[ndim], walls[ndim],
     * and seven histories of hist_cap entries each (hist_cap >= step cap) */
    double* x;
    double* D;
    uint64_t* origin;
    double* box;
    int8_t* walls;
    size_t hist_cap;
    double *phi_hist, *F_hist, *U_hist, *maxov_hist, *mu_hist, *alpha_hist;
    int8_t* muflag_hist;
    /* scalars */
    size_t steps;
    double phi, max_min_dist, force_magnitude;
    double kappa;
    int error;               /* 0 ok, 1 box length, 2 walls length */
} orc_run_out;

int orc_run_packing(const orc_run_in* in, orc_run_out* out);

#ifdef __cplusplus
}
#endif
#endif

/*
 * bfm.h -- C ABI of the MI355X mechanical step for spherocylinder biofilm ABMs.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, never throws,
 * and returns an int status (BFM_OK = 0). bfm_last_error(ctx) gives the message
 * of the last failure on that context.
 *
 * Array conventions (identical to the reference's Eigen arrays):
 *   cells      column-major n x ncols doubles with leading dimension ld >= n;
 *              element (i, c) is cells[c * ld + i]. Column indices follow
 *              ABM-3d/ABM-3d-cpp/include/indices.hpp:17-47 (id 0, r 1-3, n 4-6,
 *              dr 7-9, dn 10-12, l 13, half_l 14, t0 15, growth 16, eta0 17,
 *              eta1 18, sigma0 19, group 20) plus the optional gamma column
 *              (params.colidx_gamma, 21 in the reference: simulation.hpp:1182).
 *   neighbors  column-major m x 7 doubles with leading dimension ld_nb >= m:
 *              (i, j, dx, dy, dz, s_i, s_j), rows sorted (i asc, j asc), j < i,
 *              exactly as getCellNeighbors writes them (mechanics.hpp:536-586).
 *   forces / velocities / noise / errors
 *              column-major n x 6 doubles with leading dimension n.
 *
 * Entry points named bfm_* take HOST pointers and return once the results are
 * in host memory (the reference's synchronous call semantics). Entry points
 * named bfm_dev_* take DEVICE pointers, are enqueued on the context's stream
 * and return after enqueue; bfm_sync waits for them.
 *
 * Threading: one context per host thread. Calls on one context are
 * stream-ordered.
 */
#ifndef BFM_H
#define BFM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum bfm_status {
    BFM_OK = 0,
    BFM_ERR_ARG = 1,          /* invalid argument */
    BFM_ERR_CAPACITY = 2,     /* output or internal capacity exceeded */
    BFM_ERR_NAN = 3,          /* NaN / Inf detected in forces or velocities */
    BFM_ERR_JKR_ROOT = 4,     /* JKR contact-radius quartic has no admissible root */
    BFM_ERR_HIP = 5,          /* HIP runtime error */
    BFM_ERR_UNSUPPORTED = 6   /* mode not implemented behind this ABI */
};

/* Adhesion modes: AdhesionMode, mechanics.hpp:55-60. JKR_ANISOTROPIC takes
 * the JKR force from the precomputed tables of bfm_set_aniso_jkr_tables (the
 * reference's jkr_data path); jkr_mode does not apply to it. */
enum bfm_adhesion_mode {
    BFM_ADHESION_NONE = 0,
    BFM_ADHESION_JKR_ISOTROPIC = 1,
    BFM_ADHESION_JKR_ANISOTROPIC = 2
};

/* How the isotropic JKR contact radius is obtained for a contact pair. */
enum bfm_jkr_mode {
    BFM_JKR_EXACT = 0,      /* per-pair FP64 solve of the quartic (adhesion.hpp:171-240) */
    BFM_JKR_TABLE_1D = 1,   /* overlap table, fixed gamma (adhesion.hpp:282-370) */
    BFM_JKR_TABLE_2D = 2,   /* (overlap, gamma) table (adhesion.hpp:412-526) */
    BFM_JKR_SURROGATE = 3   /* BF16 MLP guess + FP64 polish, exact fallback outside envelope */
};

/* Physical and numerical parameters of getVelocities / stepRungeKuttaAdaptive
 * (mechanics.hpp:2011-2028, 2406-2427). */
typedef struct bfm_params {
    double R;                          /* cell radius incl. EPS */
    double Rcell;                      /* cell radius excl. EPS */
    double E0;                         /* EPS modulus */
    double repulsion_prefactors[3];    /* simulation.hpp:1271-1274 */
    double nz_threshold;
    double cell_cell_coulomb_coeff;    /* Coulomb cap of cell-cell friction (<= 0: none) */
    double cell_surface_coulomb_coeff;
    int no_surface;
    int adhesion_mode;                 /* enum bfm_adhesion_mode */
    int jkr_mode;                      /* enum bfm_jkr_mode */
    int colidx_gamma;                  /* gamma column, -1 when adhesion is off */
    int friction_mode;                 /* FrictionMode (mechanics.hpp:65-69): 0 NONE, 1 KINETIC
                                          (cellCellFrictionForces, mechanics.hpp:1031-1175) */
    int ncols;                         /* number of columns of the cells array */
    int colidx_eta_cell_cell;          /* friction coefficient column (friction on: 22 with
                                          adhesion, else 21; simulation.hpp:1185-1188) */
} bfm_params;

typedef struct bfm_ctx bfm_ctx;

/* ---- context ---------------------------------------------------------- */
int bfm_create(bfm_ctx** ctx, int device, int max_cells, int64_t max_pairs);
int bfm_destroy(bfm_ctx* ctx);
const char* bfm_last_error(bfm_ctx* ctx);
int bfm_set_params(bfm_ctx* ctx, const bfm_params* params);
/* Use an existing HIP stream (hipStream_t passed as void*); NULL = own stream. */
int bfm_set_stream(bfm_ctx* ctx, void* stream);
int bfm_sync(bfm_ctx* ctx);

/* JKR contact-radius table (IsotropicJKRDataFixedGamma, mechanics.hpp:113-153,
 * or IsotropicJKRDataVariableGamma, mechanics.hpp:155-197). n_gamma == 0 gives
 * the 1-D table; radii is n_overlap (1-D) or n_overlap x n_gamma row-major
 * radii[i * n_gamma + j] (2-D). */
int bfm_set_jkr_table(bfm_ctx* ctx, const double* overlaps, int n_overlap,
                      const double* gammas, int n_gamma, const double* radii);

/* Anisotropic JKR tables (AnisotropicJKRDataFixedGamma / VariableGamma,
 * mechanics.hpp:199-327; looked up by forcesAnisotropicJKRLagrange,
 * adhesion.hpp:734-1046): the principal radius of curvature Rx at a contact
 * point per (theta = angle between overlap vector and orientation, half_l,
 * normalized centerline coordinate), and the JKR force and contact radius
 * per (Rx1, Rx2, phi = angle between the cells, overlap[, gamma]). Grids are
 * non-decreasing, at most 8192 points together; n_gamma == 0 selects the
 * fixed-gamma (4-D) force table. Dense row-major tables:
 *   curvature_rx[(i_theta * n_half_l + i_half_l) * n_coords + i_coord]
 *   force_radius[2 * ((((i_rx1 * n_rx + i_rx2) * n_phi + i_phi) * n_delta
 *                + i_delta) * max(n_gamma, 1) + i_gamma) + {0: force, 1: radius}]
 * Combinations a precomputed file does not hold are 0 (what the reference's
 * hash-map lookup returns for them). The tables are copied to the device. */
typedef struct bfm_aniso_jkr_tables {
    const double* theta;  int n_theta;
    const double* half_l; int n_half_l;
    const double* coords; int n_coords;
    const double* curvature_rx;
    const double* rx;     int n_rx;
    const double* phi;    int n_phi;
    const double* delta;  int n_delta;
    const double* gamma;  int n_gamma;
    const double* force_radius;
} bfm_aniso_jkr_tables;
int bfm_set_aniso_jkr_tables(bfm_ctx* ctx, const bfm_aniso_jkr_tables* tables);

/* JKR surrogate network: 4 -> 64 -> 64 -> 64 -> 1 with SiLU, weights as BF16
 * bit patterns (uint16) row-major [out][in], biases FP32. envelope =
 * {delta_lo, delta_hi, gamma_lo, gamma_hi} in absolute units (um, Pa um). */
int bfm_set_surrogate(bfm_ctx* ctx, const uint16_t* w1, const float* b1,
                      const uint16_t* w2, const float* b2,
                      const uint16_t* w3, const float* b3,
                      const uint16_t* w4, const float* b4,
                      const double* envelope);

/* Surrogate training and evaluation (no counterpart in the reference: the
 * learned JKR radius is an addition of this implementation). Parameters are
 * FP32 masters in the order W1[64][4] b1[64] W2[64][64] b2[64] W3[64][64]
 * b3[64] W4[64] b4[1] (8705 floats). */
int bfm_surrogate_set_polish(bfm_ctx* ctx, int polish_steps);
/* Trains on n_samples seeded samples of the exact FP64 solve (overlap
 * U[-0.2 R, 0.3 R], gamma U[10, 1000] Pa um, samples without a JKR root
 * excluded), Adam with cosine-decayed lr; installs the weights. */
int bfm_surrogate_train(bfm_ctx* ctx, int n_samples, int epochs, int batch, double lr,
                        uint64_t seed, float* params_out, float* loss_per_epoch_out);
/* Network output t (= sqrt(a) / p^(1/3)) and the radius after `polish` FP64
 * Newton steps, for (delta, gamma) samples. */
int bfm_surrogate_eval(bfm_ctx* ctx, const double* delta, const double* gamma, int n,
                       int polish, float* t_out, double* radius_out);
/* Loss, outputs and parameter gradient of one forward/backward pass on the GPU
 * (features X[B][4], targets T[B]); for checking the training kernels. */
int bfm_surrogate_grad(bfm_ctx* ctx, const float* params, const float* X, const float* T,
                       int B, float* loss_out, float* y_out, float* grad_out);
/* The seeded training samples (features X[n][4], targets T[n], valid flags). */
int bfm_surrogate_generate(bfm_ctx* ctx, int n, uint64_t seed, float* X, float* T, int* ok);

/* ---- neighbor search -------------------------------------------------- */
/* getCellNeighbors (mechanics.hpp:520-593). Writes m rows into neighbors_out
 * (leading dimension ld_out). If m > ld_out, returns BFM_ERR_CAPACITY and sets
 * *m_out to the required m. */
int bfm_get_cell_neighbors(bfm_ctx* ctx, const double* cells, int n, int ld,
                           double neighbor_threshold, double R, double Ldiv,
                           double* neighbors_out, int64_t ld_out, int64_t* m_out);

/* updateNeighborDistances (mechanics.hpp:604-648): recomputes columns 2-6. */
int bfm_update_neighbor_distances(bfm_ctx* ctx, const double* cells, int n, int ld,
                                  double* neighbors, int64_t m, int64_t ld_nb);

/* ---- forces and velocities -------------------------------------------- */
/* getConservativeForces (mechanics.hpp:1702-1769), include_constraint = false. */
int bfm_get_conservative_forces(bfm_ctx* ctx, const double* cells, int n, int ld,
                                const double* neighbors, int64_t m, int64_t ld_nb,
                                const int* assume_2d, double* forces_out);

/* getVelocities (mechanics.hpp:2011-2347). noise may be NULL (zero noise). */
int bfm_get_velocities(bfm_ctx* ctx, const double* cells, int n, int ld,
                       const double* neighbors, int64_t m, int64_t ld_nb,
                       const int* assume_2d, const double* noise, double* vel_out);

/* stepRungeKuttaAdaptive (mechanics.hpp:2406-2612). A is s x s row-major,
 * b and bs have s entries. cells_new has the layout of cells with leading
 * dimension n. noise (n x 6) may be NULL; the caller draws it from its own RNG
 * stream, once per call, as the reference does (mechanics.hpp:2492-2510). */
int bfm_step_rk(bfm_ctx* ctx, const double* A, const double* b, const double* bs,
                int s, double dt, const double* cells, int n, int ld,
                const double* neighbors, int64_t m, int64_t ld_nb,
                const double* noise, double* cells_new, double* errors);

/* stepVelocityVerlet (mechanics.hpp:2668-2996): one velocity Verlet step of
 * size dt for cells of density rho (runSimulation sets rho = M0 / volume of a
 * cell of length L0, simulation.hpp:1280-1290), with the include_constraint
 * forms of the forces, ambient viscosity and cell-surface friction; noise
 * (n x 6, may be NULL) is added to both force evaluations as the reference
 * draws it once per call. cells_new has leading dimension n. */
int bfm_step_verlet(bfm_ctx* ctx, double dt, double rho, const double* cells, int n, int ld,
                    const double* neighbors, int64_t m, int64_t ld_nb, const double* noise,
                    double* cells_new);

/* ---- device-resident pipeline (device pointers, stream-ordered) -------- */
/* Builds the context's internal neighbor list from device cells. *m_out (host
 * pointer, may be NULL) receives m after a stream synchronisation; pass NULL
 * to stay asynchronous. */
int bfm_dev_build_neighbors(bfm_ctx* ctx, const double* d_cells, int n, int ld,
                            double neighbor_threshold, double R, double Ldiv,
                            int64_t* m_out);

/* updateNeighborDistances (mechanics.hpp:604-648) on the internal list. */
int bfm_dev_update_neighbor_distances(bfm_ctx* ctx, const double* d_cells, int n, int ld);

/* Copies the internal neighbor list into a device m x 7 column-major array. */
int bfm_dev_export_neighbors(bfm_ctx* ctx, double* d_neighbors, int64_t ld_out);

/* Loads an m x 7 device neighbor array as the internal list. */
int bfm_dev_import_neighbors(bfm_ctx* ctx, const double* d_neighbors, int64_t m,
                             int64_t ld_nb, int n);

/* getVelocities on device data with the internal neighbor list. assume_2d is
 * derived from d_cells when d_assume_2d is NULL. d_noise may be NULL. */
int bfm_dev_get_velocities(bfm_ctx* ctx, const double* d_cells, int n, int ld,
                           const int* d_assume_2d, const double* d_noise,
                           double* d_vel);

/* getVelocities on device arrays without any device allocation and without
 * a context. No counterpart in the reference: this is for callers that cannot
 * allocate device memory (code under HIP stream capture, serving loops,
 * harnesses that forbid hipMalloc). Same arguments and results as
 * bfm_dev_import_neighbors + bfm_dev_get_velocities (bit-identical
 * velocities): d_neighbors is a device m x 7 column-major array (leading
 * dimension ld_nb) in the reference's row order; d_assume_2d and d_noise may
 * be NULL; stream may be NULL (the default stream); runs on the current
 * device. Limits: n <= 65536 cells, m <= 262144 rows (BFM_ERR_CAPACITY
 * beyond); adhesion NONE or isotropic JKR with the exact solve, no friction
 * (BFM_ERR_UNSUPPORTED otherwise). Its working memory is a fixed ~21 MB
 * device array inside the library, one per GPU, held until the process ends
 * and shared by every call on that GPU: calls that overlap on one GPU (other
 * host threads, other streams) race. Outside a stream capture the call waits
 * for its kernels and returns BFM_ERR_ARG for a row index out of range,
 * BFM_ERR_JKR_ROOT or BFM_ERR_NAN; under capture it only enqueues. */
int bfm_dev_get_velocities_noalloc(const bfm_params* params, const double* d_cells, int n, int ld,
                                   const double* d_neighbors, int64_t m, int64_t ld_nb,
                                   const int* d_assume_2d, const double* d_noise, double* d_vel,
                                   void* stream);

/* One RK stage of a snapshot with its neighbor list rebuilt: binning, sort,
 * neighbor list, Hertz+JKR force and torque, surface forces and velocities,
 * all on device. With an isotropic interaction (Hertz, exact or tabulated
 * JKR), no friction and the default candidate kernel, the velocities come
 * straight from the per-cell neighbor lists of the search, in the
 * reference's row order for each cell, and the context keeps no reusable
 * list (build one with bfm_dev_build_neighbors); otherwise the list is
 * compacted into rows and kept. */
int bfm_dev_stage(bfm_ctx* ctx, const double* d_cells, int n, int ld,
                  double neighbor_threshold, double R, double Ldiv, double* d_vel);

/* stepRungeKuttaAdaptive on device data with the internal neighbor list.
 * d_cells_new has the layout of d_cells (leading dimension ld); d_errors
 * (n x 6, leading dimension n) may be NULL. d_err_scaled receives max over entries of errors / (tol * (1 + |y|)) (the
 * reference controller's quantity, simulation.hpp:1688-1695) when tol > 0. */
int bfm_dev_step_rk(bfm_ctx* ctx, const double* A, const double* b, const double* bs,
                    int s, double dt, const double* d_cells, int n, int ld,
                    const double* d_noise, double* d_cells_new, double* d_errors,
                    double tol, double* d_err_scaled);

/* stepVelocityVerlet on device arrays with the internal list, whose distances
 * are updated to the new positions (the reference's neighbors_new). */
int bfm_dev_step_verlet(bfm_ctx* ctx, double dt, double rho, const double* d_cells, int n, int ld,
                        const double* d_noise, double* d_cells_new, int ld_new);

/* ---- simulation driver (runSimulation, simulation.hpp:1106-2081) -------- */
/* Parameters of the growth/division loop; the mechanics parameters come from
 * bfm_set_params. tableau: 0 Heun-Euler, 1 Bogacki-Shampine, 2 RKF45,
 * 3 Dormand-Prince (simulation.hpp:1207-1268), or -1 for velocity Verlet
 * (stepVelocityVerlet at the fixed dt = max_stepsize, simulation.hpp:1787-1798;
 * one colony only, rho > 0). The random stream is a 32-bit
 * Mersenne twister seeded with `seed`; uniform draws are x * 2^-32 and normal
 * draws use Box-Muller as utils.hpp:648-655; draws follow the reference's
 * order (divideCells, growth.hpp:456-720). Noise is not supported (zero). */
typedef struct bfm_sim_config {
    double R, Rcell, Ldiv;
    double max_stepsize, min_stepsize, max_error_allowed, min_error;
    double neighbor_threshold;
    double growth_mean[2], growth_std[2];    /* per group (groups 1 and 2) */
    double daughter_length_std, daughter_angle_xy_bound, daughter_angle_z_bound;
    double max_time;
    int iter_update_neighbors, iter_update_stepsize, max_tries_update_stepsize;
    int n_cells, max_iter;
    int tableau;
    uint64_t seed;
    double report_seconds;     /* > 0: a progress line on stderr at this wall interval
                                  (the reference prints per dt_write, simulation.hpp:2026-2046) */
    double max_wall_seconds;   /* > 0: also stop after this much wall time */
    double rho;                /* cell density for tableau -1 (velocity Verlet): M0 / volume of
                                  a cell of length L0 (simulation.hpp:1280-1290) */
} bfm_sim_config;

typedef struct bfm_sim_stats {
    int64_t iterations, rk_steps, divisions, neighbor_builds;
    double t_end, wall_seconds, host_division_seconds;
} bfm_sim_stats;

/* Runs the loop from cells_in (n_in x ncols, leading dimension n_in) until
 * the reference's terminate rule holds: n >= n_cells, iter >= max_iter or
 * t >= max_time, over the criteria that are > 0. Cells and neighbor list stay
 * on the device; divisions are decided on the host with the reference's
 * random draws. parents_in (NULL: one -1 entry per id up to the largest id in
 * cells_in) is the lineage vector of runSimulation; parents_out must hold
 * n_parents_in + 2 (cap - n_in) ints. Writes the final cells (leading
 * dimension cap) and the accepted dt of each iteration (up to cap_trace).
 * BFM_ERR_CAPACITY when a division would exceed cap (cap <= max_cells). */
int bfm_simulate(bfm_ctx* ctx, const bfm_sim_config* cfg, const double* cells_in, int n_in,
                 int ncols, const int* parents_in, int n_parents_in, double* cells_out, int cap,
                 int* n_out, int* parents_out, int* n_parents_out, double* dt_trace,
                 int cap_trace, bfm_sim_stats* stats);

/* divideMaxLength + divideCells (growth.hpp:68-73, 456-720) on host data:
 * cells is n x ncols with leading dimension cap, daughters are appended. The
 * random stream starts fresh from seed. tries_out (may be NULL) receives the
 * number of placement attempts. Needs no device. */
int bfm_divide_cells(const bfm_sim_config* cfg, double* cells, int* n, int cap, int ncols,
                     double t, uint64_t seed, int* parents, int* n_parents, int parents_cap,
                     int* tries_out);

/* The first `count` uniform_01 draws of the random stream seeded with seed. */
int bfm_rng_uniform(uint64_t seed, int count, double* out);

/* The same stream drawn by the device generator of the growth loop's
 * divisions (one wavefront on GPU `device`): out[0 .. count) the first count
 * uniform_01 draws, out[count .. 2 count) the next count standard normals
 * (Box-Muller, two draws each). No counterpart in the reference: a check of
 * the device stream against the host one for callers and tests. */
int bfm_rng_draws_device(int device, uint64_t seed, int count, double* out);

/* ---- JKR setup (host only, no device) ------------------------------------ */
/* jkrEquilibriumDistance (jkr.hpp:1310-1340): the centerline distance 2R - delta
 * at which the JKR force of two cells of radius R (equivalent radius R / 2)
 * vanishes. newtonRaphson (rootFinding.hpp:42-75) on delta from 2R - d_init,
 * central differences of step increment_overlap, stop at |force| <= newton_tol
 * or max_iter steps; delta capped at 2R - d_min. residual_out (may be NULL)
 * receives the force at the last iterate. BFM_ERR_JKR_ROOT when an iterate
 * falls below the pull-off overlap (the reference throws there, jkr.hpp:99-105),
 * BFM_ERR_NAN when the iteration diverges. */
int bfm_jkr_equilibrium_distance(double R, double E0, double gamma, double d_init, double d_min,
                                 double increment_overlap, double newton_tol, int max_iter,
                                 double* deq_out, double* residual_out);

/* jkrOptimalSurfaceEnergyDensity (jkr.hpp:1369-1408): the surface energy density
 * gamma whose equilibrium distance (inner solves from 0.75 * 2R, capped at
 * 2 Rcell) is deq_target, by newtonRaphson on log10(gamma) from
 * log10(init_gamma). runSimulation calls it with init_gamma 100,
 * log_increment_gamma 1e-6, increment_overlap 1e-8, tolerances 1e-8 and 1000
 * iterations each (simulation.hpp:1354-1358). residual_out (may be NULL)
 * receives distance minus target at the last iterate. */
int bfm_jkr_optimal_surface_energy_density(double R, double Rcell, double E0, double deq_target,
                                           double init_gamma, double log_increment_gamma,
                                           double increment_overlap, double outer_newton_tol,
                                           double inner_newton_tol, int outer_max_iter,
                                           int inner_max_iter, double* gamma_out,
                                           double* residual_out);

/* ---- batched replicates (no counterpart in the reference: many independent
 * runSimulation loops advanced together) ---------------------------------- */
typedef struct bfm_ensemble_stats {
    int64_t iterations;          /* lockstep iterations */
    int64_t attempts;            /* replays of the captured RK attempt (all replicates) */
    int64_t divisions, neighbor_builds;
    int64_t stage_cell_updates;  /* sum over attempts of (cells of the attempting replicates) x stages */
    double wall_seconds, host_division_seconds, autotune_ms;
    double tuned_bin_scale;
    int tuned_variant;
    int n_final;
    int failed;                  /* replicates stopped by non-finite results */
    int pad;
} bfm_ensemble_stats;

/* Runs n_rep independent replicates of the bfm_simulate loop (same config and
 * parameters; replicate r has its own cells, random stream seeds[r], dt,
 * controller and lineage) until each meets the terminate rule. cells_in holds
 * the replicates back to back (sum(sizes) rows, leading dimension
 * sum(sizes)); cells_out receives them back to back (leading dimension cap),
 * sizes_out the final sizes, t_out / iters_out / failed_out (may be NULL) each
 * replicate's time, iteration count and failure flag. A replicate whose step
 * produces non-finite positions or orientations is frozen at the state its
 * iteration started from and flagged; the others continue. Every RK attempt of all replicates is one
 * replay of a captured graph (one launch per stage covering all replicates);
 * autotune != 0 first times the neighbor build over bin scales and candidate
 * kernel variants (bfm_autotune_neighbors) and keeps the fastest. */
int bfm_ensemble_simulate(bfm_ctx* ctx, const bfm_sim_config* cfg, int n_rep, const int* sizes,
                          const double* cells_in, int ncols, const uint64_t* seeds, int autotune,
                          double* cells_out, int cap, int* sizes_out, double* t_out, int64_t* iters_out,
                          int* failed_out, bfm_ensemble_stats* stats);

/* Times the device neighbor build on these cells for bin scales {0.5, 0.75,
 * 1, 1.25} x candidate-kernel variants 0..8 and keeps the fastest (every
 * choice gives the same list). Rebuilds the list with the choice. */
int bfm_autotune_neighbors(bfm_ctx* ctx, const double* d_cells, int n, int ld, double neighbor_threshold,
                           double R, double Ldiv, double* best_scale, int* best_variant, double* best_ms);

/* ---- instrumentation -------------------------------------------------- */
/* When enabled, every kernel launch is bracketed by HIP events on the launch
 * stream; bfm_kernel_times returns per-kernel accumulated milliseconds and
 * launch counts (names separated by '\n' in names_out). */
int bfm_enable_timing(bfm_ctx* ctx, int enable);
int bfm_kernel_times(bfm_ctx* ctx, char* names_out, int names_cap, double* ms_out,
                     int64_t* counts_out, int cap, int* n_kernels);
int bfm_reset_timing(bfm_ctx* ctx);

/* Number of pairs in the internal list; candidate pairs screened (center
 * distance bound) and segment-segment tests run by the neighbor builds since
 * the previous call (the counters are reset); the last build's bin geometry. */
int bfm_stats(bfm_ctx* ctx, int64_t* m, int64_t* screened, int64_t* tested, double* bin_edge,
              int* nbins);

/* Tuning knobs of the neighbor build: bin edge as a multiple of the search
 * cutoff (1/3 .. 4; 1 gives the 27-bin stencil, 0.5 the 125-bin stencil),
 * threads per workgroup (reserved, 0), and the candidate-kernel configuration
 * 0..8 (0..3 and 6: thread-per-candidate screen with (homes, LDS chunk) =
 * (32, 640), (64, 640), (64, 1024), (64, 512), (128, 512);
 * 4, 5: wave-per-candidate screen with LDS chunk 512 or 1024; 7: one
 * wavefront per home bin, lane per candidate, no work-group barriers; needs
 * max_cells < 2^25; 8: lane per home walking its own half stencil, a second
 * point-segment screen before the FP64 test, no work-group barriers; needs
 * max_cells <= 2^26; the default, 6 for larger contexts). 0 / 0 / -1 keep
 * the current values.
 * No knob changes the neighbor list, only the time to build it. */
int bfm_set_tuning(bfm_ctx* ctx, double bin_scale, int block_threads, int kernel_variant);

/* ---- the 2-D model (ABM-2d/ABM-2d-cpp) ------------------------------------
 * Cells: the 2-D tree's column-major layout (include/indices.hpp of ABM-2d):
 * id 0, r 1-2, n 3-4, dr 5-6, dn 7-8, l 9, half_l 10, t0 11, growth 12,
 * eta0 13, eta1 14, maxeta1 15, group 16 (ncols >= 17). Neighbor lists are
 * m x 6: i, j, dx, dy, s_i, s_j (getCellNeighbors, mechanics.hpp:293-374 of
 * ABM-2d). to_adhere: one int per neighbor row (the driver's to_adhere). */
enum bfm2d_adhesion_mode {
    BFM2D_ADHESION_NONE = 0,
    BFM2D_ADHESION_JKR = 1,      /* forcesJKRLagrange<T, 2> (adhesion.hpp:402-462 of ABM-2d) */
    BFM2D_ADHESION_KIHARA = 2,   /* forcesKiharaLagrange<T, 2> (adhesion.hpp:334-400 of ABM-2d) */
    /* GBK (3) is not built */
    BFM2D_ADHESION_JKR_ISOTROPIC = 4   /* the 3-D tree's dimension-templated
        forcesIsotropicJKRLagrange<T, 2> (adhesion.hpp:171-233 of ABM-3d) in the
        2-D adhesive loop: contact radius of the JKR quartic (exact FP64 solve)
        at the overlap capped at 2 (R - Rcell), as the 3-D pair loop calls it
        (mechanics.hpp:1482-1487 of ABM-3d), E0 and gamma below, strength 1
        for the reference's scale. The 2-D tree has no gamma-parametrised law;
        this composition gives 2-D growth runs a surface energy. */
};
typedef struct bfm2d_params {
    double R, Rcell;
    double prefactors[3];            /* cell_cell_prefactors of cellCellRepulsiveForces
                                        (the driver forms them at simulation.hpp:231-234) */
    double surface_contact_density;  /* (sigma0 R^2 / (4 E0))^(1/3), simulation.hpp:237 */
    int adhesion_mode;               /* enum bfm2d_adhesion_mode */
    int ncols;                       /* columns of the cells array (>= 17) */
    double strength, mindist, distance_exp;   /* adhesion_params "strength",
                                        "mindist", "distance_exp" */
    double E0, gamma;                /* BFM2D_ADHESION_JKR_ISOTROPIC: EPS modulus and
                                        surface energy density (Pa um) */
} bfm2d_params;

/* The 2-D mechanics use the context's device buffers; bfm2d_set_params may
 * be called on a context with or without bfm_set_params. */
int bfm2d_set_params(bfm_ctx* ctx, const bfm2d_params* params);

/* getCellNeighbors of the 2-D tree (mechanics.hpp:293-374 of ABM-2d): pairs
 * j < i whose segment distance is below thr, rows (i asc, j asc). The
 * reference takes the distance of non-parallel cells from CGAL's segment
 * distance; here it is the closest-point computation of the 3-D path on the
 * cells placed in the z = 0 plane (the parallel branch is the reference's). */
int bfm2d_get_cell_neighbors(bfm_ctx* ctx, const double* cells, int n, int ld, double thr, double R,
                             double Ldiv, double* neighbors_out, int64_t ld_out, int64_t* m_out);

/* getVelocities of the 2-D tree (mechanics.hpp:1503-1708 of ABM-2d) without
 * confinement (ConfinementMode::NONE): repulsion + adhesion + noise (n x 4,
 * may be NULL) divided by the composite drag prefactors. velocities_out: n x 4. */
int bfm2d_get_velocities(bfm_ctx* ctx, const double* cells, int n, int ld, const double* neighbors,
                         int64_t m, int64_t ld_nb, const int* to_adhere, const double* noise,
                         double* velocities_out);

/* stepRungeKuttaAdaptive of the 2-D tree (mechanics.hpp:1710-1817 of
 * ABM-2d): s stages with the step's neighbor rows, the stage orientations in
 * the constrained torques; cells_out gets the new coordinates, the stored
 * velocities and the other columns copied; errors_out (n x 4) the embedded
 * error. noise: n x 4 drawn by the caller once per call, or NULL. */
int bfm2d_step_rk(bfm_ctx* ctx, const double* A, const double* b, const double* bs, int s, double dt,
                  const double* cells, int n, int ld, const double* neighbors, int64_t m, int64_t ld_nb,
                  const int* to_adhere, const double* noise, double* cells_out, int ld_out,
                  double* errors_out);

/* Device-resident 2-D forms (device arrays, the context's stream, the
 * context's internal list). bfm2d_dev_build_neighbors builds the list of
 * bfm2d_get_cell_neighbors and the rows' to_adhere flags by the 2-D driver's
 * rule (simulation.hpp:242-262 of ABM-2d): the groups (min, max) of the pair
 * adhere in adhesion_map (host, n_groups x n_groups row-major, group = column
 * 16 as an integer; NULL: no row adheres) and R + Rcell < |d| < 2R at build
 * time. m_out == NULL: no synchronisation. The list exports with
 * bfm_dev_export_neighbors (7 columns, dz = 0). */
int bfm2d_dev_build_neighbors(bfm_ctx* ctx, const double* d_cells, int n, int ld, double thr, double R,
                              double Ldiv, const int* adhesion_map, int n_groups, int64_t* m_out);
/* updateNeighborDistances of the 2-D tree (mechanics.hpp:376-446 of ABM-2d). */
int bfm2d_dev_update_neighbor_distances(bfm_ctx* ctx, const double* d_cells, int n, int ld);
/* bfm2d_step_rk on device arrays with the internal list (d_noise n x 4 or
 * NULL, d_errors n x 4 or NULL). */
int bfm2d_dev_step_rk(bfm_ctx* ctx, const double* A, const double* b, const double* bs, int s, double dt,
                      const double* d_cells, int n, int ld, const double* d_noise, double* d_cells_out,
                      int ld_out, double* d_errors);

/* The growth loop of the 2-D tree, runSimulationAdaptiveLagrangian
 * (simulation.hpp:549-765 of ABM-2d), on the device 2-D step with the
 * context's bfm2d_params: per iteration divide the cells with l > Ldiv
 * (divideCells, growth.hpp:438-697 of ABM-2d, on the host with the
 * reference's random draws in its order), rebuild the list and to_adhere
 * after a division, take an adaptive RK step with the controller
 * (simulation.hpp:636-712), grow (growCells, growth.hpp:51-57), refresh the
 * distances, advance t and iter, rebuild every iter_update_neighbors.
 * cfg fields read: R, Rcell, Ldiv, max/min_stepsize, max_error_allowed,
 * min_error, neighbor_threshold (the 2-D driver uses 2R + Ldiv), growth
 * mean/std of groups 1-2, daughter_length_std, daughter_angle_xy_bound (the
 * 2-D daughter_angle_bound), n_cells / max_iter / max_time, the
 * iter_update_* and max_tries counts, tableau 0..3 (the 2-D tree fixes
 * Bogacki-Shampine, 1), seed, report_seconds, max_wall_seconds. Not built:
 * noise (max_noise = 0), switching, confinement, growth voids, pole tracking.
 * adhesion_map as in bfm2d_dev_build_neighbors. Cells: ncols of the params,
 * leading dimension n_in in, cap out. parents_in NULL: one -1 entry per id
 * up to the largest id; parents_out must hold n_parents_in + 2 (cap - n_in)
 * ints. BFM_ERR_CAPACITY when a division would exceed cap (cap <= max_cells),
 * BFM_ERR_NAN when a stage velocity is not finite. */
int bfm2d_simulate(bfm_ctx* ctx, const bfm_sim_config* cfg, const int* adhesion_map, int n_groups,
                   const double* cells_in, int n_in, const int* parents_in, int n_parents_in,
                   double* cells_out, int cap, int* n_out, int* parents_out, int* n_parents_out,
                   double* dt_trace, int cap_trace, bfm_sim_stats* stats);

/* divideMaxLength + divideCells of the 2-D tree (growth.hpp:67-71, 438-697
 * of ABM-2d) on host data, the 2-D columns: cells n x ncols (ncols >= 17)
 * with leading dimension cap, daughters appended; the random stream starts
 * fresh from seed. Needs no device. */
int bfm2d_divide_cells(const bfm_sim_config* cfg, double* cells, int* n, int cap, int ncols, double t,
                       uint64_t seed, int* parents, int* n_parents, int parents_cap, int* tries_out);

/* ---- frames, lineage, configuration and the driver (host only) ---------- */
/* Message of the last failed call below (per thread). */
const char* bfm_io_last_error(void);

/* writeCells (utils.hpp:146-198, the std::map overload): "# key = value" lines
 * sorted by key (a repeated key keeps its last value), then one row per cell,
 * tab-separated, 10 significant digits; columns 0 (id) and 20 (group) as
 * integers, 1..19 as doubles, a column >= 21 only if listed in extra_cols
 * (extra_types 0: integer, else double). cells: n x ncols, leading dim ld. */
int bfm_write_cells(const char* path, const double* cells, int n, int ld, int ncols,
                    const char* const* keys, const char* const* values, int n_params,
                    const int* extra_cols, const int* extra_types, int n_extra);

/* readCells (utils.hpp:95-143): rows and "# key = value" lines. n_out and
 * ncols_out (at least 21, the widest row otherwise; short rows are padded with
 * 0) are always set; params_out (may be NULL) receives "key = value\n" lines
 * in file order, params_len_out their length. With cells == NULL only the
 * sizes are returned; else cells (leading dimension ld >= n, room for
 * max_cols columns and cap rows) receives the rows. Empty lines are skipped. */
int bfm_read_cells(const char* path, double* cells, int cap, int ld, int max_cols, int* n_out,
                   int* ncols_out, char* params_out, int params_cap, int* params_len_out);

/* writeLineage / readLineage (utils.hpp:207-258): "i<TAB>parents[i]" lines.
 * bfm_read_lineage returns parents[child] = parent for child ids 0..n-1
 * (the vector abmConstAdhesion.cpp:386-393 builds); parents == NULL: count only. */
int bfm_write_lineage(const char* path, const int* parents, int n);
int bfm_read_lineage(const char* path, int* parents, int cap, int* n_out);

/* The JSON configuration of the 3-D driver (abmConstAdhesion.cpp:31-333):
 * keys and value kinds as there (floating-point keys need a number with a
 * fraction or exponent, integer keys an integer literal), optional keys with
 * the driver's defaults; Ldiv = 2 L0 + 2 R, neighbor_threshold = R + Ldiv / 2. */
typedef struct bfm_run_config {
    double R, Rcell, L0, Ldiv, E0, Ecell, sigma0, eta_ambient, eta_surface;
    double max_stepsize, min_stepsize, dt_write, max_time, max_error_allowed;
    double growth_mean, growth_std, daughter_length_std, daughter_angle_xy_bound,
        daughter_angle_z_bound;
    double nz_threshold, max_rxy_noise, max_rz_noise, max_nxy_noise, max_nz_noise;
    double basal_min_overlap, adhesion_eqdist, jkr_imag_tol, jkr_aberth_tol;
    double eta_cell_cell, cell_cell_coulomb_coeff, cell_surface_coulomb_coeff, M0;
    double neighbor_threshold;
    int iter_update_stepsize, iter_update_neighbors, max_iter, n_cells;
    int basal_only, adhesion_mode, friction_mode, integration_mode, no_surface;
    int n_mesh_overlap, n_mesh_gamma;
    char init_filename[1024], lineage_filename[1024];
    /* adhesion_mode 2: adhesion_precompute_jkr_forces and, when it is 0, the
     * table files (abmConstAdhesion.cpp:133-138, 242-249) */
    int precompute_jkr_forces;
    char curvature_filename[1024], jkr_forces_filename[1024];
    /* adhesion_mode 2: the optional keys of the ellipsoid pipeline, with the
     * driver's defaults, as the frames record them (abmConstAdhesion.cpp:139-240) */
    double calibrate_endpoint_radii, min_aspect_ratio, ellipsoid_project_tol, brent_tol, init_bracket_dx;
    int n_mesh_theta, n_mesh_half_l, n_mesh_centerline_coords, n_mesh_curvature_radii,
        ellipsoid_project_max_iter, brent_max_iter, n_tries_bracket;
} bfm_run_config;
int bfm_load_config(const char* json_path, bfm_run_config* out);

/* parseCurvatureRadiiTable and parseReducedJKRForceTable / parseJKRForceTable
 * (simulation.hpp:565-1008): the two tab-separated files of the anisotropic
 * JKR tables read into the dense tables of bfm_set_aniso_jkr_tables.
 * variable_gamma = 0: the force file's first line is the fixed gamma
 * (returned in *fixed_gamma_out, may be NULL) and its rows are (Rx1, Rx2,
 * phi, overlap, force, radius); else no header and rows (Rx1, Rx2, phi,
 * overlap, gamma, force, radius). *out stays valid until
 * bfm_free_aniso_jkr_tables. */
int bfm_read_aniso_jkr_tables(const char* curvature_path, const char* forces_path, int variable_gamma,
                              bfm_aniso_jkr_tables** out, double* fixed_gamma_out);
void bfm_free_aniso_jkr_tables(bfm_aniso_jkr_tables* tables);

/* The driver main() of abmConstAdhesion.cpp:31-457 with runSimulation's output
 * (simulation.hpp:1444-1572, 2023-2079): reads the config, starts from the
 * founder cell or from init_filename + lineage_filename, sets the surface
 * energy density of group 1 from adhesion_eqdist, runs the growth loop on
 * `device` (rng_seed seeds the division stream) and writes
 * <outprefix>_init.txt, <outprefix>_iter<k>.txt whenever t passes a multiple
 * of dt_write, <outprefix>_final.txt and <outprefix>_lineage.txt. jkr_mode:
 * BFM_JKR_TABLE_1D (runSimulation's table of n_mesh_overlap exact radii) or
 * BFM_JKR_EXACT (adhesion_mode 1; adhesion_mode 2 reads the anisotropic
 * tables from adhesion_curvature_filename and adhesion_jkr_forces_filename).
 * max_cells: capacity (0: 2 max(n_cells, initial) + 16).
 * verbose prints the reference's per-frame progress line on stdout (and a wall-clock
 * progress line on stderr every 30 s). Modes the device loop lacks
 * (anisotropic tables computed at start-up, basal_only, noise) return
 * BFM_ERR_UNSUPPORTED. stats_out may be NULL. */
int bfm_run(const char* json_path, const char* outprefix, int rng_seed, int device, int jkr_mode,
            int max_cells, int verbose, bfm_sim_stats* stats_out);

#ifdef __cplusplus
}
#endif

#endif /* BFM_H */

/*
 * oracle.h -- CPU oracle for the mechanical step (TEST INFRASTRUCTURE ONLY).
 *
 * A from-scratch C++ restatement of the reference's hot-path algorithms, used
 * by tests/ as the checker of the HIP path, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg. Nothing in the product links or calls it.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * ABM-3d/ABM-3d-cpp/). Arrays use the layouts documented in include/bfm.h.
 *
 * Parity pins (tests/): the 16 distBetweenCells known answers
 * (test/src/testDistBetweenCells.cpp:143-506), brute-force minimisation of the
 * segment distance, JKR roots against mpmath 50-digit fixtures and the
 * reference's closed forms (test/src/testJKR.cpp, testContactArea.cpp:102-162),
 * surface integrals against quadrature, surface forces against finite
 * differences of the surface energy, the velocity solve against the residual
 * of the 7x7 system, the RK step against tableau order tests, and the
 * controller against a scripted trace.
 */
#ifndef BFM_ORACLE_H
#define BFM_ORACLE_H

#include <stdint.h>
#include "../include/bfm.h"

#ifdef __cplusplus
extern "C" {
#endif

double oracle_nearest_cell_body_coord(const double* r, const double* n, double half_l,
                                      const double* q);
void oracle_dist_between_cells(const double* r1, const double* n1, double half_l1,
                               const double* r2, const double* n2, double half_l2,
                               double* d, double* s, double* t);
void oracle_friction_forces(const double* cells, int n, int ld, const double* nb, int64_t m,
                            int64_t ld_nb, const bfm_params* p, double* dEdq);
void oracle_cell_cell_forces_ic(const double* cells, int n, int ld, const double* nb,
                                int64_t m, int64_t ld_nb, const bfm_params* p, double* dEdq);
void oracle_conservative_forces_ic(const double* cells, int n, int ld, const double* nb, int64_t m,
                                   int64_t ld_nb, const int* assume_2d, const bfm_params* p, double* forces);
void oracle_ambient_viscosity_forces(const double* cells, int n, int ld, double* dEdq);
void oracle_surface_friction_forces(const double* cells, int n, int ld, const double* ss, double R,
                                    const int* assume_2d, double* dEdq);
void oracle_step_verlet(double dt, double rho, const double* cells, int n, int ld, const double* nb, int64_t m,
                        int64_t ld_nb, const double* noise, const bfm_params* p, double* cells_new);
int64_t oracle_get_cell_neighbors(const double* cells, int n, int ld, double thr,
                                  double R, double Ldiv, double* out, int64_t ld_out);
int64_t oracle_get_cell_neighbors_binned(const double* cells, int n, int ld, double thr,
                                         double R, double Ldiv, double* out,
                                         int64_t ld_out, int nthreads);
int64_t oracle_count_neighbor_rows(const double* cells, int n, int ld, double thr, double R,
                                   double Ldiv, int row_lo, int row_hi);
void oracle_update_neighbor_distances(const double* cells, int n, int ld,
                                      double* neighbors, int64_t m, int64_t ld_nb);

/* JKR: larger real root of a^4 - 2 R d a^2 - (4 pi R^2 g / E) a + R^2 d^2. */
double oracle_jkr_contact_radius(double delta, double Req, double E, double gamma,
                                 int* iters);
double oracle_jkr_force_magnitude(double radius, double Req, double E0, double gamma);
double oracle_jkr_equilibrium_distance(double R, double E0, double gamma, double dinit,
                                       double dmin, double increment_overlap, double newton_tol,
                                       int max_iter, double* residual);
double oracle_jkr_optimal_surface_energy_density(double R, double Rcell, double E0,
                                                 double deq_target, double init_gamma,
                                                 double log_increment_gamma,
                                                 double increment_overlap, double outer_tol,
                                                 double inner_tol, int outer_max_iter,
                                                 int inner_max_iter, double* residual);
void oracle_linspace(double lo, double hi, int n, double* out);
int oracle_nearest_value(const double* values, int n, double x);
double oracle_jkr_table_radius_1d(const double* overlaps, const double* radii, int n,
                                  double delta);
double oracle_jkr_table_radius_2d(const double* overlaps, int nd, const double* gammas,
                                  int ng, const double* radii, double delta, double gamma);
void oracle_set_jkr_table(const double* overlaps, int n_overlap, const double* gammas,
                          int n_gamma, const double* radii);

/* Anisotropic JKR tables (dense, the layouts of bfm_aniso_jkr_tables) and the
 * table form of forcesAnisotropicJKRLagrange, adhesion.hpp:734-1046. */
void oracle_set_aniso_tables(const double* theta, int nt, const double* half_l, int nh,
                             const double* coords, int nc, const double* curv, const double* rx, int nr,
                             const double* phi, int np, const double* delta, int nd,
                             const double* gamma, int ng, const double* fr);
void oracle_aniso_jkr_force(const double* n1, double half_l1, const double* n2, double half_l2,
                            const double* d12, double R, double gamma, double s, double t,
                            double max_overlap, double* force_out, double* radius_out);

/* Surface integrals, integrals.hpp:35-463. */
double oracle_integral1(double rz, double nz, double R, double half_l, double g, double ss);
double oracle_integral2(double rz, double nz, double R, double half_l, double g, double ss);
double oracle_integral3(double rz, double nz, double R, double half_l, double g, double ss);
double oracle_integral4(double nz, double half_l, double ss);
double oracle_integral5(double nz, double half_l, double ss);
double oracle_integral6(double nz, double half_l, double ss);
void oracle_area_integrals(double rz, double nz, double R, double half_l, double ss,
                           double* out3);

void oracle_cell_cell_forces(const double* cells, int n, int ld, const double* neighbors,
                             int64_t m, int64_t ld_nb, const bfm_params* p, double* dEdq);
void oracle_surface_repulsion_forces(const double* cells, int n, int ld, const double* ss,
                                     double R, double E0, const int* assume_2d, double* dEdq);
void oracle_surface_adhesion_forces(const double* cells, int n, int ld, const double* ss,
                                    double R, const int* assume_2d, double* dEdq);
void oracle_conservative_forces(const double* cells, int n, int ld, const double* neighbors,
                                int64_t m, int64_t ld_nb, const int* assume_2d,
                                const bfm_params* p, double* forces);
void oracle_get_velocities(const double* cells, int n, int ld, const double* neighbors,
                           int64_t m, int64_t ld_nb, const int* assume_2d,
                           const double* noise, const bfm_params* p, double* vel);
void oracle_normalize_orientations(double* cells, int n, int ld);
void oracle_step_rk(const double* A, const double* b, const double* bs, int s, double dt,
                    const double* cells, int n, int ld, const double* neighbors, int64_t m,
                    int64_t ld_nb, const double* noise, const bfm_params* p,
                    double* cells_new, double* errors);
double oracle_max_scaled_error(const double* cells, int n, int ld, const double* errors,
                               double tol, double min_error);
/* One pass of the step-size controller (simulation.hpp:1688-1769) on a scripted
 * sequence of errors: err_seq[k] is the error the k-th attempted step returns.
 * Writes the dt of every attempted step (initial, retries, final redo) into
 * dts and returns their number. */
int oracle_controller_trace(double dt, double error_order, int max_tries,
                            double min_stepsize, double max_stepsize,
                            const double* err_seq, int n_err, double* dts);

/* divideMaxLength + divideCells (growth.hpp:68-73, 456-720) with a fresh
 * std::mt19937(seed): cells n x ncols, leading dimension cap, daughters
 * appended; brute-force minDistToCell. */
int oracle_divide_cells(const bfm_sim_config* cfg, double* cells, int* n, int cap, int ncols,
                        double t, uint64_t seed, int* parents, int* n_parents, int parents_cap,
                        int* tries_out);
/* runSimulation's loop (simulation.hpp:1616-1864) with zero noise: writes the
 * final cells (leading dimension cap), n, the accepted dt per iteration, the
 * iteration count and the number of RK step evaluations. */
int oracle_simulate(const bfm_sim_config* cfg, const bfm_params* p, const double* cells_in, int n_in,
                    double* cells_out, int cap, int* n_out, double* dt_trace, int cap_trace,
                    int64_t* iters_out, int64_t* rk_steps_out, int nthreads);


/* The 2-D model (ABM-2d/ABM-2d-cpp): neighbor rows (i, j, dx, dy, s, t),
 * repulsive / adhesive dE/dq, getVelocities and stepRungeKuttaAdaptive. */
int64_t oracle2d_get_cell_neighbors(const double* cells, int n, int ld, double thr, double R, double Ldiv,
                                    double* out, int64_t ld_out);
void oracle2d_forces(const double* cells, int n, int ld, const double* nb, int64_t m, int64_t ld_nb,
                     const int* to_adhere, const bfm2d_params* p, double* rep, double* adh);
void oracle2d_get_velocities(const double* cells, int n, int ld, const double* nb, int64_t m, int64_t ld_nb,
                             const int* to_adhere, const double* noise, const bfm2d_params* p, double* vel);
void oracle2d_step_rk(const double* A, const double* b, const double* bs, int s, double dt, const double* cells,
                      int n, int ld, int ncols, const double* nb, int64_t m, int64_t ld_nb, const int* to_adhere,
                      const double* noise, const bfm2d_params* p, double* cells_new, double* errors);
/* divideCells of the 2-D tree on host data (bfm2d_divide_cells' contract). */
int oracle2d_divide_cells(const bfm_sim_config* cfg, double* cells, int* n, int cap, int ncols, double t,
                          uint64_t seed, int* parents, int* n_parents, int parents_cap, int* tries_out);
/* The 2-D growth loop (bfm2d_simulate's contract; tableau 0..3, one thread). */
int oracle2d_simulate(const bfm_sim_config* cfg, const bfm2d_params* p, const int* amap, int n_groups,
                      const double* cells_in, int n_in, double* cells_out, int cap, int* n_out, int* parents_out,
                      int* n_parents_out, double* dt_trace, int cap_trace, int64_t* iters_out,
                      int64_t* rk_steps_out);

#ifdef __cplusplus
}
#endif

#endif

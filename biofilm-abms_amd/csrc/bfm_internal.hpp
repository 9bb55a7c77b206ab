// bfm_internal.hpp -- library-internal entry points shared by the simulation
// drivers (bfm_sim.hip) and the mechanical-step implementation (bfm_main.hip).
// Each returns a bfm_status and sets the context's error message, like the
// C ABI functions.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "bfm_ctrl.hpp"
#include "bfm_ctx.hpp"

namespace bfm {

// Frozen-geometry part of a step (mechanics.hpp:2483-2556): assume_2d from the
// step-start orientations, pair forces and the per-cell cell-cell forces.
// Every RK attempt of one iteration reuses them. refresh_distances: the
// list's distances are first brought to the packed segments that
// iteration_tail_batched(pack_segments) wrote. *cc_in_attempt is set when
// the cell-cell sums are left to the iteration's first attempt
// (attempt_direct with gather). d_gate non-null (an iteration enqueued ahead
// of the host's check, iteration_ahead_ok): nothing runs unless *d_gate.
int step_prepare(bfm_ctx* c, const double* d_cells, int n, int ld, bool refresh_distances, bool* cc_in_attempt,
                 const int* d_gate = nullptr);

// One RK attempt of a batch of replicates after step_prepare: the stages and
// the update, with the cell count read from d_n and each replicate's dt and
// active flag from d_dt / d_active (inactive rows are copied unchanged);
// d_errmax[r] receives replicate r's scaled-error maximum as an ordered key
// (atomic maxima: the caller resets it to the key of -inf before the
// attempt); d_bad[r] is or-ed with 1 when replicate r produced a non-finite
// position or orientation (the caller clears it). Enqueues only kernels, so
// a stream capture records it. With an isotropic interaction and no
// friction the attempt is one launch.
int step_attempt_batched(bfm_ctx* c, const Tableau& tb, const double* d_cells, int ld, const int* d_n,
                         const int* d_rep, const double* d_dt, const int* d_active, int n_rep,
                         double* d_cells_new, double tol, unsigned long long* d_errmax, int* d_bad);

// The end of a batched RK iteration, from the controller state E.S (the
// replicates that stepped and their dt): the controller's end-of-iteration
// step (E), the columns the attempts did not write copied from d_cells into
// d_cells_new, growCells on d_cells_new, the indices of the grown cells with
// l > Ldiv appended to d_idx through *d_count (zeroed by the controller's
// first step of the iteration), then the iteration's values for the host
// in *d_snap (IterSnap: the division count, the status flags, the totals,
// the per-replicate t, dt, err, flags after it) and in *d_spec whether the
// next iteration may run ahead (d_gate of its kernels). n (the host's
// count, equal to *d_n) sizes the grid; d_ticket: one zeroed unsigned of
// scratch. Nothing when *d_gate == 0 (d_gate non-null); when *d_skip != 0
// only d_snap->tail_ran = 0 and *d_spec = 0. pack_segments: the new state's
// packed segments also go to
// the context's segment buffer, from which step_prepare(refresh_distances)
// brings the list's distances to the new state.
int iteration_tail_batched(bfm_ctx* c, const double* d_cells, int ld, const int* d_n, int n, const int* d_rep,
                           const CtrlEnd& E, double R, double Ldiv, double* d_cells_new, int* d_count, int* d_idx,
                           const int* d_skip, bool pack_segments, const int* d_gate, IterSnap* d_snap,
                           int* d_spec, unsigned* d_ticket);

// Whether an iteration of n cells may be enqueued before the host has read
// the previous one (its kernels gated on the previous tail's *d_spec): the
// one-kernel attempts with their controller steps in the last work-group,
// a device-built list, the exact or tabulated JKR radius.
bool iteration_ahead_ok(const bfm_ctx* c, int n);

// One attempt as a direct launch (attempt_fused configurations) followed by
// the controller's step A (ctrl_after_block), the grid sized for n cells
// (= *d_n). The attempt's replicates and dt: A.active / A.trial, or, when
// A.first_go is set (the iteration's first attempt), those with
// A.first_go[r] && !A.S.failed[r] at A.S.dt[r]. gather: the cell-cell sums
// are gathered from the pair records and stored (step_prepare left them to
// the attempt). d_ticket: one zeroed unsigned of scratch.
int attempt_direct(bfm_ctx* c, const Tableau& tb, const double* d_cells, int ld, const int* d_n, int n,
                   const int* d_rep, bool gather, double* d_cells_new, double tol, const CtrlAfter& A,
                   unsigned* d_ticket);

// Whether a batched RK attempt is one launch that reads only the step's
// frozen cell-cell sums (an isotropic interaction without friction), not
// the neighbor list.
bool attempt_fused(const bfm_ctx* c);

// growCells with per-replicate dt and active flags, cell count from d_n.
int grow_batched(bfm_ctx* c, double* d_cells, int ld, const int* d_n, double R, const int* d_rep,
                 const double* d_dt, const int* d_active);

// Switches the context to n_rep batched replicates (0: a single colony);
// allocates the replicate arrays on first use. d_rep must then hold each
// cell's replicate before a neighbor build.
int set_replicate_count(bfm_ctx* c, int n_rep);

// stepVelocityVerlet on device data with the internal list (whose distances
// end at the new positions); d_bad[0] is set to 1 when a new position or
// orientation is not finite.
int step_verlet(bfm_ctx* c, double dt, double rho, const double* d_cells, int n, int ld, double* d_new, int ld_new,
                int* d_bad);

// Ordered-key <-> double for the scaled-error maxima.
double order_key_value(unsigned long long k);

// The 2-D step for the 2-D growth loop, on the internal list
// (bfm2d_dev_build_neighbors) and the context's bfm2d_params.
// step2d_prepare: the pair records of the list's rows and to_adhere flags,
// then the per-cell frozen-geometry force sums d_sums (8 n doubles).
// step2d_rows_update: updateNeighborDistances on the rows, their records and
// the sums. step2d_attempt: one RK attempt from the sums into d_new (one
// launch); tol > 0 also folds the controller's scaled-error maximum into
// d_errmax (ordered key, starting from -inf); d_count != NULL also grows the
// new cells by dt (growCells) and adds the number with l > Ldiv to *d_count.
// step2d_publish: writes {error key, count, status} to host-mapped memory and
// resets the key and the count. Non-finite stage velocities set ST_NAN in
// the context's status word. All only enqueue.
int step2d_prepare(bfm_ctx* c, int n, double* d_sums);
int step2d_rows_update(bfm_ctx* c, const double* d_cells, int n, int ld, double* d_sums);
int step2d_attempt(bfm_ctx* c, const Tableau& tb, double dt, const double* d_cells, int n, int ld,
                   const double* d_sums, double* d_new, int ld_new, double tol, unsigned long long* d_errmax,
                   int* d_count, double R, double Ldiv);
int step2d_publish(bfm_ctx* c, unsigned long long* d_errmax, int* d_count, unsigned long long* host_mapped);

// Frame output of a one-colony run: after every iteration in which the
// simulated time passes a multiple of dt_write (simulation.hpp:2023-2063),
// fn receives the cells on the host (n rows, leading dimension ld), the
// iteration count and time after the update, the accepted dt, the last
// scaled-error maximum, the wall seconds since the previous frame and the
// lineage vector.
struct FrameHook {
    double dt_write = 0;
    std::function<void(const double* cells, int n, int ld, int64_t iter, double t, double dt, double max_err,
                       double wall_elapsed, const std::vector<int>& parents)>
        fn;
};

// bfm_simulate with an optional frame hook (NULL: none). Returns a
// bfm_status and sets the context's error message.
int simulate_with_frames(bfm_ctx* c, const bfm_sim_config* cfg, const double* cells_in, int n_in, int ncols,
                         const int* parents_in, int n_parents_in, double* cells_out, int cap, int* n_out,
                         int* parents_out, int* n_parents_out, bfm_sim_stats* stats, const FrameHook* hook);

}  // namespace bfm

// The step-size controller of runSimulation (simulation.hpp:1688-1769) for
// every replicate of a batch, on the device, so the host does not read the
// error maxima back after each RK attempt. A replicate's iteration is a state
// machine: FIRST (the attempt at dt), RETRY (attempts at dt * factor while the
// scaled error exceeds 1, at most max_tries, the factor accumulated and
// clipped to [0.2, 10], clipping ending the retries), FINAL (the
// unconditional re-run at the clamped dt * factor, or the single attempt of
// an iteration without step control), DONE. Each step sets the next
// attempt's dt and mask. The block-wide steps below run in one work-group of
// 256 threads: as kernels of their own (csrc/bfm_sim.hip) or in the last
// work-group of the attempt and tail kernels of a small batch.
#pragma once

#include <hip/hip_runtime.h>

namespace bfm {

enum : int { CS_DONE = 0, CS_FIRST = 1, CS_RETRY = 2, CS_FINAL = 3 };

struct CtrlState {
    double* dt;          // step size; after the iteration, the one taken
    double* factor;      // the iteration's accumulated factor
    double* t;           // simulated time
    double* err;         // max(scaled error, min_error) of the last attempt
    int* tries;          // retries counted as the reference's j
    int* state;
    int* stepped;        // takes this iteration's step
    int* failed;         // frozen after a non-finite attempt
    long long* iters;
    int* pending;        // one int: replicates with an attempt still to take (device copy)
    long long* totals;   // device copies of CtrlPub's cells_total, attempts, failed_total
};

struct CtrlConsts {
    double min_error, p1, min_step, max_step;   // p1 = 1 / (order + 1)
    int max_tries, n_rep, stages;
    double max_time;      // the run's simulated-time limit (0: none)
};

// Published after each controller step (host-mapped memory); per replicate
// after the iteration: t, dt, err, stepped | failed << 1.
struct CtrlPub {
    int pending;          // replicates with an attempt still to take
    int failed_total;     // replicates frozen so far
    long long cells_total;   // cells of all attempts scheduled so far (RK-stage updates / stages)
    int ndiv;             // cells past the division length (end of the iteration)
    int status;           // the context's status flags at the end of the iteration
    long long attempts;   // RK attempts taken by at least one replicate so far
};

// What the host reads at the end of an iteration (host-mapped, one per
// parity of the iteration so that the next iteration may already run):
// whether the tail ran (0: a replicate still retries), the next division
// count, the status flags, whether the next iteration may run without the
// host (spec), the running totals, then per replicate t, dt, err (doubles)
// and stepped | failed << 1 (ints).
struct IterSnap {
    int tail_ran, ndiv, status, spec;
    long long cells_total, attempts;
    long long failed_total;
};

// The ordered key of -inf, the start of an attempt's scaled-error maximum
// (the attempts take atomic maxima of ordered keys).
constexpr unsigned long long kKeyNegInf = 0x000FFFFFFFFFFFFFull;

__device__ __forceinline__ double ctrl_key_value(unsigned long long k)
{
    const unsigned long long u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}

// Block-wide sums into pub: the next attempt's replicates and cells, the
// replicates frozen by the last attempt, and whether any replicate took it.
__device__ __forceinline__ void ctrl_publish(int pending, long long cells, int newf, int took, CtrlPub* pub,
                                             int* dev_pending, long long* totals)
{
    __shared__ int s_p[4], s_f[4], s_t[4];
    __shared__ long long s_c[4];
    for (int off = 32; off > 0; off >>= 1) {
        pending += __shfl_xor(pending, off);
        cells += __shfl_xor(cells, off);
        newf += __shfl_xor(newf, off);
        took += __shfl_xor(took, off);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { s_p[w] = pending; s_c[w] = cells; s_f[w] = newf; s_t[w] = took; }
    __syncthreads();
    if (threadIdx.x == 0) {
        pub->pending = (s_p[0] + s_p[1]) + (s_p[2] + s_p[3]);
        *dev_pending = pub->pending;
        pub->cells_total += (s_c[0] + s_c[1]) + (s_c[2] + s_c[3]);
        pub->failed_total += (s_f[0] + s_f[1]) + (s_f[2] + s_f[3]);
        if ((s_t[0] + s_t[1]) + (s_t[2] + s_t[3]) > 0) pub->attempts += 1;
        totals[0] = pub->cells_total;
        totals[1] = pub->attempts;
        totals[2] = pub->failed_total;
    }
}

// What the step after an attempt reads and writes. first_go non-null: the
// attempt was the iteration's first, taken by the replicates with
// first_go[r] and not failed at dt, without a set-up step before it; the
// set-up (k_ctrl_begin's) is then done here, and div_count, into which the
// iteration's tail counts the dividing cells, is zeroed.
struct CtrlAfter {
    CtrlState S;
    CtrlConsts K;
    unsigned long long* errmax;
    int* bad;
    const int* starts;
    double* trial;
    int* active;
    CtrlPub* pub;
    const int* first_go;
    int control;
    int* div_count;
    const int* gate;      // non-null: the attempt and this step run only if *gate != 0
};

// After an attempt: each replicate that took it moves on (a non-finite
// result freezes it at the state the iteration started from); its error
// maximum is reset for the next attempt. One block of 256 threads. After an
// attempt that no replicate took (nothing pending) it changes nothing.
__device__ __forceinline__ void ctrl_after_block(const CtrlAfter& A)
{
    const CtrlState& S = A.S;
    const CtrlConsts& K = A.K;
    if (A.gate && *A.gate == 0) return;
    if (!A.first_go && *S.pending == 0) return;
    int pending = 0, newf = 0, took = 0;
    long long cells = 0;
    if (A.first_go && A.div_count && threadIdx.x == 0) *A.div_count = 0;
    for (int r = threadIdx.x; r < K.n_rep; r += blockDim.x) {
        if (A.first_go) {   // k_ctrl_begin's set-up of the first attempt
            const bool on = A.first_go[r] && !S.failed[r];
            S.stepped[r] = on;
            A.active[r] = on;
            S.state[r] = on ? (A.control ? CS_FIRST : CS_FINAL) : CS_DONE;
            S.factor[r] = 1;
            S.tries[r] = 0;
            A.trial[r] = S.dt[r];
            if (on) cells += A.starts[r + 1] - A.starts[r];
        }
        if (!A.active[r]) continue;
        ++took;
        // the attempt's atomics, possibly from other work-groups of the same kernel
        const unsigned long long ekey =
            __hip_atomic_load(&A.errmax[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        A.errmax[r] = kKeyNegInf;
        if (__hip_atomic_load(&A.bad[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            A.bad[r] = 0;
            S.failed[r] = 1;
            S.stepped[r] = 0;
            S.state[r] = CS_DONE;
            A.active[r] = 0;
            ++newf;
            continue;
        }
        const double e = fmax(ctrl_key_value(ekey), K.min_error);
        S.err[r] = e;
        const double dt = S.dt[r];
        int st = S.state[r];
        bool finish = false;
        if (st == CS_FIRST) {
            double f = 0.9 * pow(1.0 / e, K.p1);
            if (f >= 10) f = 10;
            else if (f < 0.2) f = 0.2;
            S.factor[r] = f;
            if (e > 1.0 && 0 < K.max_tries) { st = CS_RETRY; A.trial[r] = dt * f; }
            else finish = true;
        } else if (st == CS_RETRY) {
            double f = S.factor[r] * (0.9 * pow(1.0 / e, K.p1));
            bool stop = false;
            if (f >= 10) { f = 10; stop = true; }
            else if (f < 0.2) { f = 0.2; stop = true; }
            S.factor[r] = f;
            int j = S.tries[r];
            if (!stop) S.tries[r] = ++j;
            if (!stop && e > 1.0 && j < K.max_tries) A.trial[r] = dt * f;
            else finish = true;
        } else {   // CS_FINAL: the iteration's step is taken
            st = CS_DONE;
            A.active[r] = 0;
        }
        if (finish) {   // clamp to [min, max] step size, the unconditional re-run
            double f = S.factor[r];
            if (dt * f < K.min_step) f = K.min_step / dt;
            else if (dt * f > K.max_step) f = K.max_step / dt;
            S.dt[r] = dt * f;
            A.trial[r] = dt * f;
            st = CS_FINAL;
        }
        S.state[r] = st;
        if (A.active[r]) { ++pending; cells += A.starts[r + 1] - A.starts[r]; }
    }
    ctrl_publish(pending, cells, newf, took, A.pub, S.pending, S.totals);
}

// What the end of an iteration writes: time and iteration counts of the
// replicates that stepped, the stepped mask and dt (active, trial) and the
// per-replicate values for the host (host-mapped h_*).
struct CtrlEnd {
    CtrlState S;
    CtrlConsts K;
    double* trial;
    int* active;
    double* h_t;
    double* h_dt;
    double* h_err;
    int* h_flags;
};

// The end of an iteration, per replicate (threads of one block).
__device__ __forceinline__ void ctrl_end_block(const CtrlEnd& E)
{
    const CtrlState& S = E.S;
    for (int r = threadIdx.x; r < E.K.n_rep; r += blockDim.x) {
        const int on = S.stepped[r];
        if (on) {
            S.t[r] += S.dt[r];
            S.iters[r] += 1;
        }
        E.active[r] = on;
        E.trial[r] = S.dt[r];
        E.h_t[r] = S.t[r];
        E.h_dt[r] = S.dt[r];
        E.h_err[r] = S.err[r];
        E.h_flags[r] = on | (S.failed[r] << 1);
    }
}

// The end-of-iteration values for the host (IterSnap) once the tail ran:
// the division count, the status flags, the totals and whether the next
// iteration may go ahead without the host: no division, no flag, one
// replicate inside its time limit and not failed (the host's other
// termination rules are known to it in advance). One thread.
__device__ __forceinline__ int publish_iteration(const CtrlEnd& E, int ndiv, int status, IterSnap* snap)
{
    const CtrlState& S = E.S;
    int spec = ndiv == 0 && status == 0 && E.K.n_rep == 1 && !S.failed[0] &&
               (E.K.max_time <= 0 || S.t[0] < E.K.max_time);
    snap->ndiv = ndiv;
    snap->status = status;
    snap->cells_total = S.totals[0];
    snap->attempts = S.totals[1];
    snap->failed_total = S.totals[2];
    snap->spec = spec;
    snap->tail_ran = 1;
    return spec;
}

// The last work-group of a grid to get here (a ticket after a device-scope
// release by each work-group) returns true, after an acquire, with every
// other work-group's writes and atomics visible; the ticket is reset for the
// next launch. Every work-group must call it, with all its threads.
__device__ __forceinline__ bool last_block_done(unsigned* ticket)
{
    __shared__ int s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return false;
    __threadfence();
    if (threadIdx.x == 0) *ticket = 0;
    return true;
}

}  // namespace bfm

// bfm_main.hip -- MI355X (gfx950) kernels and C ABI of the mechanical step.
//
// Data layout in HBM: the cells array stays column-major (each column is a
// coalesced SoA stream, include/bfm.h). The neighbor list is SoA: int32 i, j
// and five FP64 columns (dx, dy, dz, s, t), rows sorted (i asc, j asc).
//
// Pipeline of one RK stage (bfm_dev_stage, isotropic interaction):
//   k_bbox_partial, k_plan        bounding box and uniform-grid plan (device side)
//   k_keys + scan + k_scatter     counting sort of the cells into bin order
//   k_candidates_lane             lane-per-home half-stencil search: FP32 centre and
//                                 point-segment screens, FP64 segment tests, per-cell slots
//   k_slot_order                  row values (Hertz / JKR) and each cell's row order
//   k_velocities_slots            per cell: forces summed in the reference's row
//                                 order, surface forces, velocity solve
// The neighbor list proper (bfm_dev_build_neighbors) compacts the slots into
// the reference's rows: scan (64-bit), k_compact_rows, k_compact_jrole; the
// other interactions use k_pair_forces and k_velocities over those rows.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <string>
#include <vector>

#include "bfm_ctx.hpp"
#include "bfm_ctrl.hpp"
#include "bfm_internal.hpp"
#include "bfm_device.hpp"
#include "bfm_surrogate.hpp"

using namespace bfm;

// ---------------------------------------------------------------------------
// error handling and launch instrumentation
// ---------------------------------------------------------------------------
namespace {

struct HipError {
    hipError_t e;
    const char* what;
};

inline void hipchk(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw HipError{e, what};
}
#define HIPCHK(x) hipchk((x), #x)

struct ArgError {
    int code;
    std::string msg;
};

inline void require(bool ok, int code, const std::string& msg)
{
    if (!ok) throw ArgError{code, msg};
}

hipEvent_t get_event(bfm_ctx* c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
}

struct LaunchScope {
    bfm_ctx* c;
    TimedLaunch tl;
    LaunchScope(bfm_ctx* ctx, const char* name) : c(ctx)
    {
        if (c->timing) {
            tl.name = name;
            tl.a = get_event(c);
            tl.b = get_event(c);
            HIPCHK(hipEventRecord(tl.a, c->stream));
        }
    }
    ~LaunchScope()
    {
        if (c->timing) {
            if (hipEventRecord(tl.b, c->stream) == hipSuccess) c->pending.push_back(tl);
        }
    }
};

#define KLAUNCH(ctx, name, kern, grid, block, shm, ...)                         \
    do {                                                                        \
        LaunchScope _ls((ctx), (name));                                         \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), shm, (ctx)->stream,   \
                           __VA_ARGS__);                                        \
        HIPCHK(hipGetLastError());                                              \
    } while (0)

void resolve_timing(bfm_ctx* c)
{
    if (c->pending.empty()) return;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& t : c->pending) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
        auto it = c->ktimes.find(t.name);
        if (it == c->ktimes.end()) {
            c->ktimes[t.name] = {ms, 1};
            c->korder.push_back(t.name);
        } else {
            it->second.first += ms;
            it->second.second += 1;
        }
        c->event_pool.push_back(t.a);
        c->event_pool.push_back(t.b);
    }
    c->pending.clear();
}

inline int grid_for(int64_t n, int block, int cap = 1 << 20)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

constexpr int kPairGrid = 2048;   // grid-stride launches over device-sized pair counts

}  // namespace

namespace bfm {

void launch_timed(bfm_ctx* c, const char* name, const std::function<void()>& launch)
{
    LaunchScope ls(c, name);
    launch();
    HIPCHK(hipGetLastError());
}

}  // namespace bfm

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
namespace bfm {

__device__ __forceinline__ double cellv(const double* __restrict__ cells, int ld, int i, int c)
{
    return cells[(int64_t)c * ld + i];
}

__global__ void k_fill_int(int* __restrict__ p, int v, int64_t n)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        p[q] = v;
}

__global__ void k_iota(int* __restrict__ p, int n)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) p[q] = q;
}

// Block partials of min/max of the centers and max half-length.
__device__ __forceinline__ void bbox_block(const double* __restrict__ cells, int n, int ld,
                                           double* __restrict__ partials)
{
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double hl = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double v = cellv(cells, ld, i, C_RX + k);
            lo[k] = fmin(lo[k], v);
            hi[k] = fmax(hi[k], v);
        }
        hl = fmax(hl, cellv(cells, ld, i, C_HALF_L));
    }
    __shared__ double red[7][256];
#pragma unroll
    for (int k = 0; k < 3; ++k) { red[k][threadIdx.x] = lo[k]; red[3 + k][threadIdx.x] = hi[k]; }
    red[6][threadIdx.x] = hl;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + w]);
                red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + w]);
            }
            red[6][threadIdx.x] = fmax(red[6][threadIdx.x], red[6][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 7; ++k) partials[blockIdx.x * 7 + k] = red[k][0];
}

__global__ void __launch_bounds__(256) k_bbox_partial(const double* __restrict__ cells, int n, int ld,
                                                      double* __restrict__ partials)
{
    bbox_block(cells, n, ld, partials);
}

// Final reduction and grid plan. The search cutoff covers every pair the
// reference can keep: |r_i - r_j| < thr + Ldiv + 2R (its prefilter) and,
// since |d| >= |r_i - r_j| - hl_i - hl_j, |r_i - r_j| < thr + 2 max(hl).
__device__ __forceinline__ void plan_from_partials(const double* __restrict__ partials, int nparts, double thr,
                                                   double pre, double bin_scale, int max_bins,
                                                   Plan* __restrict__ plan, RepGrid* __restrict__ grid)
{
    __shared__ double red[7][256];
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        red[k][tid] = tid < nparts ? partials[tid * 7 + k] : INFINITY;
        red[3 + k][tid] = tid < nparts ? partials[tid * 7 + 3 + k] : -INFINITY;
    }
    red[6][tid] = tid < nparts ? partials[tid * 7 + 6] : 0.0;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                red[k][tid] = fmin(red[k][tid], red[k][tid + w]);
                red[3 + k][tid] = fmax(red[3 + k][tid], red[3 + k][tid + w]);
            }
            red[6][tid] = fmax(red[6][tid], red[6][tid + w]);
        }
        __syncthreads();
    }
    if (tid != 0) return;
    double lo[3] = {red[0][0], red[1][0], red[2][0]}, hi[3] = {red[3][0], red[4][0], red[5][0]};
    double hl = red[6][0];
    double cut = fmin(pre, thr + 2 * hl) * (1 + 1e-9) + 1e-9;
    double edge = cut * bin_scale;
    Plan P;
    for (;;) {
        int reach = (int)ceil(cut / edge - 1e-12);
        if (reach < 1) reach = 1;
        if (reach * edge < cut) reach++;
        int64_t tot = 1;
        int dims[3];
        for (int k = 0; k < 3; ++k) {
            double span = hi[k] - lo[k];
            dims[k] = (int)floor(span / edge) + 1;
            tot *= dims[k];
        }
        if (tot <= max_bins) {
            for (int k = 0; k < 3; ++k) { P.lo[k] = lo[k]; P.dims[k] = dims[k]; }
            P.reach = reach;
            P.nbins = (int)tot;
            break;
        }
        edge *= 1.25;
    }
    P.edge = edge;
    P.inv_edge = 1.0 / edge;
    // FP32 copies of coordinates carry errors <= 2^-24 |x|; a difference of two
    // of them <= 2^-23 max|x| ~ 1.2e-7 max|x|, squared distances add a few ulps.
    double amax = 0;
    for (int k = 0; k < 3; ++k) amax = fmax(amax, fmax(fabs(lo[k]), fabs(hi[k])));
    P.fmargin = 1e-3 + 1e-6 * amax;
    P.pre = pre;
    P.thr = thr;
    P.ok = 1;
    *plan = P;
    RepGrid G;
    for (int k = 0; k < 3; ++k) { G.lo[k] = P.lo[k]; G.dims[k] = P.dims[k]; }
    G.base = 0;
    *grid = G;
}

__global__ void __launch_bounds__(256) k_plan(const double* __restrict__ partials, int nparts,
                                              double thr, double pre, double bin_scale, int max_bins,
                                              Plan* __restrict__ plan, RepGrid* __restrict__ grid)
{
    plan_from_partials(partials, nparts, thr, pre, bin_scale, max_bins, plan, grid);
}

__device__ __forceinline__ int bin_coord(double x, double lo, double edge, int dim)
{
    int b = (int)floor((x - lo) / edge);
    return b < 0 ? 0 : (b >= dim ? dim - 1 : b);
}

__global__ void k_keys(const double* __restrict__ cells, int n, int ld, const Plan* __restrict__ plan,
                       int* __restrict__ keys, int* __restrict__ vals, int* __restrict__ hist)
{
    const Plan P = *plan;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int bx = bin_coord(cellv(cells, ld, i, C_RX), P.lo[0], P.edge, P.dims[0]);
        int by = bin_coord(cellv(cells, ld, i, C_RY), P.lo[1], P.edge, P.dims[1]);
        int bz = bin_coord(cellv(cells, ld, i, C_RZ), P.lo[2], P.edge, P.dims[2]);
        int key = (bz * P.dims[1] + by) * P.dims[0] + bx;
        keys[i] = key;
        vals[i] = atomicAdd(&hist[key], 1);   // rank of the cell within its bin
    }
}

// Exclusive scan of the bin histogram over the plan's nbins + 1 entries only
// (its capacity is max_bins + 1, several times the bins a plan uses), over
// 2048-entry tiles. Entries: plan->nbins + 1 when plan is given, else
// `items`.
constexpr int kScanTile = 2048;
constexpr int kMaxScanTiles = 1 << 14;   // 33.5M entries

// One pass: tile b scans its entries and publishes its aggregate; then its
// threads read the aggregates of all earlier tiles in parallel (a thread per
// earlier tile, waiting for its flag), and the block sums them into the
// tile's prefix. A flag word holds epoch << 2 | 1; the epoch changes with
// every launch, so the flags need no clearing between launches. Blocks are
// dispatched in index order and a tile waits only for earlier ones, which
// publish first thing; at a few hundred tiles all are resident at once and
// the wait is about one round trip;
// it gives up after a bounded number of polls with ST_QUEUE in the status
// word instead of waiting forever.
constexpr int kLookbackPolls = 1 << 16;

template <typename T>
__global__ void __launch_bounds__(256)
k_scan_onepass(const Plan* __restrict__ plan, int64_t items_arg, const T* __restrict__ in, T* __restrict__ out,
               T* __restrict__ agg_val, unsigned* __restrict__ flag, unsigned epoch, int* __restrict__ status)
{
    __shared__ T wtot[4];
    __shared__ T wpre[4];
    const int64_t items = plan ? (int64_t)plan->nbins + 1 : items_arg;
    const int b = blockIdx.x;
    if ((int64_t)b * kScanTile >= items) return;
    constexpr int per = kScanTile / 256;
    const int64_t i0 = (int64_t)b * kScanTile + per * threadIdx.x;
    T v[per], tot = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        v[k] = i0 + k < items ? in[i0 + k] : (T)0;
        tot += v[k];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T incl = tot;
    for (int o = 1; o < 64; o <<= 1) {
        const T u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
    }
    if (lane == 63) wtot[w] = incl;
    __syncthreads();
    const unsigned tag = (epoch << 2) | 1u;
    if (threadIdx.x == 0) {
        agg_val[b] = (wtot[0] + wtot[1]) + (wtot[2] + wtot[3]);
        __hip_atomic_store(&flag[b], tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    T pre = 0;
    for (int p = threadIdx.x; p < b; p += 256) {
        int polls = 0;
        while (__hip_atomic_load(&flag[p], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != tag) {
            if (++polls > kLookbackPolls) {
                atomicOr(status, ST_QUEUE);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        pre += agg_val[p];
    }
    for (int o = 32; o > 0; o >>= 1) pre += __shfl_xor(pre, o);
    if (lane == 0) wpre[w] = pre;
    __syncthreads();
    T run = ((wpre[0] + wpre[1]) + (wpre[2] + wpre[3])) + incl - tot;
    for (int k = 0; k < w; ++k) run += wtot[k];
#pragma unroll
    for (int k = 0; k < per; ++k) {
        if (i0 + k < items) out[i0 + k] = run;
        run += v[k];
    }
}

// Counting-sort scatter: each cell goes to its bin's offset plus its rank
// within the bin (taken by the histogram atomic of k_keys), with its geometry. The
// order inside a bin depends on atomic timing, but nothing downstream does:
// every pair is tested from the larger original index, and each cell's list
// is sorted by partner index during compaction.
// It also clears the per-cell pair counters (n + 1) that the candidate kernel
// increments or writes next, and the histogram entries of the occupied bins (every
// nonzero entry has a cell), which leaves the histogram zero for the next build.
template <bool SOA>
__global__ void k_scatter(const double* __restrict__ cells, int n, int ld, const int* __restrict__ keys,
                          const int* __restrict__ rank, const int* __restrict__ bin_off,
                          double* __restrict__ sorted, int* __restrict__ sid, int* __restrict__ skey, int64_t stride,
                          unsigned long long* __restrict__ cnt64, unsigned long long* __restrict__ own64,
                          int* __restrict__ hist, float4* __restrict__ sortf, double* __restrict__ sgeom)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt64[n] = own64[n] = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        cnt64[i] = 0;
        own64[i] = 0;
        const int key = keys[i];
        hist[key] = 0;
        const int q = bin_off[key] + rank[i];
        sid[q] = i;
        skey[q] = key;
        const double rx = cellv(cells, ld, i, C_RX), ry = cellv(cells, ld, i, C_RY), rz = cellv(cells, ld, i, C_RZ);
        const double nx = cellv(cells, ld, i, C_NX), ny = cellv(cells, ld, i, C_NY), nz = cellv(cells, ld, i, C_NZ);
        const double hl = cellv(cells, ld, i, C_HALF_L);
        if constexpr (SOA) {   // the LDS-staged candidate kernels
            sorted[0 * stride + q] = rx;
            sorted[1 * stride + q] = ry;
            sorted[2 * stride + q] = rz;
            sorted[3 * stride + q] = nx;
            sorted[4 * stride + q] = ny;
            sorted[5 * stride + q] = nz;
            sorted[6 * stride + q] = hl;
        } else {   // the lane-per-home search: one 64-byte segment record per cell
            double2* g = reinterpret_cast<double2*>(sgeom + 8 * (int64_t)q);
            g[0] = make_double2(rx, ry);
            g[1] = make_double2(rz, nx);
            g[2] = make_double2(ny, nz);
            g[3] = make_double2(hl, 0.0);
            // FP32 copies for its screens
            sortf[q] = make_float4((float)rx, (float)ry, (float)rz, (float)hl);
            sortf[stride + q] = make_float4((float)nx, (float)ny, (float)nz, 0.f);
        }
    }
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t n)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        p[q] = 0ull;
}

__device__ __forceinline__ Seg load_sorted(const double* __restrict__ s, int64_t stride, int q)
{
    Seg c;
    c.r[0] = s[q]; c.r[1] = s[stride + q]; c.r[2] = s[2 * stride + q];
    c.n[0] = s[3 * stride + q]; c.n[1] = s[4 * stride + q]; c.n[2] = s[5 * stride + q];
    c.hl = s[6 * stride + q];
    return c;
}

// MI355X dispatches work-groups to its 8 XCDs round-robin (work-group b runs
// on XCD b % 8), and each XCD has its own L2. This bijection gives every XCD
// one contiguous range of logical blocks, so the bins a range stages (a slab
// of the grid plus its halo) stay in that XCD's L2 instead of being fetched
// by all eight.
__device__ __forceinline__ int xcd_block(int b, int nblocks)
{
    const int q = nblocks >> 3, r = nblocks & 7;
    const int xcd = b & 7, k = b >> 3;
    return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + k;
}

// Build statistics (screened pairs, segment tests) go to one of 256
// counter slots, 128 bytes apart, chosen by the work-group: atomics on one
// global address from every work-group serialise in L2 (~12 ns each on
// MI355X, measured), which would put a floor under the kernel's duration.
constexpr int kCounterSlots = 256;
constexpr int kCounterStride = 16;   // unsigned long longs per slot: two counters on their own 64-byte lines
__device__ __forceinline__ void count_add(unsigned long long* counters, int which, unsigned long long v)
{
    if (v) atomicAdd(&counters[(blockIdx.x & (kCounterSlots - 1)) * kCounterStride + which * 8], v);
}

// LDS-staged candidate search over a half stencil. A workgroup owns HOMES
// consecutive bin-ordered cells; homes sharing a grid row (same y, z bin)
// form a group. Every unordered pair is screened once, by the cell that
// comes later in bin order: for a group only bins ordered before the home
// cells are candidates -- the stencil rows below the home row (full x window
// [xmin - reach, xmax + reach]) and, in the home row, bins xmin - reach .. xmax
// up to the last home. These ranges are contiguous in sorted order and are
// staged into LDS in chunks of CAP cells. Each thread owns staged candidates
// and screens them against a batch of BATCH homes (LDS broadcasts); survivors
// (at most BATCH * CAP: the queue cannot overflow) are tested with the
// segment kernel by all threads. A kept pair becomes row (i, j) with
// i = max(original index), j = min, computed as distBetweenCells(cell i,
// cell j) exactly as the reference's j < i loop does; its slot in i's list is
// taken with a global atomic, and i is also appended to j's list of
// partners with larger index (the compaction sorts both lists).
template <int HOMES, int CAP, int BATCH, int QCAP>
__global__ void __launch_bounds__(256)
k_candidates_lds(int n, const Plan* __restrict__ plan, const double* __restrict__ sorted, int64_t stride,
                 const int* __restrict__ sid, const int* __restrict__ skey,
                 const int* __restrict__ rep, const RepGrid* __restrict__ grid,
                 const int* __restrict__ bin_off, int slot_cap, unsigned long long* __restrict__ cnt64,
                 int* __restrict__ slot_j, double* __restrict__ slot_rec,
                 int* __restrict__ jslot, unsigned long long* __restrict__ counters)
{
    // queue entry = (home - batch start) << TB | candidate, 16 bits
    constexpr int TB = CAP <= 512 ? 9 : 10;
    static_assert(HOMES <= (1 << (16 - TB)) && CAP <= (1 << TB) && BATCH * CAP <= QCAP,
                  "queue entry packing / bound");
    __shared__ float4 c_f[CAP];             // FP32 copy (x, y, z, hl) for the pre-screen
    __shared__ int c_id[CAP], c_q[CAP], c_bx[CAP];
    __shared__ uint16_t queue[QCAP];
    __shared__ int s_ovf;
    __shared__ double h_x[HOMES], h_y[HOMES], h_z[HOMES], h_nx[HOMES], h_ny[HOMES], h_nz[HOMES], h_hl[HOMES];
    __shared__ int h_id[HOMES], h_bx[HOMES], h_row[HOMES], h_rep[HOMES];
    __shared__ int h_gbase[HOMES], h_gd0[HOMES], h_gd1[HOMES], h_gd2[HOMES];   // the home's replicate grid
    __shared__ float4 h_f[HOMES];           // FP32 (x, y, z, thr + hl + 1e-3)
    __shared__ int seg_lo[49], seg_pre[50];
    __shared__ int s_qn;
    __shared__ unsigned long long s_screened;

    const int tid = threadIdx.x;
    const int h0 = xcd_block(blockIdx.x, gridDim.x) * HOMES;
    if (h0 >= n) return;
    const int H = min(HOMES, n - h0);
    const Plan P = *plan;
    const double thr = P.thr, pre = P.pre;
    const int reach = P.reach;
    if (tid < H) {
        const int hs = h0 + tid;
        h_x[tid] = sorted[hs]; h_y[tid] = sorted[stride + hs]; h_z[tid] = sorted[2 * stride + hs];
        h_nx[tid] = sorted[3 * stride + hs]; h_ny[tid] = sorted[4 * stride + hs];
        h_nz[tid] = sorted[5 * stride + hs]; h_hl[tid] = sorted[6 * stride + hs];
        h_id[tid] = sid[hs];
        h_f[tid] = make_float4((float)h_x[tid], (float)h_y[tid], (float)h_z[tid],
                               (float)(thr + h_hl[tid] + P.fmargin));
        // bins of replicate r are numbered grid[r].base + (z dims_y + y) dims_x + x;
        // key - x identifies the (replicate, y, z) row
        const int key = skey[hs];
        const int r = rep ? rep[h_id[tid]] : 0;
        const int gbase = grid[r].base, gd0 = grid[r].dims[0];
        const int bx = (key - gbase) % gd0;
        h_bx[tid] = bx;
        h_row[tid] = key - bx;
        h_rep[tid] = r;
        h_gbase[tid] = gbase;
        h_gd0[tid] = gd0;
        h_gd1[tid] = grid[r].dims[1];
        h_gd2[tid] = grid[r].dims[2];
    }
    if (tid == 0) s_screened = 0;
    __syncthreads();
    unsigned long long screened = 0, tested = 0;
    const int side = 2 * reach + 1;
    const int nseg = reach * side + reach + 1;   // rows before the home row, then the home row

    for (int g0 = 0; g0 < H;) {
        const int row = h_row[g0];
        int g1 = g0 + 1;
        while (g1 < H && h_row[g1] == row) ++g1;
        const int xmin = h_bx[g0], xmax = h_bx[g1 - 1];
        const int gbase = h_gbase[g0], gd0 = h_gd0[g0], gd1 = h_gd1[g0], gd2 = h_gd2[g0];
        const int rrow = (row - gbase) / gd0;
        const int by = rrow % gd1, bz = rrow / gd1;
        const int last = h0 + g1 - 1;                  // sorted position of the group's last home
        if (tid < 64) {   // the first wave: segment bounds and their prefix sums (nseg <= 49)
            int lo = 0, len = 0;
            if (tid < nseg) {
                // rows (dz, dy) lexicographically before (0, 0), then (0, 0) itself
                const int dz = tid / side - reach, dy = tid % side - reach;
                const int zz = bz + dz, yy = by + dy;
                if (zz >= 0 && zz < gd2 && yy >= 0 && yy < gd1) {
                    const int rb = gbase + (zz * gd1 + yy) * gd0;
                    const int xlo = max(0, xmin - reach);
                    const int xhi = (dz == 0 && dy == 0) ? xmax : min(gd0 - 1, xmax + reach);
                    lo = bin_off[rb + xlo];
                    int hi = bin_off[rb + xhi + 1];
                    if (dz == 0 && dy == 0) hi = min(hi, last);
                    len = max(0, hi - lo);
                }
            }
            int incl = len;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int v = __shfl_up(incl, off);
                if (tid >= off) incl += v;
            }
            if (tid < nseg) {
                seg_lo[tid] = lo;
                seg_pre[tid + 1] = incl;
            }
            if (tid == 0) seg_pre[0] = 0;
        }
        __syncthreads();
        const int S = seg_pre[nseg];
        for (int c0 = 0; c0 < S; c0 += CAP) {
            const int C = min(CAP, S - c0);
            for (int t = tid; t < C; t += 256) {
                const int gidx = c0 + t;
                int sg = 0;
                while (seg_pre[sg + 1] <= gidx) ++sg;
                const int q = seg_lo[sg] + (gidx - seg_pre[sg]);
                c_f[t] = make_float4((float)sorted[q], (float)sorted[stride + q], (float)sorted[2 * stride + q],
                                     (float)sorted[6 * stride + q]);
                c_id[t] = sid[q];
                c_q[t] = q;
                c_bx[t] = (skey[q] - gbase) % gd0;
            }
            __syncthreads();
            // First try the whole group as one batch; if the queue would
            // overflow (only in extremely dense packings), discard it and
            // redo the chunk in batches of BATCH homes, whose queue is bounded
            // by BATCH * CAP <= QCAP. No test has run before the check.
            for (int attempt = 0; attempt < 2; ++attempt) {
                const int BT = attempt == 0 ? (g1 - g0) : BATCH;
                bool overflow = false;
                for (int b0 = g0; b0 < g1; b0 += BT) {
                    const int G = min(BT, g1 - b0);
                    if (tid == 0) { s_qn = 0; s_ovf = 0; }
                    __syncthreads();
                    for (int t = tid; t < C; t += 256) {
                        const float4 cf = c_f[t];
                        const int bx = c_bx[t];
                        // homes after the candidate in sorted order, with |bx_h - bx| <= reach
                        // homes of the group are sorted by bx: binary searches for the
                        // first home with bx >= bx - reach and the first with bx > bx + reach
                        int hl0 = g0, hi0 = g1;
                        while (hl0 < hi0) {
                            const int mid = (hl0 + hi0) >> 1;
                            if (h_bx[mid] < bx - reach) hl0 = mid + 1; else hi0 = mid;
                        }
                        int hr0 = hl0;
                        hi0 = g1;
                        while (hr0 < hi0) {
                            const int mid = (hr0 + hi0) >> 1;
                            if (h_bx[mid] <= bx + reach) hr0 = mid + 1; else hi0 = mid;
                        }
                        const int first = max(max(b0, hl0), c_q[t] - h0 + 1);
                        const int stop = min(b0 + G, hr0);
                        for (int h = first; h < stop; ++h) {
                            ++screened;
                            // FP32 pre-screen: the rounding error of the FP32 distance is far
                            // below the margin folded into h_f.w (1e-3 um + 1e-6 max|x|), so no
                            // pair the FP64 test below would keep is dropped.
                            const float4 hf = h_f[h];
                            const float fx = hf.x - cf.x, fy = hf.y - cf.y, fz = hf.z - cf.z;
                            const float fb = hf.w + cf.w;
                            if (fx * fx + fy * fy + fz * fz <= fb * fb) {
                                const int pos = atomicAdd(&s_qn, 1);
                                if (pos < QCAP) queue[pos] = (uint16_t)(((h - b0) << TB) | t);
                                else s_ovf = 1;
                            }
                        }
                    }
                    __syncthreads();
                    if (s_ovf) { overflow = true; break; }
                    const int qn = s_qn;
                    for (int e = tid; e < qn; e += 256) {
                        const int ent = queue[e];
                        const int h = b0 + (ent >> TB), t = ent & ((1 << TB) - 1);
                        Seg ch, cc;
                        ch.r[0] = h_x[h]; ch.r[1] = h_y[h]; ch.r[2] = h_z[h];
                        ch.n[0] = h_nx[h]; ch.n[1] = h_ny[h]; ch.n[2] = h_nz[h]; ch.hl = h_hl[h];
                        cc = load_sorted(sorted, stride, c_q[t]);
                        const bool home_is_i = h_id[h] > c_id[t];
                        const int i = home_is_i ? h_id[h] : c_id[t];
                        const int j = home_is_i ? c_id[t] : h_id[h];
                        // |r_i - r_j| bound (below it |d| cannot reach thr) and the
                        // reference prefilter |r_i - r_j| < thr + Ldiv + 2R
                        const double dx = ch.r[0] - cc.r[0], dy = ch.r[1] - cc.r[1], dz = ch.r[2] - cc.r[2];
                        const double r2 = (dx * dx + dy * dy) + dz * dz;
                        const double bound = (thr + ch.hl + cc.hl) * (1 + 1e-9) + 1e-9;
                        if (!(r2 <= bound * bound && sqrt(r2) < pre)) continue;
                        double d[3], sv, tv;
                        if (home_is_i) dist_between_cells(ch, cc, d, sv, tv);
                        else dist_between_cells(cc, ch, d, sv, tv);
                        if (norm3(d[0], d[1], d[2]) < thr) {
                            // high half: rows of i (i-role); low half: rows where j is the partner
                            const int slot = (int)(atomicAdd(&cnt64[i], 1ull << 32) >> 32);
                            const int slot2 = (int)(atomicAdd(&cnt64[j], 1ull) & 0xffffffffull);
                            if (slot < slot_cap) {
                                const int64_t e = (int64_t)slot * stride + i;
                                slot_j[e] = j;
                                double2* r = reinterpret_cast<double2*>(slot_rec + 8 * e);
                                r[0] = make_double2(d[0], d[1]);
                                r[1] = make_double2(d[2], sv);
                                r[2] = make_double2(tv, __longlong_as_double((long long)slot2));
                            }
                            if (slot2 < slot_cap) jslot[(int64_t)slot2 * stride + j] = i;
                        }
                    }
                    if (tid == 0) tested += (unsigned long long)qn;
                    __syncthreads();
                }
                if (!overflow) break;
                __syncthreads();
            }
        }
        g0 = g1;
    }
    atomicAdd(&s_screened, screened);
    __syncthreads();
    if (tid == 0) {
        count_add(counters, 0, s_screened);
        count_add(counters, 1, tested);
    }
}

// Wavefront-per-bin candidate search (no work-group barriers; the four
// wavefronts of a work-group are independent). A wavefront takes one home bin
// at a time: it stages up to 64 of the bin's cells (one per lane) in its own
// LDS area and sweeps the bin's half stencil 64 candidates at a time, one per
// lane. Every lane screens its candidate against all homes of the chunk (FP32,
// broadcast LDS reads, a trip count uniform across the wavefront); survivors
// go to a wavefront-private LDS queue whose slices of 64 are tested with the
// FP64 segment kernel, one pair per lane. The pair rule is that of
// k_candidates_lds: a pair is screened once, by the cell later in bin order
// (candidates in the stencil rows before the home row, then in the home row
// up to the home bin, at sorted positions before the home). Bins are split
// into eight contiguous ranges, one per XCD (work-group b runs on XCD b % 8),
// and the wavefronts of an XCD stride through its range.
template <int QCAP>
__global__ void __launch_bounds__(256)
k_candidates_bins(const Plan* __restrict__ plan, const double* __restrict__ sorted, int64_t stride,
                  const int* __restrict__ sid, const RepGrid* __restrict__ grid, int n_grids,
                  const int* __restrict__ bin_off, int slot_cap, unsigned long long* __restrict__ cnt64,
                  int* __restrict__ slot_j, double* __restrict__ slot_rec, int* __restrict__ jslot,
                  unsigned long long* __restrict__ counters)
{
    constexpr int WAVES = 4;
    static_assert(QCAP >= 128 && QCAP % 64 == 0, "queue holds at least two slices");
    __shared__ float4 s_hf[WAVES][64];          // FP32 (x, y, z, thr + hl + margin) of the homes
    __shared__ double s_h[WAVES][7][64];        // FP64 geometry of the homes
    __shared__ int s_hid[WAVES][64];
    __shared__ int s_queue[WAVES][QCAP];        // (candidate sorted position << 6) | home
    __shared__ int s_seg[WAVES][2][64];         // segment start, inclusive prefix of lengths
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const Plan P = *plan;
    const double thr = P.thr, pre = P.pre;
    const float fmargin = (float)P.fmargin;
    const int reach = P.reach, side = 2 * reach + 1, nseg = reach * side + reach + 1;
    const int nb = P.nbins;
    const int xcd = blockIdx.x & 7, kx = blockIdx.x >> 3;
    const int blocks_x = (gridDim.x - xcd + 7) >> 3;          // blocks that run on this XCD
    const int waves_x = blocks_x * WAVES;
    const int r_lo = (int)((int64_t)nb * xcd / 8), r_hi = (int)((int64_t)nb * (xcd + 1) / 8);
    float4* hf = s_hf[wave];
    int* queue = s_queue[wave];
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    unsigned long long screened = 0, tested = 0;
    int qn = 0;   // queue fill, uniform across the wavefront

    // tests entries [0, qn) of the queue, 64 per round, then empties it
    auto drain = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int e0 = 0; e0 < qn; e0 += 64) {
            const int e = e0 + lane;
            if (e < qn) {
                const int ent = queue[e];
                const int h = ent & 63, q = ent >> 6;
                Seg ch, cc;
                ch.r[0] = s_h[wave][0][h]; ch.r[1] = s_h[wave][1][h]; ch.r[2] = s_h[wave][2][h];
                ch.n[0] = s_h[wave][3][h]; ch.n[1] = s_h[wave][4][h]; ch.n[2] = s_h[wave][5][h];
                ch.hl = s_h[wave][6][h];
                cc = load_sorted(sorted, stride, q);
                const int hid = s_hid[wave][h], cid = sid[q];
                const bool home_is_i = hid > cid;
                const int i = home_is_i ? hid : cid;
                const int j = home_is_i ? cid : hid;
                const double dx = ch.r[0] - cc.r[0], dy = ch.r[1] - cc.r[1], dz = ch.r[2] - cc.r[2];
                const double r2 = (dx * dx + dy * dy) + dz * dz;
                const double bound = (thr + ch.hl + cc.hl) * (1 + 1e-9) + 1e-9;
                if (r2 <= bound * bound && sqrt(r2) < pre) {
                    double d[3], sv, tv;
                    if (home_is_i) dist_between_cells(ch, cc, d, sv, tv);
                    else dist_between_cells(cc, ch, d, sv, tv);
                    if (norm3(d[0], d[1], d[2]) < thr) {
                        const int slot = (int)(atomicAdd(&cnt64[i], 1ull << 32) >> 32);
                        const int slot2 = (int)(atomicAdd(&cnt64[j], 1ull) & 0xffffffffull);
                        if (slot < slot_cap) {
                            const int64_t es = (int64_t)slot * stride + i;
                            slot_j[es] = j;
                            double2* rr = reinterpret_cast<double2*>(slot_rec + 8 * es);
                            rr[0] = make_double2(d[0], d[1]);
                            rr[1] = make_double2(d[2], sv);
                            rr[2] = make_double2(tv, __longlong_as_double((long long)slot2));
                        }
                        if (slot2 < slot_cap) jslot[(int64_t)slot2 * stride + j] = i;
                    }
                }
            }
        }
        tested += (unsigned long long)qn;
        qn = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    for (int b = r_lo + kx * WAVES + wave; b < r_hi; b += waves_x) {
        const int hb0 = bin_off[b], hb1 = bin_off[b + 1];
        if (hb0 >= hb1) continue;
        // the previous bin's LDS reads are done before this bin's writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        int r = 0;
        if (n_grids > 1) {   // the replicate whose bins contain b
            int lo = 0, hi = n_grids - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (grid[mid].base <= b) lo = mid; else hi = mid - 1;
            }
            r = lo;
        }
        const int gbase = grid[r].base, gd0 = grid[r].dims[0], gd1 = grid[r].dims[1], gd2 = grid[r].dims[2];
        const int local = b - gbase;
        const int bx = local % gd0, rrow = local / gd0, by = rrow % gd1, bz = rrow / gd1;
        // the half stencil as contiguous runs of sorted positions: rows (dz, dy)
        // before (0, 0) over [bx - reach, bx + reach], then the home row up to bx
        int lo = 0, len = 0;
        if (lane < nseg) {
            const int dz = lane / side - reach, dy = lane % side - reach;
            const int zz = bz + dz, yy = by + dy;
            if (zz >= 0 && zz < gd2 && yy >= 0 && yy < gd1) {
                const int rb = gbase + (zz * gd1 + yy) * gd0;
                const int xlo = max(0, bx - reach);
                const int xhi = (dz == 0 && dy == 0) ? bx : min(gd0 - 1, bx + reach);
                lo = bin_off[rb + xlo];
                len = bin_off[rb + xhi + 1] - lo;
            }
        }
        int incl = len;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        s_seg[wave][0][lane] = lo;
        s_seg[wave][1][lane] = incl;
        const int S = __shfl(incl, nseg - 1);
        for (int h0 = hb0; h0 < hb1; h0 += 64) {
            const int H = min(64, hb1 - h0);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane < H) {
                const int hs = h0 + lane;
                double g7[7];
#pragma unroll
                for (int k = 0; k < 7; ++k) {
                    g7[k] = sorted[k * stride + hs];
                    s_h[wave][k][lane] = g7[k];
                }
                s_hid[wave][lane] = sid[hs];
                hf[lane] = make_float4((float)g7[0], (float)g7[1], (float)g7[2], (float)(thr + g7[6] + P.fmargin));
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int qlast = h0 + H - 1;   // a candidate pairs only with homes after it
            for (int c0 = 0; c0 < S; c0 += 64) {
                const int g = c0 + lane;
                bool have = g < S;
                int q = 0;
                float4 cf = make_float4(0.f, 0.f, 0.f, 0.f);
                if (have) {
                    int sg = 0;
                    while (s_seg[wave][1][sg] <= g) ++sg;
                    q = s_seg[wave][0][sg] + (g - (sg ? s_seg[wave][1][sg - 1] : 0));
                    have = q < qlast;
                    if (have)
                        cf = make_float4((float)sorted[q], (float)sorted[stride + q], (float)sorted[2 * stride + q],
                                         (float)sorted[6 * stride + q]);
                }
                if (!__ballot(have)) continue;
                // homes at or before the candidate never pair with it: start at
                // the first home after the lowest candidate of the wavefront
                for (int h = 0; h < H; ++h) {
                    bool pass = false;
                    if (have && q < h0 + h) {
                        ++screened;
                        const float4 hh = hf[h];
                        const float fx = hh.x - cf.x, fy = hh.y - cf.y, fz = hh.z - cf.z;
                        const float fb = hh.w + cf.w;
                        pass = fx * fx + fy * fy + fz * fz <= fb * fb;
                    }
                    const uint64_t m = __ballot(pass);
                    if (m) {
                        if (pass) queue[qn + __popcll(m & below)] = (q << 6) | h;
                        qn += __popcll(m);
                        if (qn > QCAP - 64) drain();
                    }
                }
            }
            if (qn) drain();
        }
    }
    (void)fmargin;
    // one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) {
        screened += __shfl_xor(screened, off);
    }
    if (lane == 0) {
        count_add(counters, 0, screened);
        count_add(counters, 1, tested);
    }
}

// The candidate search of k_candidates_lds with the screen turned around: a
// work-group owns 64 consecutive bin-ordered homes, lane l of every wave
// holds home l (FP32 position and bound in registers), and each wave takes
// staged candidates one at a time (broadcast from LDS) and screens it against
// all 64 homes at once. Lanes whose home is not in the current group, not
// within `reach` bins in x, or not after the candidate in bin order are idle.
// Hits are appended with one LDS atomic per wave (ballot + mbcnt). The test
// of the survivors and everything written are identical to k_candidates_lds.
template <int CAP, int BATCH, int QCAP>
__global__ void __launch_bounds__(256)
k_candidates_wave(int n, const Plan* __restrict__ plan, const double* __restrict__ sorted, int64_t stride,
                  const int* __restrict__ sid, const int* __restrict__ skey,
                  const int* __restrict__ rep, const RepGrid* __restrict__ grid,
                  const int* __restrict__ bin_off, int slot_cap, unsigned long long* __restrict__ cnt64,
                  int* __restrict__ slot_j, double* __restrict__ slot_rec,
                  int* __restrict__ jslot, unsigned long long* __restrict__ counters)
{
    constexpr int HOMES = 64;
    static_assert(CAP <= 1024 && BATCH * CAP <= QCAP, "queue entry packing / bound");
    __shared__ float4 c_f[CAP];
    __shared__ int c_q[CAP], c_bx[CAP];
    __shared__ uint16_t queue[QCAP];
    __shared__ int s_ovf;
    __shared__ double h_x[HOMES], h_y[HOMES], h_z[HOMES], h_nx[HOMES], h_ny[HOMES], h_nz[HOMES], h_hl[HOMES];
    __shared__ int h_id[HOMES], h_bx[HOMES], h_row[HOMES], h_rep[HOMES];
    __shared__ int seg_lo[49], seg_pre[50];
    __shared__ int s_qn;
    __shared__ unsigned long long s_screened;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int h0 = xcd_block(blockIdx.x, gridDim.x) * HOMES;
    if (h0 >= n) return;
    const int H = min(HOMES, n - h0);
    const Plan P = *plan;
    const double thr = P.thr, pre = P.pre;
    const int reach = P.reach;
    if (tid < H) {
        const int hs = h0 + tid;
        h_x[tid] = sorted[hs]; h_y[tid] = sorted[stride + hs]; h_z[tid] = sorted[2 * stride + hs];
        h_nx[tid] = sorted[3 * stride + hs]; h_ny[tid] = sorted[4 * stride + hs];
        h_nz[tid] = sorted[5 * stride + hs]; h_hl[tid] = sorted[6 * stride + hs];
        h_id[tid] = sid[hs];
        const int key = skey[hs];
        const int r = rep ? rep[h_id[tid]] : 0;
        const int bx = (key - grid[r].base) % grid[r].dims[0];
        h_bx[tid] = bx;
        h_row[tid] = key - bx;
        h_rep[tid] = r;
    }
    if (tid == 0) s_screened = 0;
    __syncthreads();
    // this lane's home, in registers
    float4 hf = make_float4(0.f, 0.f, 0.f, 0.f);
    int my_bx = 0;
    if (lane < H) {
        hf = make_float4((float)h_x[lane], (float)h_y[lane], (float)h_z[lane],
                         (float)(thr + h_hl[lane] + P.fmargin));
        my_bx = h_bx[lane];
    }
    unsigned long long screened = 0, tested = 0;
    const int side = 2 * reach + 1;
    const int nseg = reach * side + reach + 1;

    for (int g0 = 0; g0 < H;) {
        const int row = h_row[g0];
        int g1 = g0 + 1;
        while (g1 < H && h_row[g1] == row) ++g1;
        const int xmin = h_bx[g0], xmax = h_bx[g1 - 1];
        const RepGrid Gr = grid[h_rep[g0]];
        const int rrow = (row - Gr.base) / Gr.dims[0];
        const int by = rrow % Gr.dims[1], bz = rrow / Gr.dims[1];
        const int last = h0 + g1 - 1;
        if (tid < nseg) {
            const int dz = tid / side - reach, dy = tid % side - reach;
            const int zz = bz + dz, yy = by + dy;
            int lo = 0, len = 0;
            if (zz >= 0 && zz < Gr.dims[2] && yy >= 0 && yy < Gr.dims[1]) {
                const int rb = Gr.base + (zz * Gr.dims[1] + yy) * Gr.dims[0];
                const int xlo = max(0, xmin - reach);
                const int xhi = (dz == 0 && dy == 0) ? xmax : min(Gr.dims[0] - 1, xmax + reach);
                lo = bin_off[rb + xlo];
                int hi = bin_off[rb + xhi + 1];
                if (dz == 0 && dy == 0) hi = min(hi, last);
                len = max(0, hi - lo);
            }
            seg_lo[tid] = lo;
            seg_pre[tid + 1] = len;
        }
        __syncthreads();
        if (tid == 0) {
            seg_pre[0] = 0;
            for (int k = 0; k < nseg; ++k) seg_pre[k + 1] += seg_pre[k];
        }
        __syncthreads();
        const int S = seg_pre[nseg];
        for (int c0 = 0; c0 < S; c0 += CAP) {
            const int C = min(CAP, S - c0);
            for (int t = tid; t < C; t += 256) {
                const int gidx = c0 + t;
                int sg = 0;
                while (seg_pre[sg + 1] <= gidx) ++sg;
                const int q = seg_lo[sg] + (gidx - seg_pre[sg]);
                c_f[t] = make_float4((float)sorted[q], (float)sorted[stride + q], (float)sorted[2 * stride + q],
                                     (float)sorted[6 * stride + q]);
                c_q[t] = q;
                c_bx[t] = (skey[q] - Gr.base) % Gr.dims[0];
            }
            __syncthreads();
            for (int attempt = 0; attempt < 2; ++attempt) {
                const int BT = attempt == 0 ? (g1 - g0) : BATCH;
                bool overflow = false;
                for (int b0 = g0; b0 < g1; b0 += BT) {
                    const int b1 = min(b0 + BT, g1);
                    if (tid == 0) { s_qn = 0; s_ovf = 0; }
                    __syncthreads();
                    const bool in_batch = lane >= b0 && lane < b1;
                    for (int t = wave; t < C; t += 4) {
                        const float4 cf = c_f[t];
                        const int cbx = c_bx[t];
                        const int cq = c_q[t];
                        const bool valid = in_batch && abs(my_bx - cbx) <= reach && cq < h0 + lane;
                        // FP32 pre-screen; its rounding error is far below the margin in hf.w
                        const float fx = hf.x - cf.x, fy = hf.y - cf.y, fz = hf.z - cf.z;
                        const float fb = hf.w + cf.w;
                        const bool pass = valid && (fx * fx + fy * fy + fz * fz <= fb * fb);
                        screened += valid ? 1 : 0;
                        const unsigned long long mask = __ballot(pass);
                        if (mask) {
                            const int cnt = __popcll(mask);
                            int base = 0;
                            if (lane == 0) base = atomicAdd(&s_qn, cnt);
                            base = __shfl(base, 0);
                            if (pass) {
                                const int below = __builtin_amdgcn_mbcnt_hi(
                                    (unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0));
                                const int pos = base + below;
                                if (pos < QCAP) queue[pos] = (uint16_t)(((lane - b0) << 10) | t);
                                else s_ovf = 1;
                            }
                        }
                    }
                    __syncthreads();
                    if (s_ovf) { overflow = true; break; }
                    const int qn = s_qn;
                    for (int e = tid; e < qn; e += 256) {
                        const int ent = queue[e];
                        const int h = b0 + (ent >> 10), t = ent & 1023;
                        Seg ch, cc;
                        ch.r[0] = h_x[h]; ch.r[1] = h_y[h]; ch.r[2] = h_z[h];
                        ch.n[0] = h_nx[h]; ch.n[1] = h_ny[h]; ch.n[2] = h_nz[h]; ch.hl = h_hl[h];
                        const int cq = c_q[t];
                        cc = load_sorted(sorted, stride, cq);
                        const int cid = sid[cq];
                        const bool home_is_i = h_id[h] > cid;
                        const int i = home_is_i ? h_id[h] : cid;
                        const int j = home_is_i ? cid : h_id[h];
                        const double dx = ch.r[0] - cc.r[0], dy = ch.r[1] - cc.r[1], dz = ch.r[2] - cc.r[2];
                        const double r2 = (dx * dx + dy * dy) + dz * dz;
                        const double bound = (thr + ch.hl + cc.hl) * (1 + 1e-9) + 1e-9;
                        if (!(r2 <= bound * bound && sqrt(r2) < pre)) continue;
                        double d[3], sv, tv;
                        if (home_is_i) dist_between_cells(ch, cc, d, sv, tv);
                        else dist_between_cells(cc, ch, d, sv, tv);
                        if (norm3(d[0], d[1], d[2]) < thr) {
                            const int slot = (int)(atomicAdd(&cnt64[i], 1ull << 32) >> 32);
                            const int slot2 = (int)(atomicAdd(&cnt64[j], 1ull) & 0xffffffffull);
                            if (slot < slot_cap) {
                                const int64_t e2 = (int64_t)slot * stride + i;
                                slot_j[e2] = j;
                                double2* r = reinterpret_cast<double2*>(slot_rec + 8 * e2);
                                r[0] = make_double2(d[0], d[1]);
                                r[1] = make_double2(d[2], sv);
                                r[2] = make_double2(tv, __longlong_as_double((long long)slot2));
                            }
                            if (slot2 < slot_cap) jslot[(int64_t)slot2 * stride + j] = i;
                        }
                    }
                    if (tid == 0) tested += (unsigned long long)qn;
                    __syncthreads();
                }
                if (!overflow) break;
                __syncthreads();
            }
        }
        g0 = g1;
    }
    atomicAdd(&s_screened, screened);
    __syncthreads();
    if (tid == 0) {
        count_add(counters, 0, s_screened);
        count_add(counters, 1, tested);
    }
}

struct ForceParams {
    double R, Rcell, E0, Req, max_overlap;
    double pref0, pref1, pref2;
    int adhesion;
    int jkr_mode;
    int colidx_gamma;
    int pad;
    SurrogateEnvelope env;   // rows inside it take the learned radius (jkr_mode surrogate)
};

// The main-term magnitude of the adhesive overload for an isotropic pair
// (gamma = min of the two cells'): JKR term1 - term2 at the overlap capped at
// the shell thickness, the contact radius from the exact FP64 solve, a table
// or the learned radius (rows inside its envelope); Hertz (EPS-capped) when
// gamma is 0.
__device__ __forceinline__ void isotropic_magnitudes(const ForceParams& fp, double overlap, double gamma,
                                                     const double* __restrict__ tab_ov,
                                                     const double* __restrict__ tab_g,
                                                     const double* __restrict__ tab_r, int tab_nd, int tab_ng,
                                                     const double* __restrict__ sur_radius, int64_t k,
                                                     int* __restrict__ status, double& mm, double& jkr)
{
    if (gamma > 0) {
        double delta = overlap;
        if (fp.max_overlap >= 0 && delta > fp.max_overlap) delta = fp.max_overlap;
        double radius;
        if (fp.jkr_mode == BFM_JKR_TABLE_1D)
            radius = table_radius_1d(tab_ov, tab_r, tab_nd, delta);
        else if (fp.jkr_mode == BFM_JKR_TABLE_2D)
            radius = table_radius_2d(tab_ov, tab_nd, tab_g, tab_ng, tab_r, delta, gamma);
        else if (fp.jkr_mode == BFM_JKR_SURROGATE && sur_radius && sur_inside(delta, gamma, fp.env))
            radius = sur_radius[k];
        else
            radius = jkr_radius_newton(delta, fp.Req, fp.E0, gamma, 0.0, 200, nullptr);
        if (!(radius > 0)) atomicOr(status, ST_JKR_ROOT);
        mm = jkr_magnitude(radius, fp.Req, fp.E0, gamma);
        jkr = 1.0;
    } else {
        if (overlap <= fp.max_overlap) mm = fp.pref0 * pow_half<3>(overlap);
        else mm = fp.pref2;
    }
}

// Main-term, hard-core and JKR-radius values of an isotropic row (i, j) with
// centerline vector d: Hertz with the EPS cap, or the adhesive overload's
// JKR main term plus the separate hard-core term (mechanics.hpp:1209-1350,
// 1390-1659).
__device__ __forceinline__ void isotropic_vals(const ForceParams& fp, const double* __restrict__ cells, int ld,
                                               int i, int j, double dx, double dy, double dz,
                                               const double* __restrict__ tab_ov, const double* __restrict__ tab_g,
                                               const double* __restrict__ tab_r, int tab_nd, int tab_ng,
                                               int* __restrict__ status, double& mm, double& mh, double& jkr)
{
    const double dist = norm3(dx, dy, dz);
    const double overlap = 2 * fp.R - dist;
    mm = 0;
    mh = 0;
    jkr = 0;
    if (overlap > 0) {
        if (!fp.adhesion) {
            if (overlap <= fp.max_overlap) {
                mm = fp.pref0 * pow_half<3>(overlap);
            } else {
                mm = fp.pref2;
                mm += fp.pref1 * pow_half<3>(overlap - fp.max_overlap);
            }
        } else {
            const double gamma = dmin(cellv(cells, ld, i, fp.colidx_gamma), cellv(cells, ld, j, fp.colidx_gamma));
            isotropic_magnitudes(fp, overlap, gamma, tab_ov, tab_g, tab_r, tab_nd, tab_ng, nullptr, 0, status, mm,
                                 jkr);
            if (overlap > fp.max_overlap)
                mh = fp.pref1 * pow_half<3>(overlap - fp.max_overlap);
        }
    }
}

// The 64-byte record of one row for an isotropic interaction ({dx, dy, dz,
// s_i, s_j, main, hard, jkr}, as k_pair_forces writes it).
__device__ __forceinline__ void isotropic_record(const ForceParams& fp, const double* __restrict__ cells, int ld,
                                                 int i, int j, double dx, double dy, double dz, double si,
                                                 double sj, const double* __restrict__ tab_ov,
                                                 const double* __restrict__ tab_g,
                                                 const double* __restrict__ tab_r, int tab_nd, int tab_ng,
                                                 int* __restrict__ status, double* __restrict__ rec)
{
    double mm, mh, jkr;
    isotropic_vals(fp, cells, ld, i, j, dx, dy, dz, tab_ov, tab_g, tab_r, tab_nd, tab_ng, status, mm, mh, jkr);
    double2* r = reinterpret_cast<double2*>(rec);
    r[0] = make_double2(dx, dy);
    r[1] = make_double2(dz, si);
    r[2] = make_double2(sj, mm);
    r[3] = make_double2(mh, jkr);
}

// Packs the slots into the reference's row order and builds the per-cell
// force lists. Rows of cell i start at hi(off64[i]) and are sorted by
// partner j (rank by comparison). The force list of cell c (starting at
// hi(off64[c]) + lo(off64[c])) holds the rows where c is i (ascending j, i.e.
// ascending row) followed by the rows where c is j (ascending i), i.e. all
// rows touching c in row order -- the order the reference's sequential pair
// loop adds them (mechanics.hpp:1438-1647); bit 31 marks the j role.
// L lanes per cell: lane k of a cell's group holds the cell's entries
// k, k + L, ...; an entry's rank among the cell's partners comes from
// width-L shuffles, so the work per cell is O(entries) per lane and no wave
// waits for its most connected cell's O(entries^2) loop.
// Pass 1 (per cell i): rows, i-role entries, and the row index of each pair
// stored at the pair's slot in j's partner list (jrow). FORCES: also the
// row's force record of an isotropic interaction (Hertz, or JKR with the
// exact radius or a table), the record k_pair_forces would write, while the
// row's geometry is in registers.
struct FusedForces {
    ForceParams fp;
    const double* cells;
    int ld;
    const double* tab_ov;
    const double* tab_g;
    const double* tab_r;
    int tab_nd, tab_ng;
    double* rec;
};

// The 64-byte segment record {r, n, half_l, 0} of sorted position q.
__device__ __forceinline__ Seg load_geom(const double* __restrict__ g, int q)
{
    const double2* p = reinterpret_cast<const double2*>(g + 8 * (int64_t)q);
    const double2 a = p[0], b = p[1], c = p[2], d = p[3];
    Seg s;
    s.r[0] = a.x; s.r[1] = a.y; s.r[2] = b.x;
    s.n[0] = b.y; s.n[1] = c.x; s.n[2] = c.y;
    s.hl = d.x;
    return s;
}

// Number of set bits of m in the lanes below this one.
__device__ __forceinline__ int lanes_below(uint64_t m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Squared distance from the point p to the segment c +- [0, c.w] n (FP32).
__device__ __forceinline__ float point_seg_dist2(float4 p, float4 c, float4 n)
{
    const float ux = p.x - c.x, uy = p.y - c.y, uz = p.z - c.z;
    const float t = fminf(fmaxf(ux * n.x + uy * n.y + uz * n.z, -c.w), c.w);
    const float wx = ux - t * n.x, wy = uy - t * n.y, wz = uz - t * n.z;
    return wx * wx + wy * wy + wz * wz;
}

// Second FP32 screen of a pair that passed the centre-distance screen: the
// segment-segment distance is at least dist(centre of one, segment of the
// other) minus the first's half length (every point of a segment lies within
// its half length of its centre), so a pair is kept only if both such lower
// bounds are at most ft = thr + margin. The FP32 rounding of the bounds
// (~1e-5 um at coordinates of a few hundred um) is far below the margin, and
// the reference's |d| joins two points of the segments, so it is never below
// the true segment distance: no pair with |d| < thr is dropped.
__device__ __forceinline__ bool seg_bounds_pass(float4 af, float4 an, float4 bf, float4 bn, float ft)
{
    const float ra = ft + bf.w, rb = ft + af.w;
    return point_seg_dist2(bf, af, an) <= ra * ra && point_seg_dist2(af, bf, bn) <= rb * rb;
}

// Lane-per-home candidate search, without work-group barriers: each wavefront
// owns 64 consecutive bin-ordered homes, one per lane, and every lane walks its
// own home's half stencil -- the stencil rows (dz, dy) before (0, 0) over bins
// [bx - reach, bx + reach] and the home row over [bx - reach, bx] up to the
// home's sorted position -- reading the candidates' FP32 (x, y, z, half length)
// from a packed copy in sorted order (lanes of one bin read the same lines).
// Survivors of the centre-distance screen go to a wavefront-private LDS queue
// (ballot + mbcnt, no atomics); 64 at a time they take the point-segment
// screen, whose survivors go to a second queue that is drained 64 at a time
// through the FP64 segment test, one pair per lane. The pair rule is that of
// k_candidates_lds; of a kept pair's two list entries, the partner's takes
// its slot through a device-scope atomic on the partner's counter (cnt64,
// slots from 0 up) and the home's through an LDS counter of the wavefront
// (slots from slot_cap - 1 down; the counts go to own64 at the end), which
// halves the device-scope atomics (see slot_of). Sorted positions must be
// below 2^26 (queue entries carry the home lane in the top six bits).
template <int UNROLL>
__global__ void __launch_bounds__(256)
k_candidates_lane(int n, const Plan* __restrict__ plan, const double* __restrict__ sgeom, int64_t stride,
                  const float4* __restrict__ sortf, const float4* __restrict__ sortn,
                  const int* __restrict__ sid, const int* __restrict__ skey,
                  const int* __restrict__ rep, const RepGrid* __restrict__ grid,
                  const int* __restrict__ bin_off, int slot_cap, unsigned long long* __restrict__ cnt64,
                  unsigned long long* __restrict__ own64, int* __restrict__ slot_j, double* __restrict__ slot_rec,
                  int* __restrict__ jslot, int* __restrict__ jslotk, unsigned long long* __restrict__ counters)
{
    constexpr int WAVES = 4, QA = 64 + 64 * UNROLL, QB = 128, QBITS = 26;
    constexpr unsigned QMASK = (1u << QBITS) - 1u;
    __shared__ unsigned s_qa[WAVES][QA];
    __shared__ unsigned s_qb[WAVES][QB];
    __shared__ float4 s_hf[WAVES][64], s_hn[WAVES][64];
    __shared__ int s_hid[WAVES][64];
    __shared__ int s_own[WAVES][64][2];   // the homes' own rows as i and as j (LDS atomics)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int hbase = (xcd_block(blockIdx.x, gridDim.x) * WAVES + wave) * 64;
    if (hbase >= n) return;   // the whole wavefront
    const Plan P = *plan;
    const double thr = P.thr, pre = P.pre;
    const float ft = (float)(thr + P.fmargin);
    const int reach = P.reach, side = 2 * reach + 1, nseg = reach * side + reach + 1;
    unsigned* qa = s_qa[wave];
    unsigned* qb = s_qb[wave];

    const int hs = hbase + lane;
    const bool hv = hs < n;
    float4 hf = make_float4(0.f, 0.f, 0.f, 0.f), hn = hf;
    int hid = 0, bx = 0, by = 0, bz = 0, gbase = 0, gd0 = 1, gd1 = 1, gd2 = 1;
    if (hv) {
        hf = sortf[hs];
        hn = sortn[hs];
        hid = sid[hs];
        const int key = skey[hs];
        const int r = rep ? rep[hid] : 0;
        gbase = grid[r].base;
        gd0 = grid[r].dims[0];
        gd1 = grid[r].dims[1];
        gd2 = grid[r].dims[2];
        const int local = key - gbase;
        bx = local % gd0;
        const int rr = local / gd0;
        by = rr % gd1;
        bz = rr / gd1;
    }
    s_hf[wave][lane] = hf;
    s_hn[wave][lane] = hn;
    s_hid[wave][lane] = hid;
    s_own[wave][lane][0] = s_own[wave][lane][1] = 0;
    const float hb = ft + hf.w;   // thr + margin + the home's half length
    unsigned long long screened = 0, tested = 0;
    int na = 0, nb = 0;           // queue fills, uniform across the wavefront

    // FP64 test of queue-B entries, 64 per round: all of them (all) or whole
    // rounds, the rest moved to the front
    auto drain_b = [&](bool all) __attribute__((always_inline)) {
        wave_sync();
        int e0 = 0;
        for (; e0 + 64 <= nb || (all && e0 < nb); e0 += 64) {
            const int e = e0 + lane;
            if (e < nb) {
                const unsigned ent = qb[e];
                const int h = (int)(ent >> QBITS), q = (int)(ent & QMASK);
                const Seg ch = load_geom(sgeom, hbase + h);
                const Seg cc = load_geom(sgeom, q);
                const int hid_h = s_hid[wave][h], cid = sid[q];
                const bool home_is_i = hid_h > cid;
                const int i = home_is_i ? hid_h : cid;
                const int j = home_is_i ? cid : hid_h;
                const double dx = ch.r[0] - cc.r[0], dy = ch.r[1] - cc.r[1], dz = ch.r[2] - cc.r[2];
                const double r2 = (dx * dx + dy * dy) + dz * dz;
                const double bound = (thr + ch.hl + cc.hl) * (1 + 1e-9) + 1e-9;
                if (r2 <= bound * bound && sqrt(r2) < pre) {
                    // one inlined segment test for every lane: the operands are
                    // selected per lane (cell i first) instead of two branches
                    Seg c1, c2;
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        c1.r[x] = home_is_i ? ch.r[x] : cc.r[x];
                        c2.r[x] = home_is_i ? cc.r[x] : ch.r[x];
                        c1.n[x] = home_is_i ? ch.n[x] : cc.n[x];
                        c2.n[x] = home_is_i ? cc.n[x] : ch.n[x];
                    }
                    c1.hl = home_is_i ? ch.hl : cc.hl;
                    c2.hl = home_is_i ? cc.hl : ch.hl;
                    double d[3], sv, tv;
                    dist_between_cells(c1, c2, d, sv, tv);
                    if (norm3(d[0], d[1], d[2]) < thr) {
                        // the home's entry from the top of its list through an LDS
                        // counter of its wavefront, the partner's from the bottom
                        // through a device-scope atomic
                        int slot, slot2;
                        if (home_is_i) {
                            slot = slot_cap - 1 - atomicAdd(&s_own[wave][h][0], 1);
                            slot2 = (int)(atomicAdd(&cnt64[j], 1ull) & 0xffffffffull);
                        } else {
                            slot2 = slot_cap - 1 - atomicAdd(&s_own[wave][h][1], 1);
                            slot = (int)(atomicAdd(&cnt64[i], 1ull << 32) >> 32);
                        }
                        if (slot >= 0 && slot < slot_cap) {
                            const int64_t es = (int64_t)slot * stride + i;
                            slot_j[es] = j;
                            double2* rr = reinterpret_cast<double2*>(slot_rec + 8 * es);
                            rr[0] = make_double2(d[0], d[1]);
                            rr[1] = make_double2(d[2], sv);
                            rr[2] = make_double2(tv, __longlong_as_double((long long)slot2));
                        }
                        if (slot2 >= 0 && slot2 < slot_cap) {
                            jslot[(int64_t)slot2 * stride + j] = i;
                            jslotk[(int64_t)slot2 * stride + j] = slot;   // where j finds the row's record
                        }
                    }
                }
            }
            tested += (unsigned long long)min(64, nb - e0);
        }
        const int rem = nb - e0;   // < 64; 0 when all
        wave_sync();
        const unsigned v = lane < rem ? qb[e0 + lane] : 0u;
        wave_sync();
        if (lane < rem) qb[lane] = v;
        nb = max(rem, 0);
        wave_sync();
    };
    // point-segment screen of queue-A entries, 64 per round, survivors to queue B
    auto drain_a = [&](bool all) __attribute__((always_inline)) {
        wave_sync();
        int e0 = 0;
        for (; e0 + 64 <= na || (all && e0 < na); e0 += 64) {
            const int e = e0 + lane;
            bool pass = false;
            unsigned ent = 0;
            if (e < na) {
                ent = qa[e];
                const int h = (int)(ent >> QBITS), q = (int)(ent & QMASK);
                pass = seg_bounds_pass(s_hf[wave][h], s_hn[wave][h], sortf[q], sortn[q], ft);
            }
            const uint64_t m = __ballot(pass);
            if (pass) qb[nb + lanes_below(m)] = ent;
            nb += __popcll(m);
            if (nb >= 64) drain_b(false);
        }
        const int rem = na - e0;
        wave_sync();
        const unsigned v = lane < rem ? qa[e0 + lane] : 0u;
        wave_sync();
        if (lane < rem) qa[lane] = v;
        na = max(rem, 0);
        wave_sync();
    };

    // bounds of stencil row k as a run of sorted positions (len 0 outside the grid)
    auto seg_bounds = [&](int k, int& lo, int& len) __attribute__((always_inline)) {
        const int dz = k / side - reach, dy = k % side - reach;
        lo = 0;
        len = 0;
        if (hv) {
            const int zz = bz + dz, yy = by + dy;
            if (zz >= 0 && zz < gd2 && yy >= 0 && yy < gd1) {
                const int rb = gbase + (zz * gd1 + yy) * gd0;
                const bool home_row = dz == 0 && dy == 0;
                const int xlo = max(0, bx - reach);
                const int xhi = home_row ? bx : min(gd0 - 1, bx + reach);
                lo = bin_off[rb + xlo];
                int hi = bin_off[rb + xhi + 1];
                if (home_row) hi = min(hi, hs);   // candidates before the home in bin order
                len = max(0, hi - lo);
            }
        }
    };
    int lo, len;
    seg_bounds(0, lo, len);
    for (int k = 0; k < nseg; ++k) {
        int lo_next = 0, len_next = 0;   // the next row's bounds, loaded ahead
        if (k + 1 < nseg) seg_bounds(k + 1, lo_next, len_next);
        int maxlen = len;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) maxlen = max(maxlen, __shfl_xor(maxlen, off));
        for (int i0 = 0; i0 < maxlen; i0 += UNROLL) {
            // loads past the lane's run read a valid position and are masked below
            float4 cf[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) cf[u] = sortf[min(lo + i0 + u, n - 1)];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                const float fx = hf.x - cf[u].x, fy = hf.y - cf[u].y, fz = hf.z - cf[u].z;
                const float fb = hb + cf[u].w;
                const bool pass = (i0 + u < len) & (fx * fx + fy * fy + fz * fz <= fb * fb);
                const uint64_t m = __ballot(pass);
                if (m) {
                    if (pass) qa[na + lanes_below(m)] = ((unsigned)lane << QBITS) | (unsigned)(lo + i0 + u);
                    na += __popcll(m);
                }
            }
            screened += (unsigned long long)max(0, min(UNROLL, len - i0));
            if (na >= 64) drain_a(false);
        }
        lo = lo_next;
        len = len_next;
    }
    drain_a(true);
    drain_b(true);
    wave_sync();
    if (hv)
        own64[hid] = ((unsigned long long)(unsigned)s_own[wave][lane][0] << 32) | (unsigned)s_own[wave][lane][1];
    for (int off = 32; off > 0; off >>= 1) screened += __shfl_xor(screened, off);
    if (lane == 0) {
        count_add(counters, 0, screened);
        count_add(counters, 1, tested);
    }
}


// A cell's per-cell list, one role: the f entries other cells' searches
// added fill slots upwards from 0
// (device-scope atomics on the cell's counter), those of the cell's own
// search fill them downwards from slot_cap - 1 (the lane-per-home search;
// none for the other variants). Their counts are cnt64 and own64 (as i in
// the high, as j in the low half); a cell with more than slot_cap entries in
// a role is reported and skipped (its slots may overlap).
struct SlotCounts {
    int fi, fj, ci, cj;   // entries added by other searches, entries in all (as i, as j)
    bool over;
};
__device__ __forceinline__ SlotCounts slot_counts(unsigned long long cv, unsigned long long ow, int slot_cap)
{
    SlotCounts k;
    k.fi = (int)(cv >> 32);
    k.fj = (int)(cv & 0xffffffffull);
    k.ci = k.fi + (int)(ow >> 32);
    k.cj = k.fj + (int)(ow & 0xffffffffull);
    k.over = k.ci > slot_cap || k.cj > slot_cap;
    if (k.over) k.fi = k.fj = k.ci = k.cj = 0;
    return k;
}

// Slot of entry e (0 <= e < count) of one role, f of them from the bottom.
__device__ __forceinline__ int slot_of(int e, int f, int slot_cap) { return e < f ? e : slot_cap - 1 - (e - f); }

// out = a + b over n entries (per-cell row counts: the atomically taken and
// the own-search ones, each packed as i << 32 | j).
__global__ void k_add_u64(int n, const unsigned long long* __restrict__ a, const unsigned long long* __restrict__ b,
                          unsigned long long* __restrict__ out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = a[i] + b[i];
}

// Lanes per cell of the compaction (k_compact_rows at 200k cells: 35 us with
// eight, 40 us with four).
constexpr int kCompactLanes = 8;

template <bool FORCES, int L>
__global__ void __launch_bounds__(256)
k_compact_rows(int n, int64_t stride, int slot_cap, const unsigned long long* __restrict__ cnt64,
               const unsigned long long* __restrict__ own64,
               const unsigned long long* __restrict__ off64, const int* __restrict__ slot_j,
               const double* __restrict__ slot_rec, int64_t max_pairs, int* __restrict__ pi, int* __restrict__ pj,
               double* __restrict__ pdx, double* __restrict__ pdy, double* __restrict__ pdz,
               double* __restrict__ ps, double* __restrict__ pt, int* __restrict__ csr,
               int* __restrict__ jrow, int* __restrict__ d_m, int* __restrict__ status, FusedForces ff)
{
    const bool fits = (int64_t)(off64[n] >> 32) <= max_pairs;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g / L < n;
         g += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(g / L), k = (int)(g % L);
    int ci = 0, fi = 0;
    int64_t oi = 0, base2 = 0;
    if (c < n) {
        const unsigned long long ov = off64[c];
        const SlotCounts sc = slot_counts(cnt64[c], own64[c], slot_cap);
        if (sc.over && k == 0) atomicOr(status, ST_SLOT_OVERFLOW);
        ci = sc.ci;
        fi = sc.fi;
        oi = (int64_t)(ov >> 32);
        base2 = oi + (int64_t)(ov & 0xffffffffull);
    }
    for (int a = 0; a * L < ci; ++a) {
        const int e = a * L + k;
        const int je = e < ci ? slot_j[(int64_t)slot_of(e, fi, slot_cap) * stride + c] : INT_MAX;
        int rank = 0;
        for (int b = 0; b * L < ci; ++b) {
            const int eb = b * L + k;
            const int jb =
                b == a ? je : (eb < ci ? slot_j[(int64_t)slot_of(eb, fi, slot_cap) * stride + c] : INT_MAX);
#pragma unroll
            for (int f = 0; f < L; ++f) rank += __shfl(jb, f, L) < je;
        }
        if (e < ci && fits) {
            const int64_t se = (int64_t)slot_of(e, fi, slot_cap) * stride + c;
            const int64_t row = oi + rank;
            const double2* r = reinterpret_cast<const double2*>(slot_rec + 8 * se);
            const double2 r0 = r[0], r1 = r[1], r2 = r[2];
            pi[row] = c; pj[row] = je;
            pdx[row] = r0.x; pdy[row] = r0.y; pdz[row] = r1.x;
            ps[row] = r1.y; pt[row] = r2.x;
            csr[base2 + rank] = (int)row;
            if constexpr (FORCES)
                isotropic_record(ff.fp, ff.cells, ff.ld, c, je, r0.x, r0.y, r1.x, r1.y, r2.x, ff.tab_ov, ff.tab_g,
                                 ff.tab_r, ff.tab_nd, ff.tab_ng, status, ff.rec + 8 * row);
            const int s2 = (int)__double_as_longlong(r2.y);
            if (s2 >= 0 && s2 < slot_cap) jrow[(int64_t)s2 * stride + je] = (int)row;
        }
    }
    if (c == n - 1 && k == 0) {
        int64_t m = (int64_t)(off64[n] >> 32);
        if (m > max_pairs) { atomicOr(status, ST_PAIR_OVERFLOW); m = max_pairs; }
        *d_m = (int)m;
    }
    }
}

// Pass 2 (per cell c, L lanes): j-role entries sorted by partner i.
template <int L>
__global__ void __launch_bounds__(256)
k_compact_jrole(int n, int64_t stride, int slot_cap, const unsigned long long* __restrict__ cnt64,
                const unsigned long long* __restrict__ own64,
                const unsigned long long* __restrict__ off64, const int* __restrict__ jslot,
                const int* __restrict__ jrow, int64_t max_pairs, int* __restrict__ csr)
{
    if ((int64_t)(off64[n] >> 32) > max_pairs) return;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g / L < n;
         g += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(g / L), k = (int)(g % L);
    int ci = 0, cj = 0, fj = 0;
    int64_t base2 = 0;
    if (c < n) {
        const unsigned long long ov = off64[c];
        const SlotCounts sc = slot_counts(cnt64[c], own64[c], slot_cap);
        ci = sc.ci;
        cj = sc.cj;
        fj = sc.fj;
        base2 = (int64_t)(ov >> 32) + (int64_t)(ov & 0xffffffffull);
    }
    for (int a = 0; a * L < cj; ++a) {
        const int e = a * L + k;
        const int ie = e < cj ? jslot[(int64_t)slot_of(e, fj, slot_cap) * stride + c] : INT_MAX;
        int rank = 0;
        for (int b = 0; b * L < cj; ++b) {
            const int eb = b * L + k;
            const int ib =
                b == a ? ie : (eb < cj ? jslot[(int64_t)slot_of(eb, fj, slot_cap) * stride + c] : INT_MAX);
#pragma unroll
            for (int f = 0; f < L; ++f) rank += __shfl(ib, f, L) < ie;
        }
        if (e < cj)
            csr[base2 + ci + rank] = jrow[(int64_t)slot_of(e, fj, slot_cap) * stride + c] | (int)0x80000000u;
    }
    }
}

// m x 7 column-major doubles <-> SoA pair list.
__global__ void k_import_nb(const double* __restrict__ nb, int64_t m, int64_t ld,
                            int* __restrict__ pi, int* __restrict__ pj, double* __restrict__ pdx,
                            double* __restrict__ pdy, double* __restrict__ pdz,
                            double* __restrict__ ps, double* __restrict__ pt)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < m;
         k += (int64_t)gridDim.x * blockDim.x) {
        pi[k] = (int)nb[k]; pj[k] = (int)nb[ld + k];
        pdx[k] = nb[2 * ld + k]; pdy[k] = nb[3 * ld + k]; pdz[k] = nb[4 * ld + k];
        ps[k] = nb[5 * ld + k]; pt[k] = nb[6 * ld + k];
    }
}

__global__ void k_export_nb(const int* __restrict__ d_m, double* __restrict__ nb, int64_t ld,
                            const int* __restrict__ pi, const int* __restrict__ pj,
                            const double* __restrict__ pdx, const double* __restrict__ pdy,
                            const double* __restrict__ pdz, const double* __restrict__ ps,
                            const double* __restrict__ pt)
{
    const int64_t m = *d_m;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < m;
         k += (int64_t)gridDim.x * blockDim.x) {
        nb[k] = pi[k]; nb[ld + k] = pj[k];
        nb[2 * ld + k] = pdx[k]; nb[3 * ld + k] = pdy[k]; nb[4 * ld + k] = pdz[k];
        nb[5 * ld + k] = ps[k]; nb[6 * ld + k] = pt[k];
    }
}

__device__ __forceinline__ Seg load_cell_seg(const double* __restrict__ cells, int ld, int i)
{
    Seg c;
    c.r[0] = cellv(cells, ld, i, C_RX); c.r[1] = cellv(cells, ld, i, C_RY); c.r[2] = cellv(cells, ld, i, C_RZ);
    c.n[0] = cellv(cells, ld, i, C_NX); c.n[1] = cellv(cells, ld, i, C_NY); c.n[2] = cellv(cells, ld, i, C_NZ);
    c.hl = cellv(cells, ld, i, C_HALF_L);
    return c;
}

// The segment of every cell (r, n, half_l) as one 64-byte record, so the
// distance update reads each partner with one cache line instead of seven
// column reads.
__global__ void k_pack_segs(int n, const double* __restrict__ cells, int ld, double* __restrict__ geom)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        double2* g = reinterpret_cast<double2*>(geom + (int64_t)c * 8);
        g[0] = make_double2(cellv(cells, ld, c, C_RX), cellv(cells, ld, c, C_RY));
        g[1] = make_double2(cellv(cells, ld, c, C_RZ), cellv(cells, ld, c, C_NX));
        g[2] = make_double2(cellv(cells, ld, c, C_NY), cellv(cells, ld, c, C_NZ));
        g[3] = make_double2(cellv(cells, ld, c, C_HALF_L), 0.0);
    }
}


// updateNeighborDistances (mechanics.hpp:604-648) on packed segments.
__global__ void k_update_distances_packed(const int* __restrict__ d_m, const double* __restrict__ geom, int n,
                                          const int* __restrict__ pi, const int* __restrict__ pj,
                                          double* __restrict__ pdx, double* __restrict__ pdy,
                                          double* __restrict__ pdz, double* __restrict__ ps,
                                          double* __restrict__ pt)
{
    const int m = *d_m;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        // rows of a build that overflowed its capacity (reported by status) may be stale
        if ((unsigned)pi[k] >= (unsigned)n || (unsigned)pj[k] >= (unsigned)n) continue;
        const Seg ci = load_packed_seg(geom, pi[k]);
        const Seg cj = load_packed_seg(geom, pj[k]);
        double d[3], s, t;
        dist_between_cells(ci, cj, d, s, t);
        pdx[k] = d[0]; pdy[k] = d[1]; pdz[k] = d[2]; ps[k] = s; pt[k] = t;
    }
}

// Counts rows per cell of one index column (for transposes).
__global__ void k_count_key(const int* __restrict__ d_m, const int* __restrict__ key, int* __restrict__ cnt)
{
    const int m = *d_m;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x)
        atomicAdd(&cnt[key[k]], 1);
}

__global__ void k_fill_perm(const int* __restrict__ d_m, const int* __restrict__ key,
                            const int* __restrict__ off, int* __restrict__ cursor, int* __restrict__ perm)
{
    const int m = *d_m;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        int c = key[k];
        int pos = off[c] + atomicAdd(&cursor[c], 1);
        perm[pos] = k;
    }
}

// Sorts each cell's segment of a permutation ascending (row order).
__global__ void k_sort_segments(int n, const int* __restrict__ off, int* __restrict__ perm)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        int a = off[c], b = off[c + 1];
        for (int x = a + 1; x < b; ++x) {
            int v = perm[x];
            int y = x - 1;
            while (y >= a && perm[y] > v) { perm[y + 1] = perm[y]; --y; }
            perm[y + 1] = v;
        }
    }
}




// Per-pair force magnitudes (the scalar factors of cellCellInteractionForces,
// mechanics.hpp:1209-1350 and :1390-1659), written with the pair geometry as
// one 64-byte record per row: {dx, dy, dz, s_i, s_j, main, hard, 0}. main
// multiplies the unit distance vector of the main term (Hertz, EPS-capped
// Hertz or JKR term1 - term2), hard that of the separately accumulated
// hard-core term (adhesive overload only; 0 otherwise). Rows without overlap
// get 0 and are skipped by the gather. ANISO: the anisotropic JKR tables
// give the JKR force (forcesAnisotropicJKRLagrange, adhesion.hpp:734-1046;
// the grids are staged in LDS, dynamic shared memory of an.total doubles);
// the orientations come from nst (3 x ld_nst, an RK stage's) when given,
// else from the cells.
template <bool ANISO>
__global__ void __launch_bounds__(256)
k_pair_forces(const int* __restrict__ d_m, const double* __restrict__ cells, int ld,
              const int* __restrict__ pi, const int* __restrict__ pj,
              const double* __restrict__ pdx, const double* __restrict__ pdy,
              const double* __restrict__ pdz, ForceParams fp,
              const double* __restrict__ tab_ov, const double* __restrict__ tab_g,
              const double* __restrict__ tab_r, int tab_nd, int tab_ng, AnisoJKR an,
              const double* __restrict__ nst, int ld_nst,
              const double* __restrict__ ps, const double* __restrict__ pt,
              const double* __restrict__ sur_radius, double* __restrict__ rec,
              int* __restrict__ jcnt, int* __restrict__ status, DistRefresh rf)
{
    extern __shared__ double s_grid[];
    if (rf.gate && *rf.gate == 0) return;
    if (ANISO) {
        for (int q = threadIdx.x; q < an.total; q += blockDim.x) s_grid[q] = an.grid[q];
        __syncthreads();
    }
    const int m = *d_m;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        const int i = pi[k], j = pj[k];
        const unsigned lim = nst ? (unsigned)min(ld, ld_nst) : (unsigned)ld;
        if ((unsigned)i >= lim || (unsigned)j >= lim) {
            // a row of a build that overflowed its capacity (reported by status)
            double2* r = reinterpret_cast<double2*>(rec + (int64_t)k * 8);
            r[0] = r[1] = r[2] = r[3] = make_double2(0.0, 0.0);
            continue;
        }
        if (jcnt) atomicAdd(&jcnt[j], 1);
        double d[3], sv, tv;   // updateNeighborDistances of the row first when rf asks
        row_geometry(rf, k, i, j, pdx, pdy, pdz, ps, pt, d, sv, tv);
        const double dx = d[0], dy = d[1], dz = d[2];
        const double dist = norm3(dx, dy, dz);
        const double overlap = 2 * fp.R - dist;
        double mm = 0, mh = 0, jkr = 0;
        if (overlap > 0) {
            if (!fp.adhesion) {
                if (overlap <= fp.max_overlap) {
                    mm = fp.pref0 * pow_half<3>(overlap);
                } else {
                    mm = fp.pref2;
                    mm += fp.pref1 * pow_half<3>(overlap - fp.max_overlap);
                }
            } else {
                const double gamma = dmin(cellv(cells, ld, i, fp.colidx_gamma),
                                          cellv(cells, ld, j, fp.colidx_gamma));
                if (gamma > 0 && ANISO) {
                    double delta = overlap;
                    if (fp.max_overlap >= 0 && delta > fp.max_overlap) delta = fp.max_overlap;
                    const double d[3] = {dx, dy, dz};
                    double n1[3], n2[3];
                    if (nst) {
#pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            n1[q] = nst[(int64_t)q * ld_nst + i];
                            n2[q] = nst[(int64_t)q * ld_nst + j];
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            n1[q] = cellv(cells, ld, i, C_NX + q);
                            n2[q] = cellv(cells, ld, j, C_NX + q);
                        }
                    }
                    const double2 fr = aniso_jkr_lookup(an, s_grid, d, dist, n1, cellv(cells, ld, i, C_HALF_L),
                                                        n2, cellv(cells, ld, j, C_HALF_L), sv, tv, delta,
                                                        gamma);
                    mm = fr.x;
                    jkr = 1.0;
                } else {
                    isotropic_magnitudes(fp, overlap, gamma, tab_ov, tab_g, tab_r, tab_nd, tab_ng, sur_radius, k,
                                         status, mm, jkr);
                }
                if (overlap > fp.max_overlap)
                    mh = fp.pref1 * pow_half<3>(overlap - fp.max_overlap);
            }
        }
        double2* r = reinterpret_cast<double2*>(rec + (int64_t)k * 8);
        r[0] = make_double2(dx, dy);
        r[1] = make_double2(dz, sv);
        r[2] = make_double2(tv, mm);
        r[3] = make_double2(mh, jkr);
    }
}

struct VelParams {
    SurfaceConsts sc;
    double nz_threshold;
    double coeff;
    int no_surface;
    int has_noise;
};

// Adds one row's contribution to dE/dq of one of its cells (row i: +w,
// s_i w; row j: -w, -s_j w; then the hard-core term the same way), exactly as
// the reference's sequential loop adds forces.row(0/1) and forces_hard
// (mechanics.hpp:1589-1591, 1643-1644).
__device__ __forceinline__ void add_row_vals(double2 a, double2 b, double2 c, double2 d, bool role_i, double R,
                                             double acc[6])
{
    const double dx = a.x, dy = a.y, dz = b.x;
    const double dist = norm3(dx, dy, dz);
    const double overlap = 2 * R - dist;
    if (!(overlap > 0)) return;
    const double dn[3] = {dx / dist, dy / dist, dz / dist};
    const double mm = c.y;
    const double st = role_i ? b.y : -c.x;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double w = mm * dn[q];
        acc[q] += role_i ? w : -w;
        acc[q + 3] += st * w;
    }
    const double mh = d.x;
    if (mh != 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double w = mh * dn[q];
            acc[q] += role_i ? w : -w;
            acc[q + 3] += st * w;
        }
    }
}

__device__ __forceinline__ void add_row(const double* __restrict__ rec, int k, bool role_i, double R,
                                        double acc[6])
{
    const double2* r = reinterpret_cast<const double2*>(rec + (int64_t)k * 8);
    add_row_vals(r[0], r[1], r[2], r[3], role_i, R, acc);
}

// Row lists of a cell. CSR form (lists built by the device neighbor search):
// one run of row indices in row order, bit 31 = j role. Merge form (lists
// imported from the caller, in any row order): rows where c is i (ioff,
// optionally through perm_i) and where c is j (joff / perm_j, ascending),
// merged by row index. Either way each component is accumulated in the
// order of the rows, as the reference's sequential loop does
// (mechanics.hpp:1438-1647).
struct RowLists {
    const unsigned long long* off64;
    const int* csr;
    int64_t cap;            // max_pairs: bounds for a list that overflowed (reported by status)
    const int* ioff;
    const int* perm_i;
    const int* joff;
    const int* perm_j;
};

template <bool CSR>
__device__ __forceinline__ void gather_cc(int c, const RowLists& L, const double* __restrict__ rec, double R,
                                          double acc[6])
{
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] = 0;
    if constexpr (CSR) {
        const unsigned long long o0 = L.off64[c], o1 = L.off64[c + 1];
        const int64_t e0 = (int64_t)(o0 >> 32) + (int64_t)(o0 & 0xffffffffull);
        const int64_t e1 = min((int64_t)(o1 >> 32) + (int64_t)(o1 & 0xffffffffull), 2 * L.cap);
        // rows in groups of four: the four entries, then the four records, are
        // loaded before any is used (two round trips per group instead of two
        // per row), and added in row order
        constexpr int G = 4;
        for (int64_t e = e0; e < e1; e += G) {
            int ent[G];
            bool ok[G];
#pragma unroll
            for (int u = 0; u < G; ++u) ent[u] = e + u < e1 ? L.csr[e + u] : 0;
            double2 ra[G], rb[G], rc[G], rd[G];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                // a build whose slots or pair capacity overflowed is redone or
                // reported; until then its entries may be stale: stay in bounds
                const int row = ent[u] & 0x7fffffff;
                ok[u] = e + u < e1 && row < L.cap;
                if (ok[u]) {
                    const double2* r = reinterpret_cast<const double2*>(rec + (int64_t)row * 8);
                    ra[u] = r[0]; rb[u] = r[1]; rc[u] = r[2]; rd[u] = r[3];
                }
            }
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (ok[u]) add_row_vals(ra[u], rb[u], rc[u], rd[u], ent[u] >= 0, R, acc);
        }
    } else {
        int a = L.ioff[c], ae = L.ioff[c + 1], b = L.joff[c], be = L.joff[c + 1];
        while (a < ae || b < be) {
            const int ka = (a < ae) ? (L.perm_i ? L.perm_i[a] : a) : 0x7fffffff;
            const int kb = (b < be) ? L.perm_j[b] : 0x7fffffff;
            const bool role_i = ka < kb;
            if (role_i) ++a; else ++b;
            add_row(rec, role_i ? ka : kb, role_i, R, acc);
        }
    }
}

// Friction of one contact row (cellCellFrictionForces, mechanics.hpp:1031-1175,
// include_constraint = false): the tangential relative velocity at the
// contact from the stored velocities (columns 7-12), in 2-D when the distance
// vector is nearly horizontal and both contact points are within 1.01 R of
// the surface (positions of the stage being evaluated), scaled by
// -eta sqrt(Req overlap) and truncated by Coulomb's law.
struct FricParams {
    double R, Req, E0, coulomb;
    int no_surface;
    int colidx_eta;
};

// zst: stage (rz, nz) per cell, 2 x n (nullptr: the cells' own columns).
__global__ void __launch_bounds__(256)
k_friction_rows(const int* __restrict__ d_m, const double* __restrict__ cells, int ld, int n_zst,
                const double* __restrict__ zst, const int* __restrict__ pi, const int* __restrict__ pj,
                const double* __restrict__ pdx, const double* __restrict__ pdy, const double* __restrict__ pdz,
                const double* __restrict__ ps, const double* __restrict__ pt, FricParams fq,
                double* __restrict__ frec)
{
    const int m = *d_m;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        double2* out = reinterpret_cast<double2*>(frec + (int64_t)k * 4);
        const int i = pi[k], j = pj[k];
        double d[3] = {pdx[k], pdy[k], pdz[k]};
        double dist = norm3(d[0], d[1], d[2]);
        double overlap = 2 * fq.R - dist;
        if ((unsigned)i >= (unsigned)ld || (unsigned)j >= (unsigned)ld || !(overlap > 0)) {
            out[0] = out[1] = make_double2(0.0, 0.0);
            continue;
        }
        const double si = ps[k], sj = pt[k];
        double dn[3] = {d[0] / dist, d[1] / dist, d[2] / dist};
        const double eta = dmin(cellv(cells, ld, i, fq.colidx_eta), cellv(cells, ld, j, fq.colidx_eta));
        bool a2d = false;
        if (!fq.no_surface && fabs(dn[2]) < 1e-3) {
            const double zi = zst ? zst[i] : cellv(cells, ld, i, C_RZ);
            const double nzi = zst ? zst[n_zst + i] : cellv(cells, ld, i, C_NZ);
            const double zj = zst ? zst[j] : cellv(cells, ld, j, C_RZ);
            const double nzj = zst ? zst[n_zst + j] : cellv(cells, ld, j, C_NZ);
            if (zi + si * nzi < 1.01 * fq.R && zj + sj * nzj < 1.01 * fq.R) a2d = true;
        }
        double rv[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double vi = cellv(cells, ld, i, C_DRX + q) + si * cellv(cells, ld, i, C_DNX + q);
            const double vj = cellv(cells, ld, j, C_DRX + q) + sj * cellv(cells, ld, j, C_DNX + q);
            rv[q] = vi - vj;
        }
        if (a2d) {
            d[2] = 0;
            dist = norm3(d[0], d[1], d[2]);
            dn[0] = d[0] / dist; dn[1] = d[1] / dist; dn[2] = d[2] / dist;
            overlap = 2 * fq.R - dist;
            rv[2] = 0;
        }
        const double dot = (rv[0] * dn[0] + rv[1] * dn[1]) + rv[2] * dn[2];
        const double pre = -eta * sqrt(fq.Req * overlap);
        double f[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) f[q] = pre * (rv[q] - dot * dn[q]);
        if (fq.coulomb > 0) {
            const double fn = norm3(f[0], f[1], f[2]);
            const double mx = fq.coulomb * (4. / 3.) * fq.E0 * sqrt(fq.Req) * pow_half<3>(overlap);
            if (fn > mx) {
                const double ratio = mx / fn;
#pragma unroll
                for (int q = 0; q < 3; ++q) f[q] *= ratio;
            }
        }
        out[0] = make_double2(f[0], f[1]);
        out[1] = make_double2(f[2], 1.0);   // .y: contact flag
    }
}

// Visits a cell's rows in row order (CSR or merge form, see RowLists) as
// fn(row, role_i).
template <bool CSR, class F>
__device__ __forceinline__ void for_rows(int c, const RowLists& L, F&& fn)
{
    if constexpr (CSR) {
        const unsigned long long o0 = L.off64[c], o1 = L.off64[c + 1];
        const int64_t e0 = (int64_t)(o0 >> 32) + (int64_t)(o0 & 0xffffffffull);
        const int64_t e1 = min((int64_t)(o1 >> 32) + (int64_t)(o1 & 0xffffffffull), 2 * L.cap);
        for (int64_t e = e0; e < e1; ++e) {
            const int ent = L.csr[e];
            const int row = ent & 0x7fffffff;
            if (row < L.cap) fn(row, ent >= 0);
        }
    } else {
        int a = L.ioff[c], ae = L.ioff[c + 1], b = L.joff[c], be = L.joff[c + 1];
        while (a < ae || b < be) {
            const int ka = (a < ae) ? (L.perm_i ? L.perm_i[a] : a) : 0x7fffffff;
            const int kb = (b < be) ? L.perm_j[b] : 0x7fffffff;
            const bool role_i = ka < kb;
            if (role_i) ++a; else ++b;
            fn(role_i ? ka : kb, role_i);
        }
    }
}

// The friction part of a cell's dE/dq, accumulated in row order as the
// reference's loop does: row i: -f and -s_i f; row j: +f and -(-s_j f).
template <bool CSR>
__device__ __forceinline__ void gather_fric(int c, const RowLists& L, const double* __restrict__ frec,
                                            const double* __restrict__ ps, const double* __restrict__ pt,
                                            double acc[6])
{
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] = 0;
    for_rows<CSR>(c, L, [&](int row, bool role_i) {
        const double2* r = reinterpret_cast<const double2*>(frec + (int64_t)row * 4);
        const double2 a = r[0], b = r[1];
        if (b.y == 0) return;
        const double f[3] = {a.x, a.y, b.x};
        if (role_i) {
            const double si = ps[row];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                acc[q] -= f[q];
                acc[q + 3] -= si * f[q];
            }
        } else {
            const double msj = -pt[row];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                acc[q] += f[q];
                acc[q + 3] -= msj * f[q];
            }
        }
    });
}

// Friction inputs of the per-cell kernels (frec == nullptr: friction off).
struct FricGather {
    const double* frec;
    const double* ps;
    const double* pt;
};

// ---- include_constraint = true (velocity Verlet) ---------------------------

// (s n) x v, the torque form of the constrained forces (mechanics.hpp:1300-1305).
__device__ __forceinline__ void cross_sn(double s, const double nc[3], const double v[3], double out[3])
{
    const double a[3] = {s * nc[0], s * nc[1], s * nc[2]};
    out[0] = a[1] * v[2] - a[2] * v[1];
    out[1] = a[2] * v[0] - a[0] * v[2];
    out[2] = a[0] * v[1] - a[1] * v[0];
}

// One row's constrained contribution to dE/dq of the cell with orientation
// nc: the JKR term (record flag) as s (-(n . w) n + w) / t ((n . w) n - w)
// (adhesion.hpp:225-231, v = -w), the Hertz, EPS-cap and hard-core terms as
// (s n) x (+-w) (mechanics.hpp:1298-1306, 1554-1562, 1608-1616).
__device__ __forceinline__ void add_row_ic(const double* __restrict__ rec, int k, bool role_i, double R,
                                           const double nc[3], double acc[6])
{
    const double2* r = reinterpret_cast<const double2*>(rec + (int64_t)k * 8);
    const double2 a = r[0], b = r[1], c = r[2], d = r[3];
    const double dx = a.x, dy = a.y, dz = b.x;
    const double dist = norm3(dx, dy, dz);
    const double overlap = 2 * R - dist;
    if (!(overlap > 0)) return;
    const double dn[3] = {dx / dist, dy / dist, dz / dist};
    const double s = role_i ? b.y : c.x;
    double w[3], mw[3], t[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        w[q] = c.y * dn[q];
        mw[q] = -w[q];
        acc[q] += role_i ? w[q] : mw[q];
    }
    if (d.y != 0) {
        const double wd = (nc[0] * w[0] + nc[1] * w[1]) + nc[2] * w[2];
#pragma unroll
        for (int q = 0; q < 3; ++q) t[q] = role_i ? s * ((-wd) * nc[q] + w[q]) : s * (wd * nc[q] - w[q]);
    } else {
        cross_sn(s, nc, role_i ? w : mw, t);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) acc[q + 3] += t[q];
    if (d.x != 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            w[q] = d.x * dn[q];
            mw[q] = -w[q];
            acc[q] += role_i ? w[q] : mw[q];
        }
        cross_sn(s, nc, role_i ? w : mw, t);
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[q + 3] += t[q];
    }
}

// Constrained friction of a cell (cellCellFrictionForces with
// include_constraint = true, mechanics.hpp:1152-1160): row i: -f, -(s_i n_i) x f;
// row j: +f, -(s_j n_j) x (-f).
template <bool CSR>
__device__ __forceinline__ void gather_fric_ic(int c, const RowLists& L, const double* __restrict__ frec,
                                               const double* __restrict__ ps, const double* __restrict__ pt,
                                               const double nc[3], double acc[6])
{
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] = 0;
    for_rows<CSR>(c, L, [&](int row, bool role_i) {
        const double2* r = reinterpret_cast<const double2*>(frec + (int64_t)row * 4);
        const double2 a = r[0], b = r[1];
        if (b.y == 0) return;
        const double f[3] = {a.x, a.y, b.x}, mf[3] = {-a.x, -a.y, -b.x};
        double t[3];
        cross_sn(role_i ? ps[row] : pt[row], nc, role_i ? f : mf, t);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            acc[q] += role_i ? mf[q] : f[q];
            acc[q + 3] -= t[q];
        }
    });
}

struct VerletParams {
    VelParams vp;
    double rho, dt;
    double R3;          // R^3 as the host's pow(R, 3)
    int ncols;
};

// The generalized forces of one cell in stepVelocityVerlet (mechanics.hpp:
// 2845-2872): getConservativeForces with include_constraint = true, plus
// noise, minus ambient viscosity (:1785-1806) and minus the cell-surface
// friction (:1822-1900) truncated to cs_coulomb times the norm of the
// constrained surface repulsion.
template <bool CSR>
__global__ void __launch_bounds__(256)
k_verlet_forces(int n, const double* __restrict__ cells, int ld, const double* __restrict__ noise, VerletParams P,
                RowLists L, const double* __restrict__ rec, FricGather fg, double* __restrict__ F)
{
    const VelParams& vp = P.vp;
    const double R = vp.sc.R;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const double nc[3] = {cellv(cells, ld, c, C_NX), cellv(cells, ld, c, C_NY), cellv(cells, ld, c, C_NZ)};
        const double rz = cellv(cells, ld, c, C_RZ), hl = cellv(cells, ld, c, C_HALF_L), l = cellv(cells, ld, c, C_L);
        double cc[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) cc[q] = 0;
        for_rows<CSR>(c, L, [&](int row, bool role_i) { add_row_ic(rec, row, role_i, R, nc, cc); });
        if (fg.frec) {
            double fr[6];
            gather_fric_ic<CSR>(c, L, fg.frec, fg.ps, fg.pt, nc, fr);
#pragma unroll
            for (int q = 0; q < 6; ++q) cc[q] += fr[q];
        }
        const bool a2d = nc[2] < vp.nz_threshold;
        const double ss = a2d ? NAN : (R - rz) / nc[2];
        double rep[6] = {0, 0, 0, 0, 0, 0}, adh[6] = {0, 0, 0, 0, 0, 0};
        if (!vp.no_surface)
            surface_forces_ic(vp.sc, a2d, rz, nc[0], nc[1], nc[2], hl, l, cellv(cells, ld, c, C_SIGMA0), ss, rep,
                              adh);
        double f[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) f[q] = (-cc[q] - rep[q]) - adh[q];
        if (noise) {
#pragma unroll
            for (int q = 0; q < 6; ++q) f[q] += noise[(int64_t)q * n + c];
        }
        const double dr[3] = {cellv(cells, ld, c, C_DRX), cellv(cells, ld, c, C_DRY), cellv(cells, ld, c, C_DRZ)};
        const double dnv[3] = {cellv(cells, ld, c, C_DNX), cellv(cells, ld, c, C_DNY), cellv(cells, ld, c, C_DNZ)};
        const double w[3] = {nc[1] * dnv[2] - nc[2] * dnv[1], nc[2] * dnv[0] - nc[0] * dnv[2],
                             nc[0] * dnv[1] - nc[1] * dnv[0]};
        {   // ambient viscosity, as written in the reference (l^3 and l^9)
            const double eta = cellv(cells, ld, c, C_ETA0);
            const double li = l * l * l;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                f[q] -= eta * li * dr[q];
                f[q + 3] -= eta * li * li * li * w[q] / 12;
            }
        }
        if (!vp.no_surface) {
            double sf[6] = {0, 0, 0, 0, 0, 0};
            const double eta1 = cellv(cells, ld, c, C_ETA1);
            const double qv[3] = {dr[0], dr[1], 0.0};
            const double mv[3] = {w[1] * nc[2] - w[2] * nc[1], w[2] * nc[0] - w[0] * nc[2], 0.0};
            if (a2d) {
                const double phi = R - rz;
                if (phi > 0) {
                    const double area = vp.sc.sqR * sqrt(phi) * l;
                    sf[0] = eta1 * qv[0] * area / R;
                    sf[1] = eta1 * qv[1] * area / R;
                    const double cross_z = nc[0] * mv[1] - nc[1] * mv[0];
                    const double l2 = l * l;
                    sf[5] = eta1 * cross_z * area * l2 / (12 * R);
                }
            } else {
                double ai[3];
                area_integrals(rz, nc[2], R, hl, ss, vp.sc.sqR, ai);
                sf[0] = eta1 * (qv[0] * ai[0] + mv[0] * ai[1]) / R;
                sf[1] = eta1 * (qv[1] * ai[0] + mv[1] * ai[1]) / R;
                const double c1[3] = {nc[1] * qv[2] - nc[2] * qv[1], nc[2] * qv[0] - nc[0] * qv[2],
                                      nc[0] * qv[1] - nc[1] * qv[0]};
                const double c2[3] = {nc[1] * mv[2] - nc[2] * mv[1], nc[2] * mv[0] - nc[0] * mv[2],
                                      nc[0] * mv[1] - nc[1] * mv[0]};
#pragma unroll
                for (int q = 0; q < 3; ++q) sf[3 + q] = eta1 * (c1[q] * ai[1] + c2[q] * ai[2]) / R;
            }
            if (vp.coeff > 0) {
                double fn2 = 0, rn2 = 0;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    fn2 += sf[q] * sf[q];
                    rn2 += rep[q] * rep[q];
                }
                const double fn = sqrt(fn2), mx = vp.coeff * sqrt(rn2);
                if (fn > mx) {
#pragma unroll
                    for (int q = 0; q < 6; ++q) sf[q] *= (mx / fn);
                }
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) f[q] -= sf[q];
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) F[(int64_t)q * n + c] = f[q];
    }
}

// Sets *bad when a position or orientation is not finite.
__global__ void k_nonfinite(int n, const double* __restrict__ cells, int ld, int* __restrict__ bad)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        bool ok = true;
#pragma unroll
        for (int q = 0; q < 6; ++q) ok &= isfinite(cells[(int64_t)(C_RX + q) * ld + c]);
        if (!ok) atomicOr(bad, 1);
    }
}

// Mass and the two moments of inertia of a spherocylinder of length l and
// density rho (mechanics.hpp:2800-2822).
__device__ __forceinline__ void verlet_inertia(double R, double R3, double rho, double l, double& mass,
                                               double& I0, double& I1)
{
    const double volume = kPi * ((4. / 3.) * R3 + R * R * l);
    mass = rho * volume;
    const double diam = 2 * R, d2 = diam * diam, d3 = d2 * diam, d4 = d3 * diam;
    const double l2 = l * l, l3 = l2 * l;
    I0 = d4 * (l / 32 + diam / 60);
    I1 = d2 * (4 * d3 + 15 * l * d2 + 20 * l2 * diam + 10 * l3) / 480;
    I0 *= (rho * kPi);
    I1 *= (rho * kPi);
}

// delta1 = F / m, delta2 = (t . n) n / I0 + (t - (t . n) n) / I1.
__device__ __forceinline__ void verlet_accel(const double* __restrict__ F, int n, int c, const double nc[3],
                                             double mass, double I0, double I1, double d1[3], double d2[3])
{
    const double t[3] = {F[(int64_t)3 * n + c], F[(int64_t)4 * n + c], F[(int64_t)5 * n + c]};
    const double dot = (t[0] * nc[0] + t[1] * nc[1]) + t[2] * nc[2];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double par = dot * nc[q];
        const double perp = t[q] - par;
        d1[q] = F[(int64_t)q * n + c] / mass;
        d2[q] = par / I0;
        d2[q] += perp / I1;
    }
}

__device__ __forceinline__ void normalize_orientation(double nv[3])
{
    const double nrm = norm3(nv[0], nv[1], nv[2]);
    nv[0] /= nrm; nv[1] /= nrm; nv[2] /= nrm;
    if (nv[2] < 0) { nv[0] *= -1; nv[1] *= -1; nv[2] *= -1; }
}

// First half of stepVelocityVerlet (mechanics.hpp:2892-2916): half-step
// velocities and angular velocities, end points advanced, new position,
// orientation (normalised) and stored velocities; the half-step angular
// velocities go to wh (3 x n).
__global__ void __launch_bounds__(256)
k_verlet_half(int n, const double* __restrict__ cells, int ld, const double* __restrict__ F, VerletParams P,
              double R, double* __restrict__ out, int ld_out, double* __restrict__ wh)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        for (int col = 0; col < P.ncols; ++col) out[(int64_t)col * ld_out + c] = cellv(cells, ld, c, col);
        const double nc[3] = {cellv(cells, ld, c, C_NX), cellv(cells, ld, c, C_NY), cellv(cells, ld, c, C_NZ)};
        const double l = cellv(cells, ld, c, C_L), hl = cellv(cells, ld, c, C_HALF_L);
        double mass, I0, I1, d1[3], d2[3];
        verlet_inertia(R, P.R3, P.rho, l, mass, I0, I1);
        verlet_accel(F, n, c, nc, mass, I0, I1, d1, d2);
        const double dnv[3] = {cellv(cells, ld, c, C_DNX), cellv(cells, ld, c, C_DNY), cellv(cells, ld, c, C_DNZ)};
        const double w[3] = {nc[1] * dnv[2] - nc[2] * dnv[1], nc[2] * dnv[0] - nc[0] * dnv[2],
                             nc[0] * dnv[1] - nc[1] * dnv[0]};
        const double h = 0.5 * P.dt;
        double v[3], wi[3], pp[3], qq[3], pq[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            v[q] = cellv(cells, ld, c, C_DRX + q) + h * d1[q];
            wi[q] = w[q] + h * d2[q];
            wh[(int64_t)q * n + c] = wi[q];
            const double r = cellv(cells, ld, c, C_RX + q);
            pp[q] = r - hl * nc[q];
            qq[q] = r + hl * nc[q];
            pq[q] = pp[q] - qq[q];
        }
        const double cr[3] = {(wi[1] * pq[2] - wi[2] * pq[1]) / 2, (wi[2] * pq[0] - wi[0] * pq[2]) / 2,
                              (wi[0] * pq[1] - wi[1] * pq[0]) / 2};
        double nn[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double pn = pp[q] + P.dt * (v[q] + cr[q]);
            const double qn = qq[q] + P.dt * (v[q] - cr[q]);
            out[(int64_t)(C_RX + q) * ld_out + c] = (pn + qn) / 2;
            nn[q] = (qn - pn) / l;
            out[(int64_t)(C_DRX + q) * ld_out + c] = v[q];
        }
        out[(int64_t)C_DNX * ld_out + c] = wi[1] * nc[2] - wi[2] * nc[1];
        out[(int64_t)C_DNY * ld_out + c] = wi[2] * nc[0] - wi[0] * nc[2];
        out[(int64_t)C_DNZ * ld_out + c] = wi[0] * nc[1] - wi[1] * nc[0];
        normalize_orientation(nn);
#pragma unroll
        for (int q = 0; q < 3; ++q) out[(int64_t)(C_NX + q) * ld_out + c] = nn[q];
    }
}

// Second half (mechanics.hpp:2970-2992): velocities and angular velocities
// completed with the forces at the new positions, dn = w x n, orientations
// normalised again.
__global__ void __launch_bounds__(256)
k_verlet_final(int n, double* __restrict__ cells, int ld, const double* __restrict__ F, VerletParams P, double R,
               const double* __restrict__ wh)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        double nc[3] = {cells[(int64_t)C_NX * ld + c], cells[(int64_t)C_NY * ld + c], cells[(int64_t)C_NZ * ld + c]};
        double mass, I0, I1, d1[3], d2[3];
        verlet_inertia(R, P.R3, P.rho, cells[(int64_t)C_L * ld + c], mass, I0, I1);
        verlet_accel(F, n, c, nc, mass, I0, I1, d1, d2);
        const double h = 0.5 * P.dt;
        double wf[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            cells[(int64_t)(C_DRX + q) * ld + c] += h * d1[q];
            wf[q] = wh[(int64_t)q * n + c] + h * d2[q];
        }
        cells[(int64_t)C_DNX * ld + c] = wf[1] * nc[2] - wf[2] * nc[1];
        cells[(int64_t)C_DNY * ld + c] = wf[2] * nc[0] - wf[0] * nc[2];
        cells[(int64_t)C_DNZ * ld + c] = wf[0] * nc[1] - wf[1] * nc[0];
        normalize_orientation(nc);
#pragma unroll
        for (int q = 0; q < 3; ++q) cells[(int64_t)(C_NX + q) * ld + c] = nc[q];
    }
}


__device__ __forceinline__ CellVel load_cellvel(const double* __restrict__ cells, int ld, int i)
{
    CellVel v;
    v.rz = cellv(cells, ld, i, C_RZ);
    v.nx = cellv(cells, ld, i, C_NX); v.ny = cellv(cells, ld, i, C_NY); v.nz = cellv(cells, ld, i, C_NZ);
    v.l = cellv(cells, ld, i, C_L); v.hl = cellv(cells, ld, i, C_HALF_L);
    v.eta0 = cellv(cells, ld, i, C_ETA0); v.eta1 = cellv(cells, ld, i, C_ETA1);
    v.drx = cellv(cells, ld, i, C_DRX); v.dry = cellv(cells, ld, i, C_DRY);
    v.dnx = cellv(cells, ld, i, C_DNX); v.dny = cellv(cells, ld, i, C_DNY);
    return v;
}

// Conservative forces (-F_cc - F_rep - F_adh, mechanics.hpp:1768) + noise and
// the velocity solve for one cell.
__device__ __forceinline__ void cell_velocity(const VelParams& vp, const CellVel& cv, double sigma0,
                                              bool a2d, const double cc[6], bool has_noise,
                                              const double noise6[6], double v[6])
{
    const double ss = a2d ? NAN : (vp.sc.R - cv.rz) / cv.nz;
    double rep2 = 0, rep5 = 0, adh2 = 0, adh5 = 0;
    if (!vp.no_surface)
        surface_forces(vp.sc, a2d, cv.rz, cv.nz, cv.hl, cv.l, sigma0, ss, rep2, rep5, adh2, adh5);
    double f[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) f[q] = -cc[q];
    f[2] = (f[2] - rep2) - adh2;
    f[5] = (f[5] - rep5) - adh5;
    if (has_noise) {
#pragma unroll
        for (int q = 0; q < 6; ++q) f[q] += noise6[q];
    }
    velocity_solve(cv, a2d, ss, vp.sc, vp.coeff, vp.no_surface, f, v);
}

// getVelocities for every cell (mechanics.hpp:2011-2347).
// cc_out receives the conservative cell-cell part only (the RK stages reuse
// it); with friction on, the velocities use cc + friction (mechanics.hpp:1650-1656).
template <bool CSR>
__global__ void __launch_bounds__(256)
k_velocities(int n, const double* __restrict__ cells, int ld, const int* __restrict__ assume_2d,
             const double* __restrict__ noise, VelParams vp, RowLists L, const double* __restrict__ rec,
             FricGather fg, double* __restrict__ vel, double* __restrict__ cc_out, int* __restrict__ status)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        double cc[6];
        gather_cc<CSR>(c, L, rec, vp.sc.R, cc);
        if (cc_out) {
#pragma unroll
            for (int q = 0; q < 6; ++q) cc_out[(int64_t)q * n + c] = cc[q];
        }
        if (!vel) continue;
        if (fg.frec) {
            double fr[6];
            gather_fric<CSR>(c, L, fg.frec, fg.ps, fg.pt, fr);
#pragma unroll
            for (int q = 0; q < 6; ++q) cc[q] += fr[q];
        }
        const CellVel cv = load_cellvel(cells, ld, c);
        const bool a2d = assume_2d ? (assume_2d[c] != 0) : (cv.nz < vp.nz_threshold);
        double nz6[6] = {0, 0, 0, 0, 0, 0};
        if (noise) {
#pragma unroll
            for (int q = 0; q < 6; ++q) nz6[q] = noise[(int64_t)q * n + c];
        }
        double v[6];
        cell_velocity(vp, cv, cellv(cells, ld, c, C_SIGMA0), a2d, cc, noise != nullptr, nz6, v);
        bool bad = false;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            vel[(int64_t)q * n + c] = v[q];
            bad |= !isfinite(v[q]);
        }
        if (bad) atomicOr(status, ST_NAN);
    }
}

// Lanes per cell of k_slot_order (measured on the 200k dome: 1, 2, 4, 8, 16
// lanes give 76, 55, 42, 46, 68 us).
constexpr int kSlotLanes = 4;

// Isotropic row values and row order of the lane-per-home search's per-cell
// lists, L lanes per cell: each i-role entry (i = c, j = its partner) gets
// the row's main and hard-core terms, computed as the compaction computes
// them, in the last two doubles of its record {d, s_i, s_j, partner slot,
// main, hard}; each entry's rank among the cell's partners comes from
// width-L shuffles, and ord_i[rank][c] = its slot (ascending j),
// ord_j[rank][c] = i << 32 | i's slot of the row (ascending i).
template <int L>
__global__ void __launch_bounds__(256)
k_slot_order(int n, FusedForces ff, int64_t stride, int slot_cap, const unsigned long long* __restrict__ cnt64,
             const unsigned long long* __restrict__ own64, const int* __restrict__ slot_j,
             double* __restrict__ slot_rec, const int* __restrict__ jslot, const int* __restrict__ jslotk,
             int* __restrict__ ord_i, unsigned long long* __restrict__ ord_j, int* __restrict__ status)
{
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g / L < n;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(g / L), k = (int)(g % L);
        const SlotCounts sc = slot_counts(cnt64[c], own64[c], slot_cap);   // overflow: k_velocities_slots
        const int ci = sc.ci, cj = sc.cj, fi = sc.fi, fj = sc.fj;
        for (int a = 0; a * L < ci; ++a) {
            const int e = a * L + k;
            const int pe = e < ci ? slot_of(e, fi, slot_cap) : 0;
            const int je = e < ci ? slot_j[(int64_t)pe * stride + c] : INT_MAX;
            int rank = 0;
            for (int b = 0; b * L < ci; ++b) {
                const int eb = b * L + k;
                const int jb =
                    b == a ? je : (eb < ci ? slot_j[(int64_t)slot_of(eb, fi, slot_cap) * stride + c] : INT_MAX);
#pragma unroll
                for (int f = 0; f < L; ++f) rank += __shfl(jb, f, L) < je;
            }
            if (e < ci) {
                const int64_t se = (int64_t)pe * stride + c;
                double2* rr = reinterpret_cast<double2*>(slot_rec + 8 * se);
                const double2 r0 = rr[0], r1 = rr[1];
                double mm, mh, jkr;
                isotropic_vals(ff.fp, ff.cells, ff.ld, c, je, r0.x, r0.y, r1.x, ff.tab_ov, ff.tab_g, ff.tab_r,
                               ff.tab_nd, ff.tab_ng, status, mm, mh, jkr);
                rr[3] = make_double2(mm, mh);
                ord_i[(int64_t)rank * stride + c] = pe;
            }
        }
        for (int a = 0; a * L < cj; ++a) {
            const int e = a * L + k;
            const int pe = e < cj ? slot_of(e, fj, slot_cap) : 0;
            const int ie = e < cj ? jslot[(int64_t)pe * stride + c] : INT_MAX;
            int rank = 0;
            for (int b = 0; b * L < cj; ++b) {
                const int eb = b * L + k;
                const int ib =
                    b == a ? ie : (eb < cj ? jslot[(int64_t)slot_of(eb, fj, slot_cap) * stride + c] : INT_MAX);
#pragma unroll
                for (int f = 0; f < L; ++f) rank += __shfl(ib, f, L) < ie;
            }
            if (e < cj)
                ord_j[(int64_t)rank * stride + c] =
                    ((unsigned long long)(unsigned)ie << 32) | (unsigned)jslotk[(int64_t)pe * stride + c];
        }
    }
}

// One slot-list row {d, s_i, s_j, partner slot, main, hard} added as
// k_velocities adds the compacted row.
__device__ __forceinline__ void slot_row(const double2 (&r)[4], bool role_i, double R, double acc[6])
{
    add_row_vals(r[0], r[1], make_double2(r[2].x, r[3].x), make_double2(r[3].y, 0.0), role_i, R, acc);
}

// getVelocities of one rebuilt stage straight from the lane-per-home search's
// per-cell lists, without compacting them into rows: the entries of cell c
// where c is i (partners j < c: slot_j and the record at c's slots) in
// ascending j, then those where c is j (partners i > c: jslot, the record at
// slot jslotk of i) in ascending i -- the reference's row order for c
// (mechanics.hpp:1438-1647), in the order k_slot_order wrote -- each row's
// values added as k_velocities adds them, so the result is bit-identical to
// the compacted path. Isotropic interactions without friction only.
__global__ void __launch_bounds__(256)
k_velocities_slots(int n, const double* __restrict__ cells, int ld, VelParams vp, int64_t stride,
                   int slot_cap, const unsigned long long* __restrict__ cnt64,
                   const unsigned long long* __restrict__ own64,
                   const double* __restrict__ slot_rec, const int* __restrict__ ord_i,
                   const unsigned long long* __restrict__ ord_j, double* __restrict__ vel, int* __restrict__ status)
{
    constexpr int G = 4;   // rows whose order entries, then records, are loaded before any is added
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const SlotCounts sc = slot_counts(cnt64[c], own64[c], slot_cap);
        if (sc.over) atomicOr(status, ST_SLOT_OVERFLOW);
        const int ci = sc.ci, cj = sc.cj;
        double cc[6] = {0, 0, 0, 0, 0, 0};
        // rows where c is i (ascending partner j), then where c is j
        // (ascending partner i), G at a time
        const int ct = ci + cj;
        for (int r0 = 0; r0 < ct; r0 += G) {
            int64_t rec_at[G];
            bool ok[G];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const int r = r0 + u;
                ok[u] = false;
                rec_at[u] = 0;
                if (r < ci) {
                    rec_at[u] = (int64_t)ord_i[(int64_t)r * stride + c] * stride + c;
                    ok[u] = true;
                } else if (r < ct) {
                    const unsigned long long o = ord_j[(int64_t)(r - ci) * stride + c];
                    const int si = (int)(o & 0xffffffffull);
                    rec_at[u] = (int64_t)si * stride + (int)(o >> 32);
                    ok[u] = si >= 0 && si < slot_cap;   // else i's slots overflowed: reported by i's thread
                }
            }
            double2 ra[G][4];
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (ok[u]) {
                    const double2* rr = reinterpret_cast<const double2*>(slot_rec + 8 * rec_at[u]);
                    ra[u][0] = rr[0]; ra[u][1] = rr[1]; ra[u][2] = rr[2]; ra[u][3] = rr[3];
                }
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (ok[u]) slot_row(ra[u], r0 + u < ci, vp.sc.R, cc);
        }
        const CellVel cv6 = load_cellvel(cells, ld, c);
        const bool a2d = cv6.nz < vp.nz_threshold;
        double nz6[6] = {0, 0, 0, 0, 0, 0};
        double v[6];
        cell_velocity(vp, cv6, cellv(cells, ld, c, C_SIGMA0), a2d, cc, false, nz6, v);
        bool bad = false;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            vel[(int64_t)q * n + c] = v[q];
            bad |= !isfinite(v[q]);
        }
        if (bad) atomicOr(status, ST_NAN);
    }
}

// Conservative forces only (getConservativeForces, include_constraint = false).
template <bool CSR>
__global__ void k_conservative(int n, const double* __restrict__ cells, int ld,
                               const int* __restrict__ assume_2d, VelParams vp,
                               const double* __restrict__ cc, RowLists L, FricGather fg, double* __restrict__ out)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        double fr[6] = {0, 0, 0, 0, 0, 0};
        if (fg.frec) gather_fric<CSR>(c, L, fg.frec, fg.ps, fg.pt, fr);
        const CellVel cv = load_cellvel(cells, ld, c);
        const bool a2d = assume_2d[c] != 0;
        const double ss = a2d ? NAN : (vp.sc.R - cv.rz) / cv.nz;
        double rep2 = 0, rep5 = 0, adh2 = 0, adh5 = 0;
        if (!vp.no_surface)
            surface_forces(vp.sc, a2d, cv.rz, cv.nz, cv.hl, cv.l, cellv(cells, ld, c, C_SIGMA0), ss,
                           rep2, rep5, adh2, adh5);
        for (int q = 0; q < 6; ++q) {
            double r = (q == 2) ? rep2 : (q == 5 ? rep5 : 0.0);
            double a = (q == 2) ? adh2 : (q == 5 ? adh5 : 0.0);
            const double cct = fg.frec ? cc[(int64_t)q * n + c] + fr[q] : cc[(int64_t)q * n + c];
            out[(int64_t)q * n + c] = (-cct - r) - a;
        }
    }
}

__global__ void k_assume2d(int n, const double* __restrict__ cells, int ld, double thr, int* __restrict__ a2d)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        a2d[c] = cellv(cells, ld, c, C_NZ) < thr;
}

// Coordinates of cell c at stage st > 0 (mechanics.hpp:2544-2548): cells +
// dt * sum_j A(st, j) k_j, then normalizeOrientations (mechanics.hpp:1918-1959).
__device__ __forceinline__ void stage_coords(const double* __restrict__ cells, int ld, int n, int c,
                                             const double* __restrict__ kbuf, const Tableau& tb, int st, double dt,
                                             double y[6])
{
    y[0] = cellv(cells, ld, c, C_RX); y[1] = cellv(cells, ld, c, C_RY); y[2] = cellv(cells, ld, c, C_RZ);
    y[3] = cellv(cells, ld, c, C_NX); y[4] = cellv(cells, ld, c, C_NY); y[5] = cellv(cells, ld, c, C_NZ);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        double mult = 0;
        for (int j = 0; j < st; ++j) mult += kbuf[((int64_t)j * 6 + q) * n + c] * tb.A[st * tb.s + j];
        y[q] += mult * dt;
    }
    double nrm = norm3(y[3], y[4], y[5]);
    y[3] /= nrm; y[4] /= nrm; y[5] /= nrm;
    if (y[5] < 0) { y[3] *= -1; y[4] *= -1; y[5] *= -1; }
}

// Stage heights and z-orientations (2 x n) for the friction rows of stage st > 0.
__global__ void __launch_bounds__(256)
k_stage_z(int n, const int* __restrict__ dn, const double* __restrict__ cells, int ld, const Tableau tb, int st,
          double dt, const int* __restrict__ rep, const double* __restrict__ dt_rep, const int* __restrict__ active,
          const double* __restrict__ kbuf, double* __restrict__ zst, int n_zst)
{
    if (dn) n = *dn;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        if (rep) {
            const int r = rep[c];
            if (!active[r]) continue;
            dt = dt_rep[r];
        }
        double y[6];
        stage_coords(cells, ld, n, c, kbuf, tb, st, dt, y);
        zst[c] = y[2];
        zst[n_zst + c] = y[5];
    }
}

// Stage orientations (3 x n) for the anisotropic JKR lookups of stage st > 0
// (cellCellForces reads n from the stage cells, mechanics.hpp:1446-1449).
// Cells of inactive replicates keep their orientation.
__global__ void __launch_bounds__(256)
k_stage_n(int n, const int* __restrict__ dn, const double* __restrict__ cells, int ld, const Tableau tb, int st,
          double dt, const int* __restrict__ rep, const double* __restrict__ dt_rep, const int* __restrict__ active,
          const double* __restrict__ kbuf, double* __restrict__ nst, int ld_nst)
{
    if (dn) n = *dn;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        double y[6];
        if (rep && !active[rep[c]]) {
#pragma unroll
            for (int q = 0; q < 3; ++q) y[3 + q] = cellv(cells, ld, c, C_NX + q);
        } else {
            stage_coords(cells, ld, n, c, kbuf, tb, st, rep ? dt_rep[rep[c]] : dt, y);
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) nst[(int64_t)q * ld_nst + c] = y[3 + q];
    }
}

// One RK stage st >= 0 of stepRungeKuttaAdaptive (mechanics.hpp:2534-2557):
// cells_i = cells + dt * sum_j A(st, j) k_j, normalizeOrientations, then
// getVelocities with the step's frozen neighbor geometry (the conservative
// cell-cell forces cc are identical in every stage, because they only read
// the neighbor rows; friction rows, when on, are recomputed per stage).
// With the anisotropic JKR tables the forces also read the stage
// orientations: rec (non-null) holds the stage's pair records and the
// cell-cell forces are gathered from them instead of cc.
// Batched replicates: dn (device cell count) replaces n, and rep / dt_rep /
// active give each cell its replicate's step size; cells of inactive
// replicates are skipped.
template <bool CSR>
__global__ void __launch_bounds__(256)
k_rk_stage(int n, const int* __restrict__ dn, const double* __restrict__ cells, int ld,
           const int* __restrict__ assume_2d, const double* __restrict__ noise, VelParams vp, Tableau tb, int st,
           double dt, const int* __restrict__ rep, const double* __restrict__ dt_rep,
           const int* __restrict__ active, const double* __restrict__ cc, const double* __restrict__ rec,
           RowLists L, FricGather fg, double* __restrict__ kbuf, int* __restrict__ status)
{
    if (dn) n = *dn;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        if (rep) {
            const int r = rep[c];
            if (!active[r]) continue;
            dt = dt_rep[r];
        }
        CellVel cv = load_cellvel(cells, ld, c);
        if (st > 0) {
            double y[6];
            stage_coords(cells, ld, n, c, kbuf, tb, st, dt, y);
            cv.rz = y[2]; cv.nx = y[3]; cv.ny = y[4]; cv.nz = y[5];
        }
        const bool a2d = assume_2d[c] != 0;
        double ccl[6], nz6[6] = {0, 0, 0, 0, 0, 0};
        if (rec) {
            gather_cc<CSR>(c, L, rec, vp.sc.R, ccl);
        } else {
#pragma unroll
            for (int q = 0; q < 6; ++q) ccl[q] = cc[(int64_t)q * n + c];
        }
        if (fg.frec) {
            double fr[6];
            gather_fric<CSR>(c, L, fg.frec, fg.ps, fg.pt, fr);
#pragma unroll
            for (int q = 0; q < 6; ++q) ccl[q] += fr[q];
        }
        if (noise) {
#pragma unroll
            for (int q = 0; q < 6; ++q) nz6[q] = noise[(int64_t)q * n + c];
        }
        double v[6];
        cell_velocity(vp, cv, cellv(cells, ld, c, C_SIGMA0), a2d, ccl, noise != nullptr, nz6, v);
        bool bad = false;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            kbuf[((int64_t)st * 6 + q) * n + c] = v[q];
            bad |= !isfinite(v[q]);
        }
        // batched replicates report non-finite results per replicate (k_rk_final_b)
        if (bad && !rep) atomicOr(status, ST_NAN);
    }
}

__device__ __forceinline__ unsigned long long order_key(double x)
{
    unsigned long long u = (unsigned long long)__double_as_longlong(x);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double order_val(unsigned long long k)
{
    unsigned long long u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}

__global__ void k_err_init(unsigned long long* e) { *e = order_key(-INFINITY); }

inline double order_val_host(unsigned long long k)
{
    unsigned long long u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    double d;
    std::memcpy(&d, &u, sizeof(d));
    return d;
}
__global__ void k_err_final(const unsigned long long* e, double* out) { *out = order_val(*e); }

// RK update (mechanics.hpp:2559-2611): weighted sums, position update,
// embedded error, stored velocities, normalizeOrientations; plus the
// controller's scaled-error maximum (simulation.hpp:1688-1695).
__global__ void __launch_bounds__(256)
k_rk_final(int n, int ncols, const double* __restrict__ cells, int ld, Tableau tb, double dt,
           const double* __restrict__ kbuf, double* __restrict__ cells_new, int ld_new,
           double* __restrict__ errors, double tol, unsigned long long* __restrict__ errmax)
{
    double local = -INFINITY;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        for (int col = 0; col < ncols; ++col)
            if (col < 1 || col > 12) cells_new[(int64_t)col * ld_new + c] = cellv(cells, ld, c, col);
        double y[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            double v1 = 0, v2 = 0;
            for (int st = 0; st < tb.s; ++st) {
                const double kv = kbuf[((int64_t)st * 6 + q) * n + c];
                v1 += kv * tb.b[st];
                v2 += kv * tb.bs[st];
            }
            const double d1 = v1 * dt, d2 = v2 * dt;
            const double y0 = cellv(cells, ld, c, C_RX + q);
            y[q] = y0 + d1;
            const double e = d1 - d2;
            if (errors) errors[(int64_t)q * n + c] = e;
            cells_new[(int64_t)(C_DRX + q) * ld_new + c] = v1;
            if (tol > 0) local = fmax(local, e / (tol * (1.0 + fabs(y0))));
        }
        double nrm = norm3(y[3], y[4], y[5]);
        y[3] /= nrm; y[4] /= nrm; y[5] /= nrm;
        if (y[5] < 0) { y[3] *= -1; y[4] *= -1; y[5] *= -1; }
#pragma unroll
        for (int q = 0; q < 6; ++q) cells_new[(int64_t)(C_RX + q) * ld_new + c] = y[q];
    }
    if (tol > 0) {
        __shared__ double red[256];
        red[threadIdx.x] = local;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
            __syncthreads();
        }
        if (threadIdx.x == 0) atomicMax(errmax, order_key(red[0]));
    }
}

// k_rk_final for batched replicates: device cell count, per-replicate dt and
// active mask (inactive replicates' rows are copied unchanged), and the
// scaled-error maximum per replicate (errmax[r]). Each block updates one
// contiguous chunk of rows. Replicates occupy runs of consecutive rows, so a
// wavefront keeps a per-lane maximum while its successive 64-row slices stay
// in one replicate and reduces it by shuffles into one atomic when the
// replicate changes or the chunk ends; a slice that straddles replicates
// reduces each run serially from its first lane.
__global__ void __launch_bounds__(256)
k_rk_final_b(const int* __restrict__ dn, int ncols, const double* __restrict__ cells, int ld, Tableau tb,
             const int* __restrict__ rep, const double* __restrict__ dt_rep, const int* __restrict__ active,
             const double* __restrict__ kbuf, double* __restrict__ cells_new, double tol,
             unsigned long long* __restrict__ errmax, int* __restrict__ bad_rep)
{
    __shared__ double s_val[256];
    __shared__ int s_rep[256];
    const int n = *dn;
    const int t = threadIdx.x, lane = t & 63, w0 = t & ~63;
    const int64_t chunk = (((int64_t)n + gridDim.x - 1) / gridDim.x + 255) / 256 * 256;
    const int64_t b0 = blockIdx.x * chunk;
    const int64_t b1 = b0 + chunk < n ? b0 + chunk : (int64_t)n;
    double acc = -INFINITY;
    int wr = -1;   // replicate of the wavefront's running maximum (wavefront-uniform)
    auto flush = [&] {
        if (wr < 0) return;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc = fmax(acc, __shfl_xor(acc, off));
        if (lane == 0) atomicMax(&errmax[wr], order_key(acc));
    };
    for (int64_t base = b0; base < b1; base += 256) {
        const int c = (int)(base + t);
        double local = -INFINITY;
        int r = -1;
        if (c < b1) {
            r = rep[c];
            if (!active[r]) {
                r = -1;   // rows of inactive replicates: the iteration's tail after the attempts
            } else {
                const double dt = dt_rep[r];
                double y[6];
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    double v1 = 0, v2 = 0;
                    for (int st = 0; st < tb.s; ++st) {
                        const double kv = kbuf[((int64_t)st * 6 + q) * n + c];
                        v1 += kv * tb.b[st];
                        v2 += kv * tb.bs[st];
                    }
                    const double d1 = v1 * dt, d2 = v2 * dt;
                    const double y0 = cellv(cells, ld, c, C_RX + q);
                    y[q] = y0 + d1;
                    cells_new[(int64_t)(C_DRX + q) * ld + c] = v1;
                    local = fmax(local, (d1 - d2) / (tol * (1.0 + fabs(y0))));
                }
                double nrm = norm3(y[3], y[4], y[5]);
                y[3] /= nrm; y[4] /= nrm; y[5] /= nrm;
                if (y[5] < 0) { y[3] *= -1; y[4] *= -1; y[5] *= -1; }
                bool bad = false;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    cells_new[(int64_t)(C_RX + q) * ld + c] = y[q];
                    bad |= !isfinite(y[q]);
                }
                if (bad) atomicOr(&bad_rep[r], 1);
            }
        }
        const int r0 = __shfl(r, 0);
        const bool uniform = __all(r == r0);
        if (uniform && r0 == wr) {
            acc = fmax(acc, local);
            continue;
        }
        flush();
        acc = -INFINITY;
        wr = -1;
        if (uniform) {
            if (r0 >= 0) {
                wr = r0;
                acc = local;
            }
        } else {
            s_val[t] = local;
            s_rep[t] = r;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (r >= 0 && (lane == 0 || s_rep[t - 1] != r)) {
                double m = local;
                for (int u = t + 1; u < w0 + 64 && s_rep[u] == r; ++u) m = fmax(m, s_val[u]);
                atomicMax(&errmax[r], order_key(m));
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    flush();
}

// One whole RK attempt of batched replicates when no stage needs another
// cell: with an isotropic interaction, no friction and no noise every
// stage's cell-cell forces are the step's frozen sums cc, so a cell's stages
// depend only on its own earlier stages. Per cell, the s stages of k_rk_stage
// one after the other (the stage velocities go to kbuf, written and read back
// by the same thread), then the update and the per-replicate scaled-error
// maximum of k_rk_final_b, with the same arithmetic and the same row chunks:
// the staged attempt's result, in one launch instead of s + 1.
// An iteration's first attempt (A.first_go non-null) takes its replicates
// and step sizes from the controller state instead of active / dt_rep (those
// with first_go[r] and not failed[r], at dt[r]); GATHER: the cell-cell sums
// are first gathered from the pair records in the reference's row order (as
// k_velocities<true>) and stored in cc for the later attempts. ticket
// non-null: the last work-group then runs the controller's step after the
// attempt (ctrl_after_block), so the attempt and its controller step are one
// launch (for small grids: every work-group pays a device-scope release).
// S > 0: the tableau has S stages (tb.s == S) and the stage velocities stay
// in registers; S == 0: any tb.s, the stage velocities through kbuf.
template <bool GATHER, int S>
__global__ void __launch_bounds__(256)
k_rk_attempt_b(const int* __restrict__ dn, const double* __restrict__ cells, int ld, VelParams vp, Tableau tb,
               const int* __restrict__ rep, const double* __restrict__ dt_rep, const int* __restrict__ active,
               RowLists L, const double* __restrict__ rec, double* __restrict__ cc, double* __restrict__ kbuf,
               double* __restrict__ cells_new, double tol, unsigned long long* __restrict__ errmax,
               int* __restrict__ bad_rep, CtrlAfter A, unsigned* __restrict__ ticket)
{
    const int* first_go = A.first_go;
    if (A.gate && *A.gate == 0) return;   // an iteration enqueued ahead that may not run
    // an attempt after every replicate finished its iteration: nothing to do
    // (read before any work-group can reach the controller step below)
    if (!first_go && A.S.pending && *A.S.pending == 0) return;
    __shared__ double s_val[256];
    __shared__ int s_rep[256];
    const int n = *dn;
    const int t = threadIdx.x, lane = t & 63, w0 = t & ~63;
    const int64_t chunk = (((int64_t)n + gridDim.x - 1) / gridDim.x + 255) / 256 * 256;
    const int64_t b0 = blockIdx.x * chunk;
    const int64_t b1 = b0 + chunk < n ? b0 + chunk : (int64_t)n;
    double acc = -INFINITY;
    int wr = -1;   // replicate of the wavefront's running maximum (wavefront-uniform)
    auto flush = [&] {
        if (wr < 0) return;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc = fmax(acc, __shfl_xor(acc, off));
        if (lane == 0) atomicMax(&errmax[wr], order_key(acc));
    };
    for (int64_t base = b0; base < b1; base += 256) {
        const int c = (int)(base + t);
        double local = -INFINITY;
        int r = -1;
        if (c < b1) {
            r = rep[c];
            const bool on = first_go ? (first_go[r] && !A.S.failed[r]) : active[r] != 0;
            if (!on) {
                r = -1;   // rows of inactive replicates: the iteration's tail after the attempts
            } else {
                const double dt = first_go ? A.S.dt[r] : dt_rep[r];
                const bool a2d = cellv(cells, ld, c, C_NZ) < vp.nz_threshold;
                const double sigma0 = cellv(cells, ld, c, C_SIGMA0);
                double ccl[6];
                const double nz6[6] = {0, 0, 0, 0, 0, 0};
                if (GATHER) {
                    gather_cc<true>(c, L, rec, vp.sc.R, ccl);
#pragma unroll
                    for (int q = 0; q < 6; ++q) cc[(int64_t)q * n + c] = ccl[q];
                } else {
#pragma unroll
                    for (int q = 0; q < 6; ++q) ccl[q] = cc[(int64_t)q * n + c];
                }
                double y[6];
                double kr[S > 0 ? S : 1][6];   // S > 0: the stage velocities in registers
                if (S > 0) {
                    // stage_coords' arithmetic on register copies of the cell
                    const CellVel cv0 = load_cellvel(cells, ld, c);
                    const double x0[6] = {cellv(cells, ld, c, C_RX), cellv(cells, ld, c, C_RY), cv0.rz,
                                          cv0.nx, cv0.ny, cv0.nz};
#pragma unroll
                    for (int st = 0; st < (S > 0 ? S : 1); ++st) {
                        CellVel cv = cv0;
                        if (st > 0) {
                            double ys[6];
#pragma unroll
                            for (int q = 0; q < 6; ++q) {
                                double mult = 0;
#pragma unroll
                                for (int j = 0; j < st; ++j) mult += kr[j][q] * tb.A[st * tb.s + j];
                                ys[q] = x0[q];
                                ys[q] += mult * dt;
                            }
                            double nrm = norm3(ys[3], ys[4], ys[5]);
                            ys[3] /= nrm; ys[4] /= nrm; ys[5] /= nrm;
                            if (ys[5] < 0) { ys[3] *= -1; ys[4] *= -1; ys[5] *= -1; }
                            cv.rz = ys[2]; cv.nx = ys[3]; cv.ny = ys[4]; cv.nz = ys[5];
                        }
                        cell_velocity(vp, cv, sigma0, a2d, ccl, false, nz6, kr[st]);
                    }
                } else {
                    for (int st = 0; st < tb.s; ++st) {
                        CellVel cv = load_cellvel(cells, ld, c);
                        if (st > 0) {
                            double ys[6];
                            stage_coords(cells, ld, n, c, kbuf, tb, st, dt, ys);
                            cv.rz = ys[2]; cv.nx = ys[3]; cv.ny = ys[4]; cv.nz = ys[5];
                        }
                        double v[6];
                        cell_velocity(vp, cv, sigma0, a2d, ccl, false, nz6, v);
#pragma unroll
                        for (int q = 0; q < 6; ++q) kbuf[((int64_t)st * 6 + q) * n + c] = v[q];
                    }
                }
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    double v1 = 0, v2 = 0;
                    if (S > 0) {
#pragma unroll
                        for (int st = 0; st < (S > 0 ? S : 1); ++st) {
                            v1 += kr[st][q] * tb.b[st];
                            v2 += kr[st][q] * tb.bs[st];
                        }
                    } else {
                        for (int st = 0; st < tb.s; ++st) {
                            const double kv = kbuf[((int64_t)st * 6 + q) * n + c];
                            v1 += kv * tb.b[st];
                            v2 += kv * tb.bs[st];
                        }
                    }
                    const double d1 = v1 * dt, d2 = v2 * dt;
                    const double y0 = cellv(cells, ld, c, C_RX + q);
                    y[q] = y0 + d1;
                    cells_new[(int64_t)(C_DRX + q) * ld + c] = v1;
                    local = fmax(local, (d1 - d2) / (tol * (1.0 + fabs(y0))));
                }
                double nrm = norm3(y[3], y[4], y[5]);
                y[3] /= nrm; y[4] /= nrm; y[5] /= nrm;
                if (y[5] < 0) { y[3] *= -1; y[4] *= -1; y[5] *= -1; }
                bool bad = false;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    cells_new[(int64_t)(C_RX + q) * ld + c] = y[q];
                    bad |= !isfinite(y[q]);
                }
                if (bad) atomicOr(&bad_rep[r], 1);
            }
        }
        const int r0 = __shfl(r, 0);
        const bool uniform = __all(r == r0);
        if (uniform && r0 == wr) {
            acc = fmax(acc, local);
            continue;
        }
        flush();
        acc = -INFINITY;
        wr = -1;
        if (uniform) {
            if (r0 >= 0) {
                wr = r0;
                acc = local;
            }
        } else {
            s_val[t] = local;
            s_rep[t] = r;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (r >= 0 && (lane == 0 || s_rep[t - 1] != r)) {
                double m = local;
                for (int u = t + 1; u < w0 + 64 && s_rep[u] == r; ++u) m = fmax(m, s_val[u]);
                atomicMax(&errmax[r], order_key(m));
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    flush();
    if (ticket && last_block_done(ticket)) ctrl_after_block(A);
}

__global__ void __launch_bounds__(256) k_ctrl_after_step(CtrlAfter A) { ctrl_after_block(A); }

// The end of a batched iteration in one launch, for every cell: the columns
// the RK attempts leave alone (all but the coordinates and velocities 1-12
// of the replicates that stepped, every column of the others: finished or
// failed replicates keep the state the iteration started from), growCells on
// the new array (as k_grow_b, growth.hpp:51-58) and the rows past the division
// length (divideMaxLength's test, growth.hpp:68-73) appended to idx with
// count (zeroed by the controller's first step of the iteration); geom (when
// not null) receives the new state's packed segments as k_pack_segs writes
// them. The replicates that stepped and their dt come from the controller
// state, whose end-of-iteration step (ctrl_end_block, writing the per-replicate
// values after snap) one work-group runs. Nothing when *gate == 0 (gate
// non-null: an iteration enqueued ahead that may not run); only
// snap->tail_ran = 0 and *spec_out = 0 when *skip != 0 (a replicate still
// retries). ticket non-null: the last work-group runs the end step and
// publishes the iteration (snap, *spec_out); without a ticket work-group 0
// runs the end step and k_publish_tail publishes.
__global__ void __launch_bounds__(256)
k_iter_tail_b(const int* __restrict__ dn, int ncols, const double* __restrict__ cells, int ld,
              const int* __restrict__ rep, CtrlEnd E, double R, double Ldiv, double* __restrict__ cells_new,
              int* __restrict__ count, int* __restrict__ idx, double* __restrict__ geom,
              const int* __restrict__ skip, const int* __restrict__ gate, IterSnap* snap,
              int* __restrict__ spec_out, const int* __restrict__ status, unsigned* __restrict__ ticket)
{
    if (gate && *gate == 0) return;
    if (*skip) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            snap->tail_ran = 0;
            *spec_out = 0;
        }
        return;
    }
    if (!ticket && blockIdx.x == 0) ctrl_end_block(E);
    const int n = *dn;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const int r = rep[c];
        const bool on = E.S.stepped[r] != 0;
        for (int col = 0; col < ncols; ++col)
            if (!on || col < 1 || col > 12) cells_new[(int64_t)col * ld + c] = cellv(cells, ld, c, col);
        double l = cellv(cells, ld, c, C_L);
        double hl = cellv(cells, ld, c, C_HALF_L);
        if (on) {
            const double g = cellv(cells, ld, c, C_GROWTH);
            l = l + (g * (4 * R / 3 + l)) * E.S.dt[r];
            hl = l / 2;
            cells_new[(int64_t)C_L * ld + c] = l;
            cells_new[(int64_t)C_HALF_L * ld + c] = hl;
        }
        if (l > Ldiv) idx[atomicAdd(count, 1)] = c;
        if (geom) {
            const double* src = on ? cells_new : cells;   // the new state's coordinates
            double2* gp = reinterpret_cast<double2*>(geom + (int64_t)c * 8);
            gp[0] = make_double2(cellv(src, ld, c, C_RX), cellv(src, ld, c, C_RY));
            gp[1] = make_double2(cellv(src, ld, c, C_RZ), cellv(src, ld, c, C_NX));
            gp[2] = make_double2(cellv(src, ld, c, C_NY), cellv(src, ld, c, C_NZ));
            gp[3] = make_double2(hl, 0.0);
        }
    }
    if (ticket && last_block_done(ticket)) {
        ctrl_end_block(E);
        __syncthreads();
        if (threadIdx.x == 0)
            *spec_out = publish_iteration(E, __hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                                          __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                                          snap);
    }
}

// The publication of a tail without a ticket (see k_iter_tail_b).
__global__ void k_publish_tail(const int* __restrict__ count, const int* __restrict__ status, CtrlEnd E,
                               const int* __restrict__ skip, const int* __restrict__ gate, IterSnap* snap,
                               int* __restrict__ spec_out)
{
    if ((gate && *gate == 0) || *skip) return;
    *spec_out = publish_iteration(E, *count, *status, snap);
}

// growCells (growth.hpp:51-58) with each replicate's dt (inactive: unchanged).
__global__ void k_grow_b(const int* __restrict__ dn, double* __restrict__ cells, int ld, double R,
                         const int* __restrict__ rep, const double* __restrict__ dt_rep,
                         const int* __restrict__ active)
{
    const int n = *dn;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int r = rep[i];
        if (!active[r]) continue;
        double l = cells[(int64_t)C_L * ld + i];
        const double g = cells[(int64_t)C_GROWTH * ld + i];
        l = l + (g * (4 * R / 3 + l)) * dt_rep[r];
        cells[(int64_t)C_L * ld + i] = l;
        cells[(int64_t)C_HALF_L * ld + i] = l / 2;
    }
}

__global__ void k_fill_key(unsigned long long* __restrict__ p, int n, double v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = order_key(v);
}

__global__ void k_box_init(unsigned long long* __restrict__ box, int n_rep)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rep) return;
    for (int k = 0; k < 7; ++k) box[r * 7 + k] = order_key(k < 3 ? INFINITY : -INFINITY);
}

// Per-replicate bounding boxes and max half-length (ordered-key atomics).
// Each block reduces one contiguous chunk of cells. Replicates occupy runs of
// consecutive cells, so a wavefront keeps per-lane minima and maxima while its
// successive 64-cell slices stay in one replicate and reduces them by shuffles
// into one set of atomics when the replicate changes or the chunk ends; a
// slice that straddles replicates reduces each run serially from its first
// lane (at most 64 cells). Minima and maxima are exact, so the order of the
// reduction does not change the result.
__global__ void __launch_bounds__(256)
k_bbox_rep(const double* __restrict__ cells, int n, int ld, const int* __restrict__ rep,
           unsigned long long* __restrict__ box)
{
    __shared__ double s_v[7][256];
    __shared__ int s_rep[256];
    const int t = threadIdx.x, lane = t & 63, w0 = t & ~63;
    const int64_t chunk = (((int64_t)n + gridDim.x - 1) / gridDim.x + 255) / 256 * 256;
    const int64_t b0 = blockIdx.x * chunk;
    const int64_t b1 = b0 + chunk < n ? b0 + chunk : (int64_t)n;
    double acc[7];
    int wr = -1;   // replicate of the wavefront's accumulators (wavefront-uniform)
    auto reset = [&] {
        for (int k = 0; k < 3; ++k) { acc[k] = INFINITY; acc[3 + k] = -INFINITY; }
        acc[6] = -INFINITY;
    };
    auto flush = [&] {
        if (wr < 0) return;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                acc[k] = fmin(acc[k], __shfl_xor(acc[k], off));
                acc[3 + k] = fmax(acc[3 + k], __shfl_xor(acc[3 + k], off));
            }
            acc[6] = fmax(acc[6], __shfl_xor(acc[6], off));
        }
        if (lane == 0) {
            for (int k = 0; k < 3; ++k) {
                atomicMin(&box[wr * 7 + k], order_key(acc[k]));
                atomicMax(&box[wr * 7 + 3 + k], order_key(acc[3 + k]));
            }
            atomicMax(&box[wr * 7 + 6], order_key(acc[6]));
        }
    };
    reset();
    for (int64_t base = b0; base < b1; base += 256) {
        const int64_t i = base + t;
        const int r = i < b1 ? rep[i] : -1;
        double v[7];
        if (r >= 0) {
            for (int k = 0; k < 3; ++k) v[k] = v[3 + k] = cellv(cells, ld, (int)i, C_RX + k);
            v[6] = cellv(cells, ld, (int)i, C_HALF_L);
        } else {
            for (int k = 0; k < 3; ++k) { v[k] = INFINITY; v[3 + k] = -INFINITY; }
            v[6] = -INFINITY;
        }
        const int r0 = __shfl(r, 0);
        const bool uniform = __all(r == r0);
        if (uniform && r0 == wr) {
            for (int k = 0; k < 3; ++k) { acc[k] = fmin(acc[k], v[k]); acc[3 + k] = fmax(acc[3 + k], v[3 + k]); }
            acc[6] = fmax(acc[6], v[6]);
            continue;
        }
        flush();
        reset();
        wr = -1;
        if (uniform) {
            if (r0 >= 0) {
                wr = r0;
                for (int k = 0; k < 7; ++k) acc[k] = v[k];
            }
            continue;
        }
        for (int k = 0; k < 7; ++k) s_v[k][t] = v[k];
        s_rep[t] = r;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (r >= 0 && (lane == 0 || s_rep[t - 1] != r)) {
            for (int u = t + 1; u < w0 + 64 && s_rep[u] == r; ++u) {
                for (int k = 0; k < 3; ++k) {
                    v[k] = fmin(v[k], s_v[k][u]);
                    v[3 + k] = fmax(v[3 + k], s_v[3 + k][u]);
                }
                v[6] = fmax(v[6], s_v[6][u]);
            }
            for (int k = 0; k < 3; ++k) {
                atomicMin(&box[r * 7 + k], order_key(v[k]));
                atomicMax(&box[r * 7 + 3 + k], order_key(v[3 + k]));
            }
            atomicMax(&box[r * 7 + 6], order_key(v[6]));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    flush();
}

// Grid plans of all replicates with one shared edge (so one stencil reach):
// the edge grows by 1.25 until the replicates' bins fit in max_bins; bases are
// the exclusive prefix sums of the replicates' bin counts. One block.
__global__ void __launch_bounds__(256)
k_plan_rep(const unsigned long long* __restrict__ box, int n_rep, double thr, double pre, double bin_scale,
           int max_bins, Plan* __restrict__ plan, RepGrid* __restrict__ grid)
{
    __shared__ double s_red[256];
    __shared__ long long s_tot[256];
    __shared__ double s_edge;
    const int t = threadIdx.x;
    double hl = 0, amax = 0;
    for (int r = t; r < n_rep; r += 256) {
        const double lo0 = order_val(box[r * 7]);
        if (!(lo0 <= order_val(box[r * 7 + 3]))) continue;   // empty replicate
        hl = fmax(hl, order_val(box[r * 7 + 6]));
        for (int k = 0; k < 3; ++k)
            amax = fmax(amax, fmax(fabs(order_val(box[r * 7 + k])), fabs(order_val(box[r * 7 + 3 + k]))));
    }
    s_red[t] = hl;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) { if (t < w) s_red[t] = fmax(s_red[t], s_red[t + w]); __syncthreads(); }
    hl = s_red[0];
    __syncthreads();
    s_red[t] = amax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) { if (t < w) s_red[t] = fmax(s_red[t], s_red[t + w]); __syncthreads(); }
    amax = s_red[0];
    const double cut = fmin(pre, thr + 2 * hl) * (1 + 1e-9) + 1e-9;
    if (t == 0) s_edge = cut * bin_scale;
    __syncthreads();
    for (;;) {
        const double edge = s_edge;
        long long tot = 0;
        for (int r = t; r < n_rep; r += 256) {
            long long b = 1;
            const double lo0 = order_val(box[r * 7]);
            if (lo0 <= order_val(box[r * 7 + 3]))
                for (int k = 0; k < 3; ++k)
                    b *= (long long)floor((order_val(box[r * 7 + 3 + k]) - order_val(box[r * 7 + k])) / edge) + 1;
            tot += b;
        }
        s_tot[t] = tot;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) { if (t < w) s_tot[t] += s_tot[t + w]; __syncthreads(); }
        const bool fits = s_tot[0] <= max_bins;
        __syncthreads();
        if (fits) break;
        if (t == 0) s_edge = edge * 1.25;
        __syncthreads();
    }
    const double edge = s_edge;
    // each thread builds its replicates' grids; bases are the exclusive
    // prefix sums of the bin counts, 256 replicates per block scan
    __shared__ int s_scan[256];
    int base = 0;
    for (int c0 = 0; c0 < n_rep; c0 += 256) {
        const int r = c0 + t;
        RepGrid G;
        int cnt = 0;
        if (r < n_rep) {
            const double lo0 = order_val(box[r * 7]);
            if (lo0 <= order_val(box[r * 7 + 3])) {
                for (int k = 0; k < 3; ++k) {
                    G.lo[k] = order_val(box[r * 7 + k]);
                    G.dims[k] = (int)floor((order_val(box[r * 7 + 3 + k]) - G.lo[k]) / edge) + 1;
                }
            } else {
                for (int k = 0; k < 3; ++k) { G.lo[k] = 0; G.dims[k] = 1; }
            }
            cnt = G.dims[0] * G.dims[1] * G.dims[2];
        }
        s_scan[t] = cnt;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int v = t >= off ? s_scan[t - off] : 0;
            __syncthreads();
            s_scan[t] += v;
            __syncthreads();
        }
        if (r < n_rep) {
            G.base = base + s_scan[t] - cnt;
            grid[r] = G;
        }
        base += s_scan[255];
        __syncthreads();
    }
    if (t != 0) return;
    int reach = (int)ceil(cut / edge - 1e-12);
    if (reach < 1) reach = 1;
    if (reach * edge < cut) reach++;
    Plan P;
    for (int k = 0; k < 3; ++k) { P.lo[k] = 0; P.dims[k] = 1; }
    P.edge = edge;
    P.inv_edge = 1.0 / edge;
    P.reach = reach;
    P.nbins = base;
    P.fmargin = 1e-3 + 1e-6 * amax;
    P.pre = pre;
    P.thr = thr;
    P.ok = 1;
    *plan = P;
}

// Bin keys of batched replicates: grid[rep].base + local bin.
__global__ void k_keys_rep(const double* __restrict__ cells, int n, int ld, const Plan* __restrict__ plan,
                           const int* __restrict__ rep, const RepGrid* __restrict__ grid,
                           int* __restrict__ keys, int* __restrict__ vals, int* __restrict__ hist)
{
    const double edge = plan->edge;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const RepGrid G = grid[rep[i]];
        int bx = bin_coord(cellv(cells, ld, i, C_RX), G.lo[0], edge, G.dims[0]);
        int by = bin_coord(cellv(cells, ld, i, C_RY), G.lo[1], edge, G.dims[1]);
        int bz = bin_coord(cellv(cells, ld, i, C_RZ), G.lo[2], edge, G.dims[2]);
        int key = G.base + (bz * G.dims[1] + by) * G.dims[0] + bx;
        keys[i] = key;
        vals[i] = atomicAdd(&hist[key], 1);   // rank of the cell within its bin
    }
}

}  // namespace bfm

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
namespace {

constexpr int kMaxCols = 32;
constexpr int kBBoxParts = 256;

template <typename T>
void dalloc(T** p, size_t count)
{
    HIPCHK(hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1)));
}

int nbits(int64_t v)
{
    int b = 1;
    while ((int64_t(1) << b) <= v) ++b;
    return b;
}

int take_status(bfm_ctx* c)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    int st = 0;
    HIPCHK(hipMemcpy(&st, c->d_status, sizeof(int), hipMemcpyDeviceToHost));
    if (st) {
        int zero = 0;
        HIPCHK(hipMemcpy(c->d_status, &zero, sizeof(int), hipMemcpyHostToDevice));
    }
    return st;
}

// Doubles the per-cell slot capacity of the neighbor build.
void grow_slots(bfm_ctx* c)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    const int cap = c->slot_cap * 2;
    int* sj = nullptr;
    double* sr = nullptr;
    HIPCHK(hipMalloc((void**)&sj, sizeof(int) * (size_t)c->max_cells * cap));
    HIPCHK(hipMalloc((void**)&sr, sizeof(double) * (size_t)c->max_cells * cap * 8));
    int *js = nullptr, *jr = nullptr;
    HIPCHK(hipMalloc((void**)&js, sizeof(int) * (size_t)c->max_cells * cap));
    HIPCHK(hipMalloc((void**)&jr, sizeof(int) * (size_t)c->max_cells * cap));
    int* jk = nullptr;
    unsigned long long* oj = nullptr;
    HIPCHK(hipMalloc((void**)&jk, sizeof(int) * (size_t)c->max_cells * cap));
    HIPCHK(hipMalloc((void**)&oj, sizeof(unsigned long long) * (size_t)c->max_cells * cap));
    for (void* q : {(void*)c->d_slot_j, (void*)c->d_slot_rec, (void*)c->d_jslot, (void*)c->d_jrow,
                    (void*)c->d_jslotk, (void*)c->d_ordj})
        HIPCHK(hipFree(q));
    c->d_jslotk = jk;
    c->d_ordj = oj;
    c->d_slot_j = sj;
    c->d_slot_rec = sr;
    c->d_jslot = js;
    c->d_jrow = jr;
    c->slot_cap = cap;
}

void raise_status(bfm_ctx* c, int st)
{
    if (!st) return;
    if (st & ST_SLOT_OVERFLOW) {
        grow_slots(c);
        throw ArgError{BFM_ERR_CAPACITY, "a cell had more neighbors than the per-cell slot capacity; "
                                         "the capacity has been doubled, run the call again"};
    }
    if (st & ST_PAIR_OVERFLOW)
        throw ArgError{BFM_ERR_CAPACITY, "neighbor pairs exceed max_pairs of the context"};
    if (st & ST_JKR_ROOT)
        throw ArgError{BFM_ERR_JKR_ROOT,
                       "JKR contact radius is undefined; invalid number of real roots found "
                       "from radius-overlap polynomial"};
    if (st & ST_NAN) throw ArgError{BFM_ERR_NAN, "Found nan in velocities"};
    throw ArgError{BFM_ERR_HIP, "internal device status error"};
}

void check_status(bfm_ctx* c) { raise_status(c, take_status(c)); }

void need_params(bfm_ctx* c)
{
    require(c->params_set, BFM_ERR_ARG, "bfm_set_params has not been called");
}

void check_cells(bfm_ctx* c, int n, int ld)
{
    require(n >= 0 && n <= c->max_cells, BFM_ERR_CAPACITY, "n exceeds max_cells of the context");
    require(ld >= n, BFM_ERR_ARG, "ld < n");
}

ForceParams force_params_of(const bfm_params& p)
{
    ForceParams fp;
    fp.R = p.R; fp.Rcell = p.Rcell; fp.E0 = p.E0;
    fp.Req = p.R / 2;
    fp.max_overlap = 2 * (p.R - p.Rcell);
    fp.pref0 = p.repulsion_prefactors[0];
    fp.pref1 = p.repulsion_prefactors[1];
    fp.pref2 = p.repulsion_prefactors[2];
    fp.adhesion = p.adhesion_mode != BFM_ADHESION_NONE;
    fp.jkr_mode = p.jkr_mode;
    fp.colidx_gamma = p.colidx_gamma;
    fp.pad = 0;
    fp.env = SurrogateEnvelope{1, 0, 1, 0, 1};   // contains nothing
    return fp;
}

ForceParams force_params(const bfm_ctx* c)
{
    ForceParams fp = force_params_of(c->p);
    fp.env = bfm::surrogate_envelope(c);
    return fp;
}

AnisoJKR aniso_params(const bfm_ctx* c)
{
    require(c->d_an_grid != nullptr, BFM_ERR_ARG, "anisotropic JKR tables not set (bfm_set_aniso_jkr_tables)");
    AnisoJKR an;
    an.grid = c->d_an_grid;
    an.curv = c->d_an_curv;
    an.fr = reinterpret_cast<const double2*>(c->d_an_fr);
    int off = 0;
    for (int k = 0; k < kAnisoGrids; ++k) {
        an.n[k] = c->an_n[k];
        an.off[k] = off;
        an.inv_h[k] = c->an_inv_h[k];
        off += c->an_n[k];
    }
    an.total = off;
    return an;
}

VelParams vel_params_of(const bfm_params& p)
{
    VelParams vp;
    const double R = p.R, E0 = p.E0;
    vp.sc.R = R; vp.sc.E0 = E0; vp.sc.sqR = std::sqrt(R);
    vp.sc.rep0 = 2 * E0;
    vp.sc.rep1 = (8. / 3.) * E0 * std::pow(R, 0.5);
    vp.sc.rep2 = 2 * E0 * std::pow(R, 0.5);
    vp.sc.adh0 = std::pow(R, 0.5) / 2;
    vp.sc.adh1 = 2 * bfm::kPi * R;
    vp.sc.adh2 = 2 * std::pow(R, 0.5);
    vp.nz_threshold = p.nz_threshold;
    vp.coeff = p.cell_surface_coulomb_coeff;
    vp.no_surface = p.no_surface;
    vp.has_noise = 0;
    return vp;
}

VelParams vel_params(const bfm_ctx* c) { return vel_params_of(c->p); }

void scan(bfm_ctx* c, const int* in, int* out, int64_t items, const char* name)
{
    LaunchScope ls(c, name);
    size_t tb = c->temp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(c->d_temp, tb, in, out, (int)items, c->stream));
}

void launch_candidates(bfm_ctx* c, int n, int64_t stride)
{
#define BFM_CAND(HOMES, CAP, BATCH, QCAP)                                                              \
    KLAUNCH(c, "candidates", (k_candidates_lds<HOMES, CAP, BATCH, QCAP>), (n + HOMES - 1) / HOMES, 256, 0, n, \
            c->d_plan, c->d_sorted, stride, c->d_sid, c->d_skey, c->n_rep > 0 ? c->d_rep : (const int*)nullptr, \
            c->d_grid, c->d_bin_off, c->slot_cap, c->d_cnt64,                                             \
            c->d_slot_j, c->d_slot_rec, c->d_jslot, c->d_counters)
#define BFM_CANDW(CAP, BATCH, QCAP)                                                                      \
    KLAUNCH(c, "candidates", (k_candidates_wave<CAP, BATCH, QCAP>), (n + 63) / 64, 256, 0, n, c->d_plan,     \
            c->d_sorted, stride, c->d_sid, c->d_skey, c->n_rep > 0 ? c->d_rep : (const int*)nullptr,         \
            c->d_grid, c->d_bin_off, c->slot_cap, c->d_cnt64, c->d_slot_j, c->d_slot_rec, c->d_jslot,        \
            c->d_counters)
    switch (c->cand_variant) {
        case 8:
            KLAUNCH(c, "candidates", k_candidates_lane<4>, grid_for(n, 256), 256, 0, n, c->d_plan, c->d_sgeom,
                    stride, c->d_sortf, c->d_sortf + stride, c->d_sid, c->d_skey,
                    c->n_rep > 0 ? c->d_rep : (const int*)nullptr, c->d_grid, c->d_bin_off, c->slot_cap, c->d_cnt64,
                    c->d_own64, c->d_slot_j, c->d_slot_rec, c->d_jslot, c->d_jslotk, c->d_counters);
            break;
        case 7:
            KLAUNCH(c, "candidates", k_candidates_bins<256>, 2048, 256, 0, c->d_plan, c->d_sorted, stride, c->d_sid,
                    c->d_grid, std::max(c->n_rep, 1), c->d_bin_off, c->slot_cap, c->d_cnt64, c->d_slot_j,
                    c->d_slot_rec, c->d_jslot, c->d_counters);
            break;
        case 4: BFM_CANDW(512, 8, 4096); break;
        case 5: BFM_CANDW(1024, 8, 8192); break;
        case 6: BFM_CAND(128, 512, 8, 4096); break;
        case 0: BFM_CAND(32, 640, 8, 8192); break;
        case 2: BFM_CAND(64, 1024, 8, 8192); break;
        case 3: BFM_CAND(64, 512, 8, 4096); break;
        default: BFM_CAND(64, 640, 8, 8192); break;
    }
#undef BFM_CAND
#undef BFM_CANDW
}

// Device neighbor build into the context's internal list.
// fuse: the compaction also writes the pair-force records of an isotropic
// interaction (see fused_forces_ok); n > 1 only.
void build_neighbors(bfm_ctx* c, const double* d_cells, int n, int ld, double thr, double R, double Ldiv,
                     bool fuse = false, bool compact = true)
{
    c->transposed = false;
    c->list_n = n;
    c->list_sorted_i = true;
    c->use_perm_i = false;
    if (n <= 1) {
        int zero = 0;
        HIPCHK(hipMemcpyAsync(c->d_m, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
        KLAUNCH(c, "fill", k_fill_u64, 1, 64, 0, c->d_off64, (int64_t)(n + 1));
        c->h_m = 0;
        c->list_csr = true;
        return;
    }
    const double pre = thr + Ldiv + 2 * R;
    if (c->n_rep > 0) {
        KLAUNCH(c, "bbox", k_box_init, grid_for(c->n_rep, 256), 256, 0, c->d_rep_box, c->n_rep);
        KLAUNCH(c, "bbox", k_bbox_rep, std::min(grid_for(n, 2048), 1024), 256, 0, d_cells, n, ld, c->d_rep,
                c->d_rep_box);
        KLAUNCH(c, "plan", k_plan_rep, 1, 256, 0, c->d_rep_box, c->n_rep, thr, pre, c->bin_scale, c->max_bins,
                c->d_plan, c->d_grid);
    } else {
        const int nparts = std::min(kBBoxParts, grid_for(n, 256));
        KLAUNCH(c, "bbox", k_bbox_partial, nparts, 256, 0, d_cells, n, ld, c->d_partials);
        KLAUNCH(c, "plan", k_plan, 1, 256, 0, c->d_partials, nparts, thr, pre, c->bin_scale, c->max_bins,
                c->d_plan, c->d_grid);
    }
    // k_scatter clears the histogram entries it used, so the histogram is
    // zero for the next build; a fill is needed only for the first build or
    // after one that did not reach the scatter
    if (c->bins_dirty)
        KLAUNCH(c, "fill", k_fill_int, grid_for(c->max_bins + 1, 256, 4096), 256, 0, c->d_bin_count, 0,
                (int64_t)c->max_bins + 1);
    c->bins_dirty = true;
    if (c->n_rep > 0)
        KLAUNCH(c, "keys", k_keys_rep, grid_for(n, 256), 256, 0, d_cells, n, ld, c->d_plan, c->d_rep, c->d_grid,
                c->d_keys, c->d_vals, c->d_bin_count);
    else
        KLAUNCH(c, "keys", k_keys, grid_for(n, 256), 256, 0, d_cells, n, ld, c->d_plan, c->d_keys, c->d_vals,
                c->d_bin_count);
    {
        const int tiles = (int)((c->max_bins + kScanTile) / kScanTile);   // tiles of max_bins + 1 entries
        const unsigned epoch = ++c->scan_epoch & 0x3fffffffu;
        KLAUNCH(c, "scan_bins", k_scan_onepass<int>, tiles, 256, 0, c->d_plan, (int64_t)0, c->d_bin_count,
                c->d_bin_off, c->d_tile_sum, c->d_scan_flag, epoch, c->d_status);
    }
    const int64_t stride = c->max_cells;
    if (c->cand_variant == 8)
        KLAUNCH(c, "scatter", k_scatter<false>, grid_for(n, 256), 256, 0, d_cells, n, ld, c->d_keys, c->d_vals,
                c->d_bin_off, c->d_sorted, c->d_sid, c->d_skey, stride, c->d_cnt64, c->d_own64, c->d_bin_count, c->d_sortf,
                c->d_sgeom);
    else
        KLAUNCH(c, "scatter", k_scatter<true>, grid_for(n, 256), 256, 0, d_cells, n, ld, c->d_keys, c->d_vals,
                c->d_bin_off, c->d_sorted, c->d_sid, c->d_skey, stride, c->d_cnt64, c->d_own64, c->d_bin_count, c->d_sortf,
                c->d_sgeom);
    c->bins_dirty = false;
    launch_candidates(c, n, stride);
    if (!compact) {   // per-cell lists only: no rows, no reusable list in the context
        c->list_n = -1;
        c->h_m = -1;
        c->list_csr = false;
        return;
    }
    {
        // (the one-pass scan measured no faster than hipCUB on the n + 1 row counts)
        LaunchScope ls(c, "scan_counts");
        size_t tb = c->temp_bytes;
        hipLaunchKernelGGL(k_add_u64, dim3(grid_for(n + 1, 256)), dim3(256), 0, c->stream, n + 1, c->d_cnt64,
                           c->d_own64, c->d_tot64);
        HIPCHK(hipGetLastError());
        HIPCHK(hipcub::DeviceScan::ExclusiveSum(c->d_temp, tb, c->d_tot64, c->d_off64, n + 1, c->stream));
    }
    if (fuse) {
        const FusedForces ff{force_params(c), d_cells, ld, c->d_tab_ov, c->d_tab_g, c->d_tab_r, c->tab_nd, c->tab_ng,
                             c->d_rec};
        KLAUNCH(c, "compact_rows", (k_compact_rows<true, kCompactLanes>), grid_for(kCompactLanes * (int64_t)n, 256), 256, 0, n, stride,
                c->slot_cap, c->d_cnt64, c->d_own64, c->d_off64, c->d_slot_j, c->d_slot_rec, c->max_pairs, c->pi,
                c->pj, c->pdx, c->pdy, c->pdz, c->ps, c->pt, c->d_csr, c->d_jrow, c->d_m, c->d_status, ff);
    } else {
        KLAUNCH(c, "compact_rows", (k_compact_rows<false, kCompactLanes>), grid_for(kCompactLanes * (int64_t)n, 256), 256, 0, n, stride,
                c->slot_cap, c->d_cnt64, c->d_own64, c->d_off64, c->d_slot_j, c->d_slot_rec, c->max_pairs, c->pi,
                c->pj, c->pdx, c->pdy, c->pdz, c->ps, c->pt, c->d_csr, c->d_jrow, c->d_m, c->d_status,
                FusedForces{});
    }
    KLAUNCH(c, "compact_jrole", k_compact_jrole<kCompactLanes>, grid_for(kCompactLanes * (int64_t)n, 256), 256, 0, n, stride, c->slot_cap,
            c->d_cnt64, c->d_own64, c->d_off64, c->d_jslot, c->d_jrow, c->max_pairs, c->d_csr);
    c->h_m = -1;
    c->list_csr = true;
}

// Synchronous build: grows the per-cell slot capacity until the list fits.
void build_neighbors_sync(bfm_ctx* c, const double* d_cells, int n, int ld, double thr, double R,
                          double Ldiv)
{
    for (;;) {
        build_neighbors(c, d_cells, n, ld, thr, R, Ldiv);
        int st = take_status(c);
        if ((st & ST_SLOT_OVERFLOW) && !(st & ~ST_SLOT_OVERFLOW) && c->slot_cap < 4096) {
            grow_slots(c);
            continue;
        }
        raise_status(c, st);
        return;
    }
}

// Transposes of the current list: rows per cell as j (sorted), and as i when
// the list is not known to be sorted by i. k_pair_forces counts the j rows.
void build_transposes(bfm_ctx* c, int n);

void pair_forces_and_transpose(bfm_ctx* c, const double* d_cells, int ld, int n,
                               const double* sur_radius, bool refresh = false, const int* gate = nullptr)
{
    ForceParams fp = force_params(c);
    DistRefresh rf{};
    if (refresh) rf = DistRefresh{c->d_geom, n, c->pdx, c->pdy, c->pdz, c->ps, c->pt, gate};
    const bool need_t = !c->list_csr && !c->transposed;
    if (need_t)
        KLAUNCH(c, "fill", k_fill_int, grid_for(n + 1, 256, 4096), 256, 0, c->d_jcnt, 0, (int64_t)n + 1);
    if (c->p.adhesion_mode == BFM_ADHESION_JKR_ANISOTROPIC) {
        const AnisoJKR an = aniso_params(c);
        KLAUNCH(c, "pair_forces", k_pair_forces<true>, kPairGrid, 256, sizeof(double) * an.total, c->d_m, d_cells,
                ld, c->pi, c->pj, c->pdx, c->pdy, c->pdz, fp, c->d_tab_ov, c->d_tab_g, c->d_tab_r, c->tab_nd,
                c->tab_ng, an, (const double*)nullptr, 0, c->ps, c->pt, sur_radius, c->d_rec,
                need_t ? c->d_jcnt : (int*)nullptr, c->d_status, DistRefresh{});
    } else {
        KLAUNCH(c, "pair_forces", k_pair_forces<false>, kPairGrid, 256, 0, c->d_m, d_cells, ld, c->pi, c->pj,
                c->pdx, c->pdy, c->pdz, fp, c->d_tab_ov, c->d_tab_g, c->d_tab_r, c->tab_nd, c->tab_ng,
                AnisoJKR{}, (const double*)nullptr, 0, c->ps, c->pt, sur_radius, c->d_rec,
                need_t ? c->d_jcnt : (int*)nullptr, c->d_status, rf);
    }
    if (!need_t) return;
    build_transposes(c, n);
}

// Transposes of a list imported from the caller for the merge-form gather,
// after d_jcnt holds the rows per cell as j.
void build_transposes(bfm_ctx* c, int n)
{
    scan(c, c->d_jcnt, c->d_joff, (int64_t)n + 1, "scan_j");
    KLAUNCH(c, "fill", k_fill_int, grid_for(n + 1, 256, 4096), 256, 0, c->d_jcnt, 0, (int64_t)n + 1);
    KLAUNCH(c, "fill_perm", k_fill_perm, kPairGrid, 256, 0, c->d_m, c->pj, c->d_joff, c->d_jcnt, c->d_perm_j);
    KLAUNCH(c, "sort_segments", k_sort_segments, grid_for(n, 256), 256, 0, n, c->d_joff, c->d_perm_j);
    KLAUNCH(c, "fill", k_fill_int, grid_for(n + 1, 256, 4096), 256, 0, c->d_icnt, 0, (int64_t)n + 1);
    KLAUNCH(c, "count_i", k_count_key, kPairGrid, 256, 0, c->d_m, c->pi, c->d_icnt);
    scan(c, c->d_icnt, c->d_offsets, (int64_t)n + 1, "scan_i");
    if (!c->list_sorted_i) {
        KLAUNCH(c, "fill", k_fill_int, grid_for(n + 1, 256, 4096), 256, 0, c->d_icnt, 0, (int64_t)n + 1);
        KLAUNCH(c, "fill_perm", k_fill_perm, kPairGrid, 256, 0, c->d_m, c->pi, c->d_offsets, c->d_icnt,
                c->d_perm_i);
        KLAUNCH(c, "sort_segments", k_sort_segments, grid_for(n, 256), 256, 0, n, c->d_offsets, c->d_perm_i);
    }
    c->use_perm_i = !c->list_sorted_i;
    c->transposed = true;
}

RowLists row_lists(bfm_ctx* c)
{
    RowLists L;
    L.off64 = c->d_off64;
    L.csr = c->d_csr;
    L.cap = c->max_pairs;
    L.ioff = c->d_offsets;
    L.perm_i = c->use_perm_i ? c->d_perm_i : nullptr;
    L.joff = c->d_joff;
    L.perm_j = c->d_perm_j;
    return L;
}

bool friction_on(const bfm_ctx* c) { return c->p.friction_mode == 1; }

FricGather fric_gather(const bfm_ctx* c)
{
    FricGather fg;
    fg.frec = friction_on(c) ? c->d_frec : nullptr;
    fg.ps = c->ps;
    fg.pt = c->pt;
    return fg;
}

// Friction rows of the current list at the cells' own positions (zst ==
// nullptr) or at stage positions (zst: 2 x n_zst).
void launch_friction(bfm_ctx* c, const double* d_cells, int ld, const double* zst, int n_zst)
{
    FricParams fq;
    fq.R = c->p.R;
    fq.Req = c->p.R / 2;
    fq.E0 = c->p.E0;
    fq.coulomb = c->p.cell_cell_coulomb_coeff;
    fq.no_surface = c->p.no_surface;
    fq.colidx_eta = c->p.colidx_eta_cell_cell;
    KLAUNCH(c, "friction_rows", k_friction_rows, grid_for(c->max_pairs, 256, 4096), 256, 0, c->d_m, d_cells, ld,
            n_zst, zst, c->pi, c->pj, c->pdx, c->pdy, c->pdz, c->ps, c->pt, fq, c->d_frec);
}



// The interactions whose pair forces the compaction kernel can write: no
// adhesion, or isotropic JKR with the exact radius or a table (not the
// learned radius, which runs on the finished list, nor the anisotropic
// tables).
bool fused_forces_ok(const bfm_ctx* c)
{
    if (c->p.adhesion_mode == BFM_ADHESION_NONE) return true;
    if (c->p.adhesion_mode != BFM_ADHESION_JKR_ISOTROPIC) return false;
    if (c->p.jkr_mode == BFM_JKR_EXACT) return true;
    return (c->p.jkr_mode == BFM_JKR_TABLE_1D || c->p.jkr_mode == BFM_JKR_TABLE_2D) && c->d_tab_ov != nullptr;
}

// refresh: the list's distances are first brought to the cells' packed
// segments in c->d_geom (updateNeighborDistances), inside the first kernel
// that reads them (the pair forces, or the learned radius's staging).
void compute_forces(bfm_ctx* c, const double* d_cells, int n, int ld, bool refresh = false,
                    const int* gate = nullptr)
{
    const double* sur = nullptr;
    const bool sur_on = c->p.adhesion_mode == BFM_ADHESION_JKR_ISOTROPIC && c->p.jkr_mode == BFM_JKR_SURROGATE;
    if (refresh && c->p.adhesion_mode == BFM_ADHESION_JKR_ANISOTROPIC) {
        if (n > 0)
            KLAUNCH(c, "update_distances", k_update_distances_packed, kPairGrid, 256, 0, c->d_m, c->d_geom, n,
                    c->pi, c->pj, c->pdx, c->pdy, c->pdz, c->ps, c->pt);
        refresh = false;
    }
    if (sur_on) {
        require(c->sur != nullptr, BFM_ERR_ARG, "surrogate weights not set (bfm_set_surrogate)");
        // the learned radius's staging reads the list first and refreshes it;
        // its iterations are never enqueued ahead (no gate)
        require(gate == nullptr, BFM_ERR_ARG, "learned JKR radius with a gated iteration");
        DistRefresh rf{};
        if (refresh) rf = DistRefresh{c->d_geom, n, c->pdx, c->pdy, c->pdz, c->ps, c->pt, nullptr};
        sur = bfm::surrogate_pair_radii(c, d_cells, ld, n, rf);
        refresh = false;
    }
    if ((c->p.jkr_mode == BFM_JKR_TABLE_1D || c->p.jkr_mode == BFM_JKR_TABLE_2D) &&
        c->p.adhesion_mode == BFM_ADHESION_JKR_ISOTROPIC)
        require(c->d_tab_ov != nullptr, BFM_ERR_ARG, "JKR table not set (bfm_set_jkr_table)");
    pair_forces_and_transpose(c, d_cells, ld, n, sur, refresh, gate);
}

void launch_velocities(bfm_ctx* c, const double* d_cells, int n, int ld, const int* d_a2d,
                       const double* d_noise, double* d_vel, double* d_cc)
{
    VelParams vp = vel_params(c);
    RowLists L = row_lists(c);
    FricGather fg = fric_gather(c);
    if (fg.frec && d_vel) launch_friction(c, d_cells, ld, nullptr, 0);
    if (c->list_csr)
        KLAUNCH(c, "velocities", k_velocities<true>, grid_for(n, 256), 256, 0, n, d_cells, ld, d_a2d, d_noise,
                vp, L, c->d_rec, fg, d_vel, d_cc, c->d_status);
    else
        KLAUNCH(c, "velocities", k_velocities<false>, grid_for(n, 256), 256, 0, n, d_cells, ld, d_a2d, d_noise,
                vp, L, c->d_rec, fg, d_vel, d_cc, c->d_status);
}

// One RK stage's friction rows and velocities (single colony or batched).
void launch_stage(bfm_ctx* c, int n, const int* d_n, const double* d_cells, int ld, const double* d_noise,
                  const VelParams& vp, const Tableau& tb, int st, double dt, const int* d_rep, const double* d_dt,
                  const int* d_active, int grid, bool timed)
{
    RowLists L = row_lists(c);
    FricGather fg = fric_gather(c);
    const double* rec = nullptr;
    if (c->p.adhesion_mode == BFM_ADHESION_JKR_ANISOTROPIC) {
        // the pair records at the stage orientations (stage 0: the cells'),
        // recomputed in every attempt
        const AnisoJKR an = aniso_params(c);
        const double* nst = nullptr;
        if (st > 0) {
            hipLaunchKernelGGL(k_stage_n, dim3(grid), dim3(256), 0, c->stream, n, d_n, d_cells, ld, tb, st, dt,
                               d_rep, d_dt, d_active, c->d_k, c->d_nst, c->max_cells);
            HIPCHK(hipGetLastError());
            nst = c->d_nst;
        }
        hipLaunchKernelGGL(k_pair_forces<true>, dim3(kPairGrid), dim3(256), sizeof(double) * an.total, c->stream,
                           c->d_m, d_cells, ld, c->pi, c->pj, c->pdx, c->pdy, c->pdz, force_params(c),
                           c->d_tab_ov, c->d_tab_g, c->d_tab_r, c->tab_nd, c->tab_ng, an, nst, c->max_cells,
                           c->ps, c->pt, (const double*)nullptr, c->d_rec, (int*)nullptr, c->d_status,
                           DistRefresh{});
        HIPCHK(hipGetLastError());
        rec = c->d_rec;
    }
    if (fg.frec) {
        const double* zst = nullptr;
        if (st > 0) {
            hipLaunchKernelGGL(k_stage_z, dim3(grid), dim3(256), 0, c->stream, n, d_n, d_cells, ld, tb, st, dt,
                               d_rep, d_dt, d_active, c->d_k, c->d_zst, c->max_cells);
            HIPCHK(hipGetLastError());
            zst = c->d_zst;
        }
        FricParams fq;
        fq.R = c->p.R;
        fq.Req = c->p.R / 2;
        fq.E0 = c->p.E0;
        fq.coulomb = c->p.cell_cell_coulomb_coeff;
        fq.no_surface = c->p.no_surface;
        fq.colidx_eta = c->p.colidx_eta_cell_cell;
        hipLaunchKernelGGL(k_friction_rows, dim3(grid_for(c->max_pairs, 256, 4096)), dim3(256), 0, c->stream, c->d_m,
                           d_cells, ld, c->max_cells, zst, c->pi, c->pj, c->pdx, c->pdy, c->pdz, c->ps, c->pt, fq,
                           c->d_frec);
        HIPCHK(hipGetLastError());
    }
    if (timed) {
        if (c->list_csr)
            KLAUNCH(c, "rk_stage", k_rk_stage<true>, grid, 256, 0, n, d_n, d_cells, ld, c->d_assume2d, d_noise, vp,
                    tb, st, dt, d_rep, d_dt, d_active, c->d_cc, rec, L, fg, c->d_k, c->d_status);
        else
            KLAUNCH(c, "rk_stage", k_rk_stage<false>, grid, 256, 0, n, d_n, d_cells, ld, c->d_assume2d, d_noise, vp,
                    tb, st, dt, d_rep, d_dt, d_active, c->d_cc, rec, L, fg, c->d_k, c->d_status);
    } else {
        if (c->list_csr)
            hipLaunchKernelGGL(k_rk_stage<true>, dim3(grid), dim3(256), 0, c->stream, n, d_n, d_cells, ld,
                               c->d_assume2d, d_noise, vp, tb, st, dt, d_rep, d_dt, d_active, c->d_cc, rec, L,
                               fg, c->d_k, c->d_status);
        else
            hipLaunchKernelGGL(k_rk_stage<false>, dim3(grid), dim3(256), 0, c->stream, n, d_n, d_cells, ld,
                               c->d_assume2d, d_noise, vp, tb, st, dt, d_rep, d_dt, d_active, c->d_cc, rec, L,
                               fg, c->d_k, c->d_status);
        HIPCHK(hipGetLastError());
    }
}

void import_host_neighbors(bfm_ctx* c, const double* nb, int64_t m, int64_t ld_nb, int n)
{
    require(m >= 0 && m <= c->max_pairs, BFM_ERR_CAPACITY, "m exceeds max_pairs of the context");
    require(m == 0 || ld_nb >= m, BFM_ERR_ARG, "ld of neighbors < m");
    bool sorted = true;
    for (int64_t k = 0; k < m; ++k) {
        double vi = nb[k], vj = nb[ld_nb + k];
        require(vi >= 0 && vi < n && vj >= 0 && vj < n && vi == (int)vi && vj == (int)vj, BFM_ERR_ARG,
                "neighbor indices out of range");
        if (k > 0 && nb[k] < nb[k - 1]) sorted = false;
    }
    if (m > 0) {
        HIPCHK(hipMemcpy2DAsync(c->d_nb, sizeof(double) * m, nb, sizeof(double) * ld_nb, sizeof(double) * m, 7,
                                hipMemcpyHostToDevice, c->stream));
        KLAUNCH(c, "import_nb", k_import_nb, grid_for(m, 256, 4096), 256, 0, c->d_nb, m, m, c->pi, c->pj, c->pdx,
                c->pdy, c->pdz, c->ps, c->pt);
    }
    int mi = (int)m;
    HIPCHK(hipMemcpyAsync(c->d_m, &mi, sizeof(int), hipMemcpyHostToDevice, c->stream));
    c->h_m = m;
    c->list_n = n;
    c->list_sorted_i = sorted;
    c->transposed = false;
    c->list_csr = false;
}

void upload_cells(bfm_ctx* c, const double* cells, int n, int ld)
{
    require(c->p.ncols > 0 && c->p.ncols <= kMaxCols, BFM_ERR_ARG, "params.ncols must be in 1..32");
    if (n > 0)
        HIPCHK(hipMemcpy2DAsync(c->d_cells, sizeof(double) * c->max_cells, cells, sizeof(double) * ld,
                                sizeof(double) * n, c->p.ncols, hipMemcpyHostToDevice, c->stream));
}

void download6(bfm_ctx* c, const double* d, double* h, int n)
{
    if (n > 0)
        HIPCHK(hipMemcpy2DAsync(h, sizeof(double) * n, d, sizeof(double) * n, sizeof(double) * n, 6,
                                hipMemcpyDeviceToHost, c->stream));
}

Tableau make_tableau(const double* A, const double* b, const double* bs, int s)
{
    require(s >= 1 && s <= 7, BFM_ERR_ARG, "RK stages must be 1..7");
    Tableau tb;
    std::memset(&tb, 0, sizeof(tb));
    tb.s = s;
    for (int i = 0; i < s * s; ++i) tb.A[i] = A[i];
    for (int i = 0; i < s; ++i) { tb.b[i] = b[i]; tb.bs[i] = bs[i]; }
    return tb;
}

// updateNeighborDistances of the internal list for cells d_cells (n cells).
void update_distances(bfm_ctx* c, const double* d_cells, int ld, int n)
{
    if (n <= 0) return;
    KLAUNCH(c, "pack_segs", k_pack_segs, grid_for(n, 256), 256, 0, n, d_cells, ld, c->d_geom);
    KLAUNCH(c, "update_distances", k_update_distances_packed, kPairGrid, 256, 0, c->d_m, c->d_geom, n, c->pi,
            c->pj, c->pdx, c->pdy, c->pdz, c->ps, c->pt);
}

// RK step on device data with the internal list (stepRungeKuttaAdaptive).
void step_rk_device(bfm_ctx* c, const Tableau& tb, double dt, const double* d_cells, int n, int ld,
                    const double* d_noise, double* d_cells_new, int ld_new, double* d_errors, double tol,
                    double* d_err_scaled)
{
    need_params(c);
    require(c->list_n == n, BFM_ERR_ARG, "neighbor list was built for a different number of cells");
    compute_forces(c, d_cells, n, ld);
    KLAUNCH(c, "assume2d", k_assume2d, grid_for(n, 256), 256, 0, n, d_cells, ld, c->p.nz_threshold,
            c->d_assume2d);
    launch_velocities(c, d_cells, n, ld, c->d_assume2d, d_noise, nullptr, c->d_cc);
    VelParams vp = vel_params(c);
    for (int st = 0; st < tb.s; ++st)
        launch_stage(c, n, nullptr, d_cells, ld, d_noise, vp, tb, st, dt, nullptr, nullptr, nullptr,
                     grid_for(n, 256), true);
    KLAUNCH(c, "err_init", k_err_init, 1, 1, 0, c->d_errmax);
    KLAUNCH(c, "rk_final", k_rk_final, grid_for(n, 256), 256, 0, n, c->p.ncols, d_cells, ld, tb, dt, c->d_k,
            d_cells_new, ld_new, d_errors, tol, c->d_errmax);
    if (tol > 0 && d_err_scaled) KLAUNCH(c, "err_final", k_err_final, 1, 1, 0, c->d_errmax, d_err_scaled);
}

// stepVelocityVerlet on device data with the internal list
// (mechanics.hpp:2668-2996). The list's distances are updated to the half
// step (the reference's neighbors_new), so after the call the internal list
// describes d_cells_new. Work arrays: the generalized forces (6 x n) and the
// half-step angular velocities (3 x n) in the stage buffer.
void step_verlet_device(bfm_ctx* c, double dt, double rho, const double* d_cells, int n, int ld,
                        const double* d_noise, double* d_new, int ld_new)
{
    need_params(c);
    require(c->list_n == n, BFM_ERR_ARG, "neighbor list was built for a different number of cells");
    require(dt > 0 && rho > 0, BFM_ERR_ARG, "need dt > 0 and rho > 0");
    VerletParams P;
    P.vp = vel_params(c);
    P.rho = rho;
    P.dt = dt;
    P.R3 = std::pow(c->p.R, 3);
    P.ncols = c->p.ncols;
    double* F = c->d_k;
    double* wh = c->d_k + (int64_t)6 * n;
    RowLists L = row_lists(c);
    FricGather fg = fric_gather(c);
    auto forces = [&](const double* cells, int ldc) {
        compute_forces(c, cells, n, ldc);
        if (fg.frec) launch_friction(c, cells, ldc, nullptr, 0);
        if (c->list_csr)
            KLAUNCH(c, "verlet_forces", k_verlet_forces<true>, grid_for(n, 256), 256, 0, n, cells, ldc, d_noise, P, L,
                    c->d_rec, fg, F);
        else
            KLAUNCH(c, "verlet_forces", k_verlet_forces<false>, grid_for(n, 256), 256, 0, n, cells, ldc, d_noise, P,
                    L, c->d_rec, fg, F);
    };
    forces(d_cells, ld);
    KLAUNCH(c, "verlet_half", k_verlet_half, grid_for(n, 256), 256, 0, n, d_cells, ld, F, P, c->p.R, d_new, ld_new,
            wh);
    update_distances(c, d_new, ld_new, n);
    forces(d_new, ld_new);
    KLAUNCH(c, "verlet_final", k_verlet_final, grid_for(n, 256), 256, 0, n, d_new, ld_new, F, P, c->p.R, wh);
}

template <typename F>
int guarded(bfm_ctx* c, F&& f)
{
    if (!c) return BFM_ERR_ARG;
    try {
        HIPCHK(hipSetDevice(c->device));
        f();
        return BFM_OK;
    } catch (const ArgError& e) {
        c->err = e.msg;
        return e.code;
    } catch (const HipError& e) {
        c->err = std::string(e.what) + ": " + hipGetErrorString(e.e);
        return BFM_ERR_HIP;
    } catch (const std::exception& e) {
        c->err = e.what();
        return BFM_ERR_ARG;
    }
}


// ---------------------------------------------------------------------------
// the 2-D model (ABM-2d/ABM-2d-cpp): getVelocities and stepRungeKuttaAdaptive
// ---------------------------------------------------------------------------
// 2-D cell columns (include/indices.hpp of ABM-2d).
enum Col2D { D_RX = 1, D_RY = 2, D_NX = 3, D_NY = 4, D_DRX = 5, D_L = 9, D_HALF_L = 10, D_ETA0 = 13, D_ETA1 = 14 };

struct Params2D {
    double R, Rcell, p0, p1, p2, scd, strength, mindist, dexp, E0, gamma;
    int mode;
};

// One row's record for the 2-D stage kernel: {dx, dy, s, t, repulsion
// prefactor, adhesion magnitude, repulsion on, adhesion on}.
__device__ __forceinline__ void pair_record_2d(const Params2D& P, double dx, double dy, double s_i, double s_j,
                                               int adhere, double* __restrict__ rec)
{
    double2* r = reinterpret_cast<double2*>(rec);
    const double mag = sqrt(dx * dx + dy * dy);
    const double overlap = 2 * P.R - mag;
    const double cap = 2 * (P.R - P.Rcell);
    double pref = 0;
    if (overlap > 0 && overlap < cap) pref = P.p0 * pow(overlap, 1.5);
    else if (overlap >= cap) pref = P.p1 + P.p2 * pow(overlap - cap, 1.5);
    double a = 0, adh = 0;
    if (P.mode != BFM2D_ADHESION_NONE && adhere && overlap > 0) {
        adh = 1;
        if (P.mode == BFM2D_ADHESION_JKR) {
            a = kPi * P.R;
            a *= (mag <= P.mindist) ? (2 * P.R - P.mindist) : (2 * P.R - mag);
        } else if (P.mode == BFM2D_ADHESION_JKR_ISOTROPIC) {
            // forcesIsotropicJKRLagrange<T, 2> (adhesion.hpp:171-233 of
            // ABM-3d) at the overlap capped at 2 (R - Rcell): a = term2 - term1
            const double delta = overlap > cap ? cap : overlap;
            const double Req = P.R / 2;
            const double radius = jkr_radius_newton(delta, Req, P.E0, P.gamma, 0.0, 200, nullptr);
            a = -jkr_magnitude(radius, Req, P.E0, P.gamma);
        } else {
            const double t1 = (mag <= P.mindist) ? 1.0 / pow(P.mindist, P.dexp + 1) : 1.0 / pow(mag, P.dexp + 1);
            const double t2 = 1.0 / pow(2 * P.R, P.dexp + 1);
            a = P.dexp * (t1 - t2);
        }
    }
    r[0] = make_double2(dx, dy);
    r[1] = make_double2(s_i, s_j);
    r[2] = make_double2(pref, a);
    r[3] = make_double2(overlap > 0 ? 1.0 : 0.0, adh);
}

// Per-row magnitudes of cellCellRepulsiveForces and cellCellAdhesiveForces
// (mechanics.hpp:448-694 of ABM-2d), which read only the frozen distance:
// {dx, dy, s, t, repulsion prefactor, adhesion magnitude a (v = a d/|d|,
// before the strength factor), repulsion on, adhesion on}. The torques read
// the (stage) orientations and are formed in the per-cell gather.
__global__ void __launch_bounds__(256)
k2d_pair(const int* __restrict__ d_m, int n, const int* __restrict__ pi, const int* __restrict__ pj,
         const double* __restrict__ pdx, const double* __restrict__ pdy, const double* __restrict__ ps,
         const double* __restrict__ pt, const int* __restrict__ to_adhere, Params2D P, double* __restrict__ rec,
         int* __restrict__ jcnt)
{
    const int m = *d_m;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        double2* r = reinterpret_cast<double2*>(rec + (int64_t)k * 8);
        const int i = pi[k], j = pj[k];
        if ((unsigned)i >= (unsigned)n || (unsigned)j >= (unsigned)n) {
            r[0] = r[1] = r[2] = r[3] = make_double2(0.0, 0.0);
            continue;
        }
        if (jcnt) atomicAdd(&jcnt[j], 1);
        pair_record_2d(P, pdx[k], pdy[k], ps[k], pt[k], P.mode != BFM2D_ADHESION_NONE ? to_adhere[k] : 0,
                       rec + (int64_t)k * 8);
    }
}

// A cell's coordinates at stage st (cells + dt sum_j A(st, j) k_j, then the
// 2-D normalizeOrientations, mechanics.hpp:1458-1465 of ABM-2d: no flip).
__device__ __forceinline__ void stage_coords_2d(const double* __restrict__ cells, int ld, int n, int c,
                                                const double* __restrict__ kbuf, const Tableau& tb, int st,
                                                double dt, double y[4])
{
#pragma unroll
    for (int q = 0; q < 4; ++q) y[q] = cellv(cells, ld, c, D_RX + q);
    if (st == 0) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double mult = 0;
        for (int j = 0; j < st; ++j) mult += kbuf[((int64_t)j * 4 + q) * n + c] * tb.A[st * tb.s + j];
        y[q] += mult * dt;
    }
    const double nrm = sqrt(y[2] * y[2] + y[3] * y[3]);
    y[2] /= nrm;
    y[3] /= nrm;
}

// One stage of the 2-D stepRungeKuttaAdaptive (mechanics.hpp:1770-1792 of
// ABM-2d), or getVelocities alone (tb.s == 0 form: st = 0, out = kbuf):
// repulsion and adhesion summed per cell in the order of the rows (the
// reference's two sequential loops), the constrained torque forms with the
// stage orientation, (repulsion + adhesion) + noise, then -dE/dq divided by
// the composite drag l (eta0 + eta1 scd / R) and l^3 (...) / 12.
template <bool CSR>
__global__ void __launch_bounds__(256)
k2d_stage(int n, const double* __restrict__ cells, int ld, const double* __restrict__ noise, Params2D P, Tableau tb,
          int st, double dt, const double* __restrict__ rec, RowLists L, double* __restrict__ kbuf,
          int* __restrict__ status)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        double y[4];
        stage_coords_2d(cells, ld, n, c, kbuf, tb, st, dt, y);
        const double nx = y[2], ny = y[3];
        double rep[4] = {0, 0, 0, 0}, adh[4] = {0, 0, 0, 0};
        for_rows<CSR>(c, L, [&](int row, bool role_i) {
            const double2* r = reinterpret_cast<const double2*>(rec + (int64_t)row * 8);
            const double2 d = r[0], st2 = r[1], mg = r[2], fl = r[3];
            if (fl.x == 0) return;   // no overlap: neither term
            const double mag = sqrt(d.x * d.x + d.y * d.y);
            const double ux = d.x / mag, uy = d.y / mag;
            {
                const double vx = mg.x * ux, vy = mg.x * uy;
                const double w = nx * vx + ny * vy;
                if (role_i) {
                    const double s = st2.x;
                    rep[0] += vx; rep[1] += vy;
                    rep[2] += s * (-w * nx + vx);
                    rep[3] += s * (-w * ny + vy);
                } else {
                    const double t = st2.y;
                    rep[0] += -vx; rep[1] += -vy;
                    rep[2] += t * (w * nx - vx);
                    rep[3] += t * (w * ny - vy);
                }
            }
            if (fl.y != 0) {
                const double vx = mg.y * ux, vy = mg.y * uy;
                const double w = nx * (-vx) + ny * (-vy);
                double f[4];
                if (role_i) {
                    const double s = st2.x;
                    f[0] = -vx; f[1] = -vy;
                    f[2] = s * (-w * nx - vx);
                    f[3] = s * (-w * ny - vy);
                } else {
                    const double t = st2.y;
                    f[0] = vx; f[1] = vy;
                    f[2] = t * (w * nx + vx);
                    f[3] = t * (w * ny + vy);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) adh[q] += P.strength * f[q];
            }
        });
        const double l = cellv(cells, ld, c, D_L);
        const double drag = cellv(cells, ld, c, D_ETA0) + cellv(cells, ld, c, D_ETA1) * P.scd / P.R;
        const double K = l * drag, Lr = l * l * l * drag / 12;
        bool bad = false;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double e = rep[q] + adh[q];
            if (noise) e += noise[(int64_t)q * n + c];
            const double v = -e / (q < 2 ? K : Lr);
            kbuf[((int64_t)st * 4 + q) * n + c] = v;
            bad |= !isfinite(v);
        }
        if (bad) atomicOr(status, ST_NAN);
    }
}

// Growth of the 2-D loop fused into the accepted RK update: growCells
// (growth.hpp:51-57 of ABM-2d) on the new cells and divideMaxLength's count
// (l > Ldiv, growth.hpp:67-71) into *count.
struct Grow2D {
    int on;
    double R, Ldiv, dt;
    int* count;
};

// RK update of the 2-D step (mechanics.hpp:1794-1816 of ABM-2d); errors may
// be NULL; tol > 0 also reduces the controller's scaled error
// max e / (tol (1 + |y|)) over the four coordinates (simulation.hpp:639-650
// of ABM-2d) into errmax (ordered key); gr.on applies the growth step.
__global__ void __launch_bounds__(256)
k2d_final(int n, int ncols, const double* __restrict__ cells, int ld, Tableau tb, double dt,
          const double* __restrict__ kbuf, double* __restrict__ out, int ld_out, double* __restrict__ errors,
          double tol, unsigned long long* __restrict__ errmax, Grow2D gr)
{
    double local = -INFINITY;
    int divide = 0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        for (int col = 0; col < ncols; ++col)
            if (col < D_RX || col > D_DRX + 3) out[(int64_t)col * ld_out + c] = cellv(cells, ld, c, col);
        if (gr.on) {
            double l = cellv(cells, ld, c, D_L);
            const double g = cellv(cells, ld, c, 12);
            l = l + (g * (4 * gr.R / 3 + l)) * gr.dt;
            out[(int64_t)D_L * ld_out + c] = l;
            out[(int64_t)D_HALF_L * ld_out + c] = l / 2;
            divide += l > gr.Ldiv;
        }
        double y[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double v1 = 0, v2 = 0;
            for (int st = 0; st < tb.s; ++st) {
                const double kv = kbuf[((int64_t)st * 4 + q) * n + c];
                v1 += kv * tb.b[st];
                v2 += kv * tb.bs[st];
            }
            const double d1 = v1 * dt, d2 = v2 * dt;
            const double y0 = cellv(cells, ld, c, D_RX + q);
            y[q] = y0 + d1;
            if (errors) errors[(int64_t)q * n + c] = d1 - d2;
            if (tol > 0) local = fmax(local, (d1 - d2) / (tol * (1.0 + fabs(y0))));
            out[(int64_t)(D_DRX + q) * ld_out + c] = v1;
        }
        const double nrm = sqrt(y[2] * y[2] + y[3] * y[3]);
        y[2] /= nrm;
        y[3] /= nrm;
#pragma unroll
        for (int q = 0; q < 4; ++q) out[(int64_t)(D_RX + q) * ld_out + c] = y[q];
    }
    if (tol > 0) {
        for (int off = 32; off > 0; off >>= 1) local = fmax(local, __shfl_xor(local, off));
        if ((threadIdx.x & 63) == 0) atomicMax(errmax, order_key(local));
    }
    if (gr.on) {
        for (int off = 32; off > 0; off >>= 1) divide += __shfl_xor(divide, off);
        if ((threadIdx.x & 63) == 0 && divide) atomicAdd(gr.count, divide);
    }
}

// updateNeighborDistances of the 2-D tree (mechanics.hpp:376-446 of ABM-2d)
// on the internal list's rows from the 2-D columns (the segments in the
// z = 0 plane, as the neighbor search sees them), followed by each row's
// pair record (k2d_pair) for the next step.
__global__ void __launch_bounds__(256)
k2d_rows_update(const int* __restrict__ d_m, int n, const double* __restrict__ cells, int ld,
                const int* __restrict__ pi, const int* __restrict__ pj, double* __restrict__ pdx,
                double* __restrict__ pdy, double* __restrict__ pdz, double* __restrict__ ps, double* __restrict__ pt,
                const int* __restrict__ to_adhere, Params2D P, double* __restrict__ rec)
{
    const int m = *d_m;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        const int i = pi[k], j = pj[k];
        if ((unsigned)i >= (unsigned)n || (unsigned)j >= (unsigned)n) {
            double2* r = reinterpret_cast<double2*>(rec + (int64_t)k * 8);
            r[0] = r[1] = r[2] = r[3] = make_double2(0.0, 0.0);
            continue;
        }
        Seg a, b;
        a.r[0] = cellv(cells, ld, i, D_RX); a.r[1] = cellv(cells, ld, i, D_RY); a.r[2] = 0.0;
        a.n[0] = cellv(cells, ld, i, D_NX); a.n[1] = cellv(cells, ld, i, D_NY); a.n[2] = 0.0;
        a.hl = cellv(cells, ld, i, D_HALF_L);
        b.r[0] = cellv(cells, ld, j, D_RX); b.r[1] = cellv(cells, ld, j, D_RY); b.r[2] = 0.0;
        b.n[0] = cellv(cells, ld, j, D_NX); b.n[1] = cellv(cells, ld, j, D_NY); b.n[2] = 0.0;
        b.hl = cellv(cells, ld, j, D_HALF_L);
        double d[3], si, sj;
        dist_between_cells(a, b, d, si, sj);
        pdx[k] = d[0]; pdy[k] = d[1]; pdz[k] = d[2]; ps[k] = si; pt[k] = sj;
        pair_record_2d(P, d[0], d[1], si, sj, P.mode != BFM2D_ADHESION_NONE ? to_adhere[k] : 0,
                       rec + (int64_t)k * 8);
    }
}

// Frozen-geometry sums of the 2-D forces for the whole RK step. The pair
// geometry (d, s, t) of every row is frozen during a step and only the
// cell's own stage orientation n enters its constrained torques, each of
// the form c (v - (n . v) n) (cellCellRepulsiveForces / the adhesive loop,
// mechanics.hpp:448-694 of ABM-2d). Summed over a cell's rows in row order
// that is V - (n . V) n with V = sum c v, so per cell and step four vectors
// carry all of them: the repulsive and adhesive position forces and their
// torque sums V (adhesion before the strength factor). sums: 8 doubles per
// cell {Frep, Vrep, Fadh, Vadh}.
template <bool CSR>
__global__ void __launch_bounds__(256)
k2d_gather(int n, const double* __restrict__ rec, RowLists L, double* __restrict__ sums)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        double f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for_rows<CSR>(c, L, [&](int row, bool role_i) {
            const double2* r = reinterpret_cast<const double2*>(rec + (int64_t)row * 8);
            const double2 d = r[0], st2 = r[1], mg = r[2], fl = r[3];
            if (fl.x == 0) return;
            const double mag = sqrt(d.x * d.x + d.y * d.y);
            const double ux = d.x / mag, uy = d.y / mag;
            const double vx = mg.x * ux, vy = mg.x * uy;
            if (role_i) {           // +v, s (v - (n.v) n)
                f[0] += vx; f[1] += vy;
                f[2] += st2.x * vx; f[3] += st2.x * vy;
            } else {                // -v, -t (v - (n.v) n)
                f[0] += -vx; f[1] += -vy;
                f[2] += -(st2.y * vx); f[3] += -(st2.y * vy);
            }
            if (fl.y != 0) {
                const double ax = mg.y * ux, ay = mg.y * uy;
                if (role_i) {       // -v, -s (v - (n.v) n)
                    f[4] += -ax; f[5] += -ay;
                    f[6] += -(st2.x * ax); f[7] += -(st2.x * ay);
                } else {            // +v, +t (v - (n.v) n)
                    f[4] += ax; f[5] += ay;
                    f[6] += st2.y * ax; f[7] += st2.y * ay;
                }
            }
        });
        double2* o = reinterpret_cast<double2*>(sums + (int64_t)c * 8);
        o[0] = make_double2(f[0], f[1]);
        o[1] = make_double2(f[2], f[3]);
        o[2] = make_double2(f[4], f[5]);
        o[3] = make_double2(f[6], f[7]);
    }
}

// One whole RK attempt of the 2-D step per cell (stepRungeKuttaAdaptive,
// mechanics.hpp:1710-1817 of ABM-2d) from the frozen-geometry sums: the S
// stages in registers (stage coordinates, 2-D normalizeOrientations,
// (repulsion + adhesion) + noise, the composite drag), then the update of
// k2d_final (error key, growth and division count as there). Cells are
// independent within a step, so the attempt is one launch.
template <int S>
__global__ void __launch_bounds__(256)
k2d_attempt(int n, int ncols, const double* __restrict__ cells, int ld, const double* __restrict__ sums,
            const double* __restrict__ noise, double strength, double scd, double R, Tableau tb, double dt,
            double* __restrict__ out, int ld_out, double* __restrict__ errors, double tol,
            unsigned long long* __restrict__ errmax, Grow2D gr, int* __restrict__ status)
{
    double local = -INFINITY;
    int divide = 0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const double2* sp = reinterpret_cast<const double2*>(sums + (int64_t)c * 8);
        const double2 Fr = sp[0], Vr = sp[1], Fa = sp[2], Va = sp[3];
        double y0[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) y0[q] = cellv(cells, ld, c, D_RX + q);
        double nz[4] = {0, 0, 0, 0};
        if (noise) {
#pragma unroll
            for (int q = 0; q < 4; ++q) nz[q] = noise[(int64_t)q * n + c];
        }
        const double l = cellv(cells, ld, c, D_L);
        const double drag = cellv(cells, ld, c, D_ETA0) + cellv(cells, ld, c, D_ETA1) * scd / R;
        const double K = l * drag, Lr = l * l * l * drag / 12;
        double k[S][4];
        bool bad = false;
#pragma unroll
        for (int st = 0; st < S; ++st) {
            double y[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                y[q] = y0[q];
                if (st > 0) {
                    double mult = 0;
#pragma unroll
                    for (int j = 0; j < st; ++j) mult += k[j][q] * tb.A[st * S + j];
                    y[q] += mult * dt;
                }
            }
            if (st > 0) {
                const double nrm = sqrt(y[2] * y[2] + y[3] * y[3]);
                y[2] /= nrm;
                y[3] /= nrm;
            }
            const double nx = y[2], ny = y[3];
            const double pr = nx * Vr.x + ny * Vr.y, pa = nx * Va.x + ny * Va.y;
            const double e[4] = {Fr.x + strength * Fa.x, Fr.y + strength * Fa.y,
                                 (Vr.x - pr * nx) + strength * (Va.x - pa * nx),
                                 (Vr.y - pr * ny) + strength * (Va.y - pa * ny)};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double v = -(e[q] + nz[q]) / (q < 2 ? K : Lr);
                k[st][q] = v;
                bad |= !isfinite(v);
            }
        }
        if (bad) atomicOr(status, ST_NAN);
        for (int col = 0; col < ncols; ++col)
            if (col < D_RX || col > D_DRX + 3) out[(int64_t)col * ld_out + c] = cellv(cells, ld, c, col);
        if (gr.on) {
            const double g = cellv(cells, ld, c, 12);
            const double l1 = l + (g * (4 * gr.R / 3 + l)) * gr.dt;
            out[(int64_t)D_L * ld_out + c] = l1;
            out[(int64_t)D_HALF_L * ld_out + c] = l1 / 2;
            divide += l1 > gr.Ldiv;
        }
        double y[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double v1 = 0, v2 = 0;
#pragma unroll
            for (int st = 0; st < S; ++st) {
                v1 += k[st][q] * tb.b[st];
                v2 += k[st][q] * tb.bs[st];
            }
            const double d1 = v1 * dt, d2 = v2 * dt;
            y[q] = y0[q] + d1;
            if (errors) errors[(int64_t)q * n + c] = d1 - d2;
            if (tol > 0) local = fmax(local, (d1 - d2) / (tol * (1.0 + fabs(y0[q]))));
            out[(int64_t)(D_DRX + q) * ld_out + c] = v1;
        }
        const double nrm = sqrt(y[2] * y[2] + y[3] * y[3]);
        y[2] /= nrm;
        y[3] /= nrm;
#pragma unroll
        for (int q = 0; q < 4; ++q) out[(int64_t)(D_RX + q) * ld_out + c] = y[q];
    }
    if (tol > 0) {
        for (int off = 32; off > 0; off >>= 1) local = fmax(local, __shfl_xor(local, off));
        if ((threadIdx.x & 63) == 0) atomicMax(errmax, order_key(local));
    }
    if (gr.on) {
        for (int off = 32; off > 0; off >>= 1) divide += __shfl_xor(divide, off);
        if ((threadIdx.x & 63) == 0 && divide) atomicAdd(gr.count, divide);
    }
}

// Readback of one 2-D loop attempt: the scaled-error maximum, the division
// count and the status word into host-mapped memory; the maximum and the
// count start again from their identities.
__global__ void k2d_publish(unsigned long long* __restrict__ errmax, int* __restrict__ count,
                            const int* __restrict__ status, unsigned long long* __restrict__ host)
{
    if (threadIdx.x == 0) {
        host[0] = *errmax;
        host[1] = (unsigned long long)(unsigned)*count;
        host[2] = (unsigned long long)(unsigned)*status;
        *errmax = order_key(-INFINITY);
        *count = 0;
    }
}

// Places 2-D cells in the z = 0 plane in the 3-D layout the neighbor search
// reads (r, n, l, half_l).
__global__ void k2d_embed(int n, const double* __restrict__ c2, int ld2, double* __restrict__ c3, int ld3)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        c3[(int64_t)C_ID * ld3 + c] = c2[c];
        c3[(int64_t)C_RX * ld3 + c] = c2[(int64_t)D_RX * ld2 + c];
        c3[(int64_t)C_RY * ld3 + c] = c2[(int64_t)D_RY * ld2 + c];
        c3[(int64_t)C_RZ * ld3 + c] = 0.0;
        c3[(int64_t)C_NX * ld3 + c] = c2[(int64_t)D_NX * ld2 + c];
        c3[(int64_t)C_NY * ld3 + c] = c2[(int64_t)D_NY * ld2 + c];
        c3[(int64_t)C_NZ * ld3 + c] = 0.0;
        c3[(int64_t)C_L * ld3 + c] = c2[(int64_t)D_L * ld2 + c];
        c3[(int64_t)C_HALF_L * ld3 + c] = c2[(int64_t)D_HALF_L * ld2 + c];
    }
}

// to_adhere of the 2-D driver (simulation.hpp:242-262 of ABM-2d) for the
// rows of the internal list: the pair's groups (min, max) adhere in the map
// and R + Rcell < |d| < 2R.
__global__ void k2d_adhere(const int* __restrict__ d_m, const double* __restrict__ cells, int ld,
                           const int* __restrict__ pi, const int* __restrict__ pj, const double* __restrict__ pdx,
                           const double* __restrict__ pdy, const int* __restrict__ amap, int ng, double R,
                           double Rcell, int* __restrict__ adhere)
{
    const int m = *d_m;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        const int i = pi[k], j = pj[k];
        int flag = 0;
        if ((unsigned)i < (unsigned)ld && (unsigned)j < (unsigned)ld && amap) {
            const int gi = (int)cellv(cells, ld, i, 16), gj = (int)cellv(cells, ld, j, 16);
            const int g0 = min(gi, gj), g1 = max(gi, gj);
            const double dx = pdx[k], dy = pdy[k];
            const double dist = sqrt(dx * dx + dy * dy);
            flag = g0 >= 0 && g1 < ng && amap[g0 * ng + g1] != 0 && dist > R + Rcell && dist < 2 * R;
        }
        adhere[k] = flag;
    }
}

Params2D params_2d(const bfm_ctx* c)
{
    require(c->p2d_set, BFM_ERR_ARG, "bfm2d_set_params has not been called");
    const bfm2d_params& q = c->p2d;
    Params2D P;
    P.R = q.R; P.Rcell = q.Rcell;
    P.p0 = q.prefactors[0]; P.p1 = q.prefactors[1]; P.p2 = q.prefactors[2];
    P.scd = q.surface_contact_density;
    P.strength = q.strength; P.mindist = q.mindist; P.dexp = q.distance_exp;
    P.E0 = q.E0; P.gamma = q.gamma;
    P.mode = q.adhesion_mode;
    return P;
}

// Host arrays in: cells to c->d_cells, the 6-column list as the internal
// list (dz = 0), to_adhere and noise to device buffers; then the pair
// records and the transposes of the list.
void prepare_2d(bfm_ctx* c, const double* cells, int n, int ld, const double* nb, int64_t m, int64_t ld_nb,
                const int* to_adhere, const double* noise)
{
    require(n >= 0 && n <= c->max_cells, BFM_ERR_CAPACITY, "n exceeds max_cells of the context");
    require(ld >= n, BFM_ERR_ARG, "ld < n");
    const int ncols = c->p2d.ncols;
    if (n > 0)
        HIPCHK(hipMemcpy2DAsync(c->d_cells, sizeof(double) * c->max_cells, cells, sizeof(double) * ld,
                                sizeof(double) * n, ncols, hipMemcpyHostToDevice, c->stream));
    require(m >= 0 && m <= c->max_pairs, BFM_ERR_CAPACITY, "m exceeds max_pairs of the context");
    require(m == 0 || ld_nb >= m, BFM_ERR_ARG, "ld of neighbors < m");
    std::vector<double> nb7((size_t)7 * std::max<int64_t>(m, 1), 0.0);
    const int src[7] = {0, 1, 2, 3, -1, 4, 5};
    for (int q = 0; q < 7; ++q)
        if (src[q] >= 0)
            for (int64_t k = 0; k < m; ++k) nb7[(size_t)q * m + k] = nb[(size_t)src[q] * ld_nb + k];
    import_host_neighbors(c, nb7.data(), m, std::max<int64_t>(m, 1), n);
    const bool adh = c->p2d.adhesion_mode != BFM2D_ADHESION_NONE;
    require(!adh || m == 0 || to_adhere != nullptr, BFM_ERR_ARG, "to_adhere is required with adhesion");
    if (adh && m > 0)
        HIPCHK(hipMemcpyAsync(c->d_adhere, to_adhere, sizeof(int) * m, hipMemcpyHostToDevice, c->stream));
    if (noise && n > 0)
        HIPCHK(hipMemcpyAsync(c->d_noise, noise, sizeof(double) * 4 * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));   // the staged host arrays (nb7) end with this call
    const Params2D P = params_2d(c);
    KLAUNCH(c, "fill", k_fill_int, grid_for(n + 1, 256, 4096), 256, 0, c->d_jcnt, 0, (int64_t)n + 1);
    KLAUNCH(c, "pair_2d", k2d_pair, kPairGrid, 256, 0, c->d_m, n, c->pi, c->pj, c->pdx, c->pdy, c->ps, c->pt,
            c->d_adhere, P, c->d_rec, c->d_jcnt);
    build_transposes(c, n);
}

// The s stages of the 2-D step (s == 1 with tb.A = 0: getVelocities) on
// cells d_cells (leading dimension ld), with the current list's row lists.
void stages_2d(bfm_ctx* c, const Tableau& tb, double dt, const double* d_cells, int ld, int n,
               const double* d_noise)
{
    const Params2D P = params_2d(c);
    const RowLists L = row_lists(c);
    for (int st = 0; st < tb.s; ++st) {
        if (c->list_csr)
            KLAUNCH(c, "stage_2d", k2d_stage<true>, grid_for(n, 256), 256, 0, n, d_cells, ld, d_noise, P, tb, st,
                    dt, c->d_rec, L, c->d_k, c->d_status);
        else
            KLAUNCH(c, "stage_2d", k2d_stage<false>, grid_for(n, 256), 256, 0, n, d_cells, ld, d_noise, P, tb, st,
                    dt, c->d_rec, L, c->d_k, c->d_status);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// library-internal entry points (bfm_internal.hpp)
// ---------------------------------------------------------------------------
namespace bfm {

constexpr int kBatchGrid = 2048;   // fixed grid of the graph-captured batched kernels
// Grids of at most this many work-groups run the controller's steps in
// their last work-group (a ticket after a device-scope release per group)
// instead of a launch of their own.
constexpr int kCtrlFuseGroups = 16;

int step_verlet(bfm_ctx* c, double dt, double rho, const double* d_cells, int n, int ld, double* d_new, int ld_new,
                int* d_bad)
{
    return guarded(c, [&] {
        check_cells(c, n, ld);
        step_verlet_device(c, dt, rho, d_cells, n, ld, nullptr, d_new, ld_new);
        KLAUNCH(c, "nonfinite", k_nonfinite, grid_for(n, 256), 256, 0, n, d_new, ld_new, d_bad);
    });
}

// Whether a batched RK attempt runs as one launch (k_rk_attempt_b): every
// stage's cell-cell forces are the step's frozen sums (no anisotropic
// tables, no friction).
bool attempt_fused(const bfm_ctx* c)
{
    return c->p.adhesion_mode != BFM_ADHESION_JKR_ANISOTROPIC && !friction_on(c);
}

int step_prepare(bfm_ctx* c, const double* d_cells, int n, int ld, bool refresh_distances, bool* cc_in_attempt,
                 const int* d_gate)
{
    return guarded(c, [&] {
        need_params(c);
        check_cells(c, n, ld);
        require(c->list_n == n, BFM_ERR_ARG, "neighbor list was built for a different number of cells");
        require(!d_gate || (refresh_distances && iteration_ahead_ok(c, n)), BFM_ERR_ARG,
                "a gated step needs the one-kernel attempts and the list's refresh");
        compute_forces(c, d_cells, n, ld, refresh_distances, d_gate);
        const bool fused = attempt_fused(c);
        if (!fused)   // the one-launch attempt tests nz itself
            KLAUNCH(c, "assume2d", k_assume2d, grid_for(n, 256), 256, 0, n, d_cells, ld, c->p.nz_threshold,
                    c->d_assume2d);
        // the gather inside the first attempt saves a launch for small
        // batches; large ones keep the separate, fully parallel gather
        *cc_in_attempt = fused && c->list_csr && grid_for(n, 256) <= kCtrlFuseGroups;
        if (!*cc_in_attempt) launch_velocities(c, d_cells, n, ld, c->d_assume2d, nullptr, nullptr, c->d_cc);
    });
}

int attempt_direct(bfm_ctx* c, const Tableau& tb, const double* d_cells, int ld, const int* d_n, int n,
                   const int* d_rep, bool gather, double* d_cells_new, double tol, const CtrlAfter& A,
                   unsigned* d_ticket)
{
    return guarded(c, [&] {
        require(attempt_fused(c), BFM_ERR_ARG, "this interaction needs the staged attempt");
        require(!gather || c->list_csr, BFM_ERR_ARG, "the cell-cell gather needs a device-built list");
        require(tol > 0, BFM_ERR_ARG, "tol must be > 0");
        const VelParams vp = vel_params(c);
        // 256-row slices, several per work-group above kBatchGrid groups
        // (faster than one group per slice on large batches)
        const int grid = grid_for(n, 256, kBatchGrid);
        // the controller's step in the attempt's last work-group for small
        // grids; a launch of its own otherwise
        unsigned* ticket = grid <= kCtrlFuseGroups ? d_ticket : nullptr;
        auto launch = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(256), 0, c->stream, d_n, d_cells, ld, vp, tb, d_rep, A.trial,
                               A.active, row_lists(c), c->d_rec, c->d_cc, c->d_k, d_cells_new, tol, A.errmax, A.bad,
                               A, ticket);
        };
        // the stage counts of the reference's tableaux with the stage
        // velocities in registers
        switch (tb.s * 2 + (gather ? 1 : 0)) {
        case 4: launch(k_rk_attempt_b<false, 2>); break;
        case 5: launch(k_rk_attempt_b<true, 2>); break;
        case 8: launch(k_rk_attempt_b<false, 4>); break;
        case 9: launch(k_rk_attempt_b<true, 4>); break;
        case 12: launch(k_rk_attempt_b<false, 6>); break;
        case 13: launch(k_rk_attempt_b<true, 6>); break;
        case 14: launch(k_rk_attempt_b<false, 7>); break;
        case 15: launch(k_rk_attempt_b<true, 7>); break;
        default:
            if (gather) launch(k_rk_attempt_b<true, 0>);
            else launch(k_rk_attempt_b<false, 0>);
        }
        if (!ticket) hipLaunchKernelGGL(k_ctrl_after_step, dim3(1), dim3(256), 0, c->stream, A);
        HIPCHK(hipGetLastError());
    });
}

int step_attempt_batched(bfm_ctx* c, const Tableau& tb, const double* d_cells, int ld, const int* d_n,
                         const int* d_rep, const double* d_dt, const int* d_active, int n_rep,
                         double* d_cells_new, double tol, unsigned long long* d_errmax, int* d_bad)
{
    return guarded(c, [&] {
        require(tol > 0, BFM_ERR_ARG, "tol must be > 0");
        require(n_rep >= 1, BFM_ERR_ARG, "n_rep must be >= 1");
        VelParams vp = vel_params(c);
        if (attempt_fused(c)) {
            hipLaunchKernelGGL((k_rk_attempt_b<false, 0>), dim3(kBatchGrid), dim3(256), 0, c->stream, d_n, d_cells,
                               ld, vp, tb, d_rep, d_dt, d_active, RowLists{}, (const double*)nullptr, c->d_cc,
                               c->d_k, d_cells_new, tol, d_errmax, d_bad, CtrlAfter{}, (unsigned*)nullptr);
        } else {
            for (int st = 0; st < tb.s; ++st)
                launch_stage(c, 0, d_n, d_cells, ld, nullptr, vp, tb, st, 0.0, d_rep, d_dt, d_active, kBatchGrid,
                             false);
            hipLaunchKernelGGL(k_rk_final_b, dim3(kBatchGrid), dim3(256), 0, c->stream, d_n, c->p.ncols, d_cells,
                               ld, tb, d_rep, d_dt, d_active, c->d_k, d_cells_new, tol, d_errmax, d_bad);
        }
        HIPCHK(hipGetLastError());
    });
}

bool iteration_ahead_ok(const bfm_ctx* c, int n)
{
    return attempt_fused(c) && c->list_csr && c->p.jkr_mode != BFM_JKR_SURROGATE &&
           grid_for(n, 256) <= kCtrlFuseGroups;
}

int iteration_tail_batched(bfm_ctx* c, const double* d_cells, int ld, const int* d_n, int n, const int* d_rep,
                           const CtrlEnd& E, double R, double Ldiv, double* d_cells_new, int* d_count, int* d_idx,
                           const int* d_skip, bool pack_segments, const int* d_gate, IterSnap* d_snap,
                           int* d_spec, unsigned* d_ticket)
{
    return guarded(c, [&] {
        check_cells(c, n, ld);
        const int grid = grid_for(n, 256);
        unsigned* ticket = grid <= kCtrlFuseGroups ? d_ticket : nullptr;
        hipLaunchKernelGGL(k_iter_tail_b, dim3(grid), dim3(256), 0, c->stream, d_n, c->p.ncols, d_cells, ld, d_rep,
                           E, R, Ldiv, d_cells_new, d_count, d_idx, pack_segments ? c->d_geom : (double*)nullptr,
                           d_skip, d_gate, d_snap, d_spec, (const int*)c->d_status, ticket);
        if (!ticket)
            hipLaunchKernelGGL(k_publish_tail, dim3(1), dim3(1), 0, c->stream, (const int*)d_count,
                               (const int*)c->d_status, E, d_skip, d_gate, d_snap, d_spec);
        HIPCHK(hipGetLastError());
    });
}

int grow_batched(bfm_ctx* c, double* d_cells, int ld, const int* d_n, double R, const int* d_rep,
                 const double* d_dt, const int* d_active)
{
    return guarded(c, [&] {
        hipLaunchKernelGGL(k_grow_b, dim3(kBatchGrid), dim3(256), 0, c->stream, d_n, d_cells, ld, R, d_rep, d_dt,
                           d_active);
        HIPCHK(hipGetLastError());
    });
}

int set_replicate_count(bfm_ctx* c, int n_rep)
{
    return guarded(c, [&] {
        require(n_rep >= 0, BFM_ERR_ARG, "n_rep < 0");
        if (n_rep > c->max_reps) {
            HIPCHK(hipStreamSynchronize(c->stream));
            for (void* q : {(void*)c->d_grid, (void*)c->d_rep_box})
                if (q) HIPCHK(hipFree(q));
            c->d_grid = nullptr;
            c->d_rep_box = nullptr;
            dalloc(&c->d_grid, (size_t)n_rep);
            dalloc(&c->d_rep_box, (size_t)n_rep * 7);
            if (!c->d_rep) dalloc(&c->d_rep, (size_t)c->max_cells);
            c->max_reps = n_rep;
        }
        c->n_rep = n_rep;
    });
}

double order_key_value(unsigned long long k) { return order_val_host(k); }

int step2d_prepare(bfm_ctx* c, int n, double* d_sums)
{
    return guarded(c, [&] {
        require(c->list_n == n && c->list_csr, BFM_ERR_ARG,
                "no device neighbor list for these cells (bfm2d_dev_build_neighbors)");
        const Params2D P = params_2d(c);
        KLAUNCH(c, "pair_2d", k2d_pair, kPairGrid, 256, 0, c->d_m, n, c->pi, c->pj, c->pdx, c->pdy, c->ps, c->pt,
                c->d_adhere, P, c->d_rec, (int*)nullptr);
        KLAUNCH(c, "gather_2d", k2d_gather<true>, grid_for(n, 256), 256, 0, n, c->d_rec, row_lists(c), d_sums);
    });
}

int step2d_rows_update(bfm_ctx* c, const double* d_cells, int n, int ld, double* d_sums)
{
    return guarded(c, [&] {
        require(c->list_n == n && c->list_csr, BFM_ERR_ARG,
                "no device neighbor list for these cells (bfm2d_dev_build_neighbors)");
        const Params2D P = params_2d(c);
        KLAUNCH(c, "rows_2d", k2d_rows_update, kPairGrid, 256, 0, c->d_m, n, d_cells, ld, c->pi, c->pj, c->pdx,
                c->pdy, c->pdz, c->ps, c->pt, c->d_adhere, P, c->d_rec);
        KLAUNCH(c, "gather_2d", k2d_gather<true>, grid_for(n, 256), 256, 0, n, c->d_rec, row_lists(c), d_sums);
    });
}

int step2d_attempt(bfm_ctx* c, const Tableau& tb, double dt, const double* d_cells, int n, int ld,
                   const double* d_sums, double* d_new, int ld_new, double tol, unsigned long long* d_errmax,
                   int* d_count, double R, double Ldiv)
{
    return guarded(c, [&] {
        require(n >= 1 && n <= c->max_cells && ld >= n && ld_new >= n, BFM_ERR_ARG, "invalid n, ld or ld_new");
        const Params2D P = params_2d(c);
        Grow2D gr{};
        if (d_count) gr = Grow2D{1, R, Ldiv, dt, d_count};
        const int g = grid_for(n, 256);
        const int ncols = c->p2d.ncols;
        const double* nz = nullptr;
        switch (tb.s) {
        case 2:
            KLAUNCH(c, "attempt_2d", k2d_attempt<2>, g, 256, 0, n, ncols, d_cells, ld, d_sums, nz, P.strength, P.scd,
                    P.R, tb, dt, d_new, ld_new, (double*)nullptr, tol, d_errmax, gr, c->d_status);
            break;
        case 4:
            KLAUNCH(c, "attempt_2d", k2d_attempt<4>, g, 256, 0, n, ncols, d_cells, ld, d_sums, nz, P.strength, P.scd,
                    P.R, tb, dt, d_new, ld_new, (double*)nullptr, tol, d_errmax, gr, c->d_status);
            break;
        case 6:
            KLAUNCH(c, "attempt_2d", k2d_attempt<6>, g, 256, 0, n, ncols, d_cells, ld, d_sums, nz, P.strength, P.scd,
                    P.R, tb, dt, d_new, ld_new, (double*)nullptr, tol, d_errmax, gr, c->d_status);
            break;
        case 7:
            KLAUNCH(c, "attempt_2d", k2d_attempt<7>, g, 256, 0, n, ncols, d_cells, ld, d_sums, nz, P.strength, P.scd,
                    P.R, tb, dt, d_new, ld_new, (double*)nullptr, tol, d_errmax, gr, c->d_status);
            break;
        default:
            require(false, BFM_ERR_ARG, "the 2-D loop's tableaux have 2, 4, 6 or 7 stages");
        }
    });
}

int step2d_publish(bfm_ctx* c, unsigned long long* d_errmax, int* d_count, unsigned long long* host_mapped)
{
    return guarded(c, [&] {
        KLAUNCH(c, "publish_2d", k2d_publish, 1, 64, 0, d_errmax, d_count, c->d_status, host_mapped);
    });
}

}  // namespace bfm

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int bfm_create(bfm_ctx** out, int device, int max_cells, int64_t max_pairs)
{
    if (!out || max_cells < 1 || max_pairs < 1 || max_cells >= (1 << 27) || max_pairs >= (int64_t(1) << 31))
        return BFM_ERR_ARG;
    bfm_ctx* c = new bfm_ctx();
    c->device = device;
    c->max_cells = max_cells;
    if (max_cells > (1 << 26)) c->cand_variant = 6;   // the lane-per-home search packs positions in 26 bits
    c->max_pairs = max_pairs;
    c->max_bins = std::max(2 * max_cells, 4096);
    int rc = guarded(c, [&] {
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
        const size_t N = max_cells, P = max_pairs;
        dalloc(&c->d_status, 1);
        dalloc(&c->d_counters, (size_t)kCounterSlots * kCounterStride);
        dalloc(&c->d_cells, N * kMaxCols);
        dalloc(&c->d_cells_new, N * kMaxCols);
        dalloc(&c->d_nb, P * 7);
        dalloc(&c->d_noise, N * 6);
        dalloc(&c->d_out6, N * 6);
        dalloc(&c->d_err6, N * 6);
        dalloc(&c->d_assume2d, N);
        dalloc(&c->d_partials, kBBoxParts * 7);
        dalloc(&c->d_plan, 1);
        dalloc(&c->d_grid, 1);
        dalloc(&c->d_keys, N); dalloc(&c->d_keys_alt, N);
        dalloc(&c->d_vals, N); dalloc(&c->d_vals_alt, N);
        dalloc(&c->d_bin_count, (size_t)c->max_bins + 1);
        dalloc(&c->d_bin_off, (size_t)c->max_bins + 1);
        require((c->max_bins + kScanTile) / kScanTile <= kMaxScanTiles, BFM_ERR_CAPACITY,
                "max_cells too large for the bin-offset scan");
        dalloc(&c->d_tile_sum, (size_t)(c->max_bins + kScanTile) / kScanTile);
        dalloc(&c->d_scan_flag, (size_t)kMaxScanTiles);
        HIPCHK(hipMemset(c->d_scan_flag, 0, sizeof(unsigned) * kMaxScanTiles));
        dalloc(&c->d_sorted, N * 7);
        dalloc(&c->d_sortf, N * 2);
        dalloc(&c->d_sgeom, N * 8);
        dalloc(&c->d_geom, N * 8);
        dalloc(&c->d_sid, N);
        dalloc(&c->d_skey, N);
        dalloc(&c->d_count, N + 1);
        dalloc(&c->d_offsets, N + 1);
        dalloc(&c->d_slot_j, N * c->slot_cap);
        dalloc(&c->d_slot_rec, N * c->slot_cap * 8);
        dalloc(&c->pi, P); dalloc(&c->pj, P);
        dalloc(&c->pdx, P); dalloc(&c->pdy, P); dalloc(&c->pdz, P);
        dalloc(&c->ps, P); dalloc(&c->pt, P);
        dalloc(&c->d_m, 1);
        dalloc(&c->d_joff, N + 1);
        dalloc(&c->d_icnt, N + 1);
        dalloc(&c->d_jcnt, N + 1);
        dalloc(&c->d_perm_i, P);
        dalloc(&c->d_perm_j, P);
        dalloc(&c->d_rec, 8 * P);
        dalloc(&c->d_cnt64, N + 1);
        dalloc(&c->d_own64, N + 1);
        dalloc(&c->d_tot64, N + 1);
        dalloc(&c->d_off64, N + 1);
        dalloc(&c->d_jslot, N * c->slot_cap);
        dalloc(&c->d_jslotk, N * c->slot_cap);
        dalloc(&c->d_ordj, N * c->slot_cap);
        dalloc(&c->d_jrow, N * c->slot_cap);
        dalloc(&c->d_csr, 2 * P);
        dalloc(&c->d_cc, 6 * N);
        dalloc(&c->d_k, 7 * 6 * N);
        dalloc(&c->d_errmax, 1);
        HIPCHK(hipMemset(c->d_status, 0, sizeof(int)));
        HIPCHK(hipMemset(c->d_counters, 0, sizeof(unsigned long long) * kCounterSlots * kCounterStride));
        HIPCHK(hipMemset(c->d_m, 0, sizeof(int)));
        // temporary storage: max over the sort and the scans
        size_t t1 = 0, t2 = 0, t3 = 0;
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, c->d_keys, c->d_keys_alt, c->d_vals,
                                                  c->d_vals_alt, max_cells, 0, nbits(c->max_bins)));
        HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, c->d_bin_count, c->d_bin_off, c->max_bins + 1));
        HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, t3, c->d_count, c->d_offsets, max_cells + 1));
        size_t t4 = 0;
        HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, t4, c->d_tot64, c->d_off64, max_cells + 1));
        c->temp_bytes = std::max(std::max(t1, t4), std::max(t2, t3));
        HIPCHK(hipMalloc(&c->d_temp, c->temp_bytes));
        HIPCHK(hipDeviceSynchronize());
    });
    if (rc != BFM_OK) {
        *out = c;   // keep the context so the caller can read bfm_last_error
        return rc;
    }
    *out = c;
    return BFM_OK;
}

int bfm_destroy(bfm_ctx* c)
{
    if (!c) return BFM_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void* ptrs[] = {c->d_status, c->d_counters, c->d_cells, c->d_cells_new, c->d_nb, c->d_noise, c->d_out6,
                    c->d_err6, c->d_assume2d, c->d_partials, c->d_plan, c->d_keys, c->d_keys_alt, c->d_vals,
                    c->d_vals_alt, c->d_bin_count, c->d_bin_off, c->d_sorted, c->d_sortf, c->d_sgeom, c->d_sid, c->d_skey, c->d_count,
                    c->d_offsets, c->d_slot_j, c->d_slot_rec, c->pi, c->pj, c->pdx, c->pdy, c->pdz, c->ps, c->pt,
                    c->d_m, c->d_joff, c->d_icnt, c->d_jcnt, c->d_perm_i, c->d_perm_j, c->d_rec, c->d_cnt64, c->d_own64, c->d_tot64, c->d_off64, c->d_jslot, c->d_jslotk, c->d_ordj, c->d_jrow, c->d_csr, c->d_cc, c->d_rep, c->d_grid, c->d_rep_box,
                    c->d_k, c->d_errmax, c->d_tab_ov, c->d_tab_g, c->d_tab_r, c->d_temp, c->d_frec, c->d_zst, c->d_geom,
                    c->d_an_grid, c->d_an_curv, c->d_an_fr, c->d_nst, c->d_adhere, c->d_amap, c->d_tile_sum, c->d_scan_flag};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    bfm::surrogate_free(c);
    for (auto& t : c->pending) { c->event_pool.push_back(t.a); c->event_pool.push_back(t.b); }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BFM_OK;
}

const char* bfm_last_error(bfm_ctx* c) { return c ? c->err.c_str() : "null context"; }

int bfm_set_params(bfm_ctx* c, const bfm_params* p)
{
    return guarded(c, [&] {
        require(p != nullptr, BFM_ERR_ARG, "params is NULL");
        require(p->R > 0 && p->Rcell > 0 && p->Rcell <= p->R && p->E0 > 0, BFM_ERR_ARG,
                "invalid radii or modulus");
        require(p->friction_mode == 0 || p->friction_mode == 1, BFM_ERR_UNSUPPORTED,
                "friction_mode must be 0 (none) or 1 (kinetic)");
        require(p->adhesion_mode >= BFM_ADHESION_NONE && p->adhesion_mode <= BFM_ADHESION_JKR_ANISOTROPIC,
                BFM_ERR_ARG, "invalid adhesion_mode");
        require(p->jkr_mode >= 0 && p->jkr_mode <= 3, BFM_ERR_ARG, "invalid jkr_mode");
        require(p->ncols >= 21 && p->ncols <= kMaxCols, BFM_ERR_ARG, "ncols must be in 21..32");
        if (p->adhesion_mode != BFM_ADHESION_NONE)
            require(p->colidx_gamma >= 21 && p->colidx_gamma < p->ncols, BFM_ERR_ARG, "invalid colidx_gamma");
        if (p->friction_mode == 1) {
            require(p->colidx_eta_cell_cell >= 21 && p->colidx_eta_cell_cell < p->ncols &&
                        p->colidx_eta_cell_cell != p->colidx_gamma,
                    BFM_ERR_ARG, "invalid colidx_eta_cell_cell");
            if (!c->d_frec) {
                HIPCHK(hipStreamSynchronize(c->stream));
                dalloc(&c->d_frec, (size_t)c->max_pairs * 4);
                dalloc(&c->d_zst, (size_t)c->max_cells * 2);
            }
        }
        c->p = *p;
        c->params_set = true;
    });
}

int bfm_set_stream(bfm_ctx* c, void* stream)
{
    return guarded(c, [&] {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->own_stream) HIPCHK(hipStreamDestroy(c->stream));
        if (stream) {
            c->stream = (hipStream_t)stream;
            c->own_stream = false;
        } else {
            HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            c->own_stream = true;
        }
    });
}

int bfm_sync(bfm_ctx* c)
{
    return guarded(c, [&] { check_status(c); });
}

int bfm_set_jkr_table(bfm_ctx* c, const double* ov, int nd, const double* gm, int ng, const double* rad)
{
    return guarded(c, [&] {
        require(ov && rad && nd >= 2, BFM_ERR_ARG, "invalid JKR table");
        require(ng == 0 || (gm && ng >= 2), BFM_ERR_ARG, "invalid gamma mesh");
        for (void* q : {(void*)c->d_tab_ov, (void*)c->d_tab_g, (void*)c->d_tab_r})
            if (q) HIPCHK(hipFree(q));
        c->d_tab_ov = c->d_tab_g = c->d_tab_r = nullptr;
        dalloc(&c->d_tab_ov, nd);
        HIPCHK(hipMemcpy(c->d_tab_ov, ov, sizeof(double) * nd, hipMemcpyHostToDevice));
        if (ng > 0) {
            dalloc(&c->d_tab_g, ng);
            HIPCHK(hipMemcpy(c->d_tab_g, gm, sizeof(double) * ng, hipMemcpyHostToDevice));
        }
        size_t nr = (size_t)nd * std::max(ng, 1);
        dalloc(&c->d_tab_r, nr);
        HIPCHK(hipMemcpy(c->d_tab_r, rad, sizeof(double) * nr, hipMemcpyHostToDevice));
        c->tab_nd = nd;
        c->tab_ng = ng;
    });
}

int bfm_set_aniso_jkr_tables(bfm_ctx* c, const bfm_aniso_jkr_tables* t)
{
    return guarded(c, [&] {
        require(t != nullptr, BFM_ERR_ARG, "tables is NULL");
        const double* grids[kAnisoGrids] = {t->theta, t->half_l, t->coords, t->rx, t->phi, t->delta, t->gamma};
        const int ns[kAnisoGrids] = {t->n_theta, t->n_half_l, t->n_coords, t->n_rx, t->n_phi, t->n_delta,
                                     t->n_gamma};
        int64_t total = 0;
        for (int k = 0; k < kAnisoGrids; ++k) {
            const bool optional = k == kAnisoGrids - 1;
            require(ns[k] >= (optional ? 0 : 1) && (ns[k] == 0 || grids[k] != nullptr), BFM_ERR_ARG,
                    "invalid anisotropic JKR grid");
            for (int q = 0; q < ns[k]; ++q) {
                require(std::isfinite(grids[k][q]), BFM_ERR_ARG, "non-finite anisotropic JKR grid value");
                require(q == 0 || grids[k][q - 1] <= grids[k][q], BFM_ERR_ARG,
                        "anisotropic JKR grids must be non-decreasing");
            }
            total += ns[k];
        }
        require(total <= kAnisoMaxGrid, BFM_ERR_ARG, "anisotropic JKR grids exceed 8192 points in total");
        require(t->curvature_rx != nullptr && t->force_radius != nullptr, BFM_ERR_ARG,
                "anisotropic JKR tables are NULL");
        const int64_t n_curv = (int64_t)t->n_theta * t->n_half_l * t->n_coords;
        const int64_t n_fr =
            (int64_t)t->n_rx * t->n_rx * t->n_phi * t->n_delta * std::max(t->n_gamma, 1) * 2;
        HIPCHK(hipStreamSynchronize(c->stream));
        for (double** q : {&c->d_an_grid, &c->d_an_curv, &c->d_an_fr}) {
            if (*q) HIPCHK(hipFree(*q));
            *q = nullptr;
        }
        if (!c->d_nst) dalloc(&c->d_nst, (size_t)c->max_cells * 3);
        dalloc(&c->d_an_grid, (size_t)total);
        dalloc(&c->d_an_curv, (size_t)n_curv);
        dalloc(&c->d_an_fr, (size_t)n_fr);
        int64_t off = 0;
        for (int k = 0; k < kAnisoGrids; ++k) {
            if (ns[k] > 0)
                HIPCHK(hipMemcpy(c->d_an_grid + off, grids[k], sizeof(double) * ns[k], hipMemcpyHostToDevice));
            off += ns[k];
            c->an_n[k] = ns[k];
            // a first guess of the lookups' interval when the grid is evenly
            // spaced to within a quarter step (it only saves bisection steps)
            c->an_inv_h[k] = 0.0;
            if (ns[k] >= 3) {
                const double* v = grids[k];
                const double h = (v[ns[k] - 1] - v[0]) / (ns[k] - 1);
                bool even = h > 0;
                for (int q = 0; even && q < ns[k]; ++q) even = std::fabs(v[q] - (v[0] + q * h)) <= 0.25 * h;
                if (even) c->an_inv_h[k] = 1.0 / h;
            }
        }
        HIPCHK(hipMemcpy(c->d_an_curv, t->curvature_rx, sizeof(double) * n_curv, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_an_fr, t->force_radius, sizeof(double) * n_fr, hipMemcpyHostToDevice));
    });
}

int bfm_get_cell_neighbors(bfm_ctx* c, const double* cells, int n, int ld, double thr, double R, double Ldiv,
                           double* out, int64_t ld_out, int64_t* m_out)
{
    return guarded(c, [&] {
        need_params(c);
        check_cells(c, n, ld);
        require(m_out != nullptr, BFM_ERR_ARG, "m_out is NULL");
        upload_cells(c, cells, n, ld);
        build_neighbors_sync(c, c->d_cells, n, c->max_cells, thr, R, Ldiv);
        int m = 0;
        HIPCHK(hipMemcpy(&m, c->d_m, sizeof(int), hipMemcpyDeviceToHost));
        c->h_m = m;
        *m_out = m;
        require(m <= ld_out, BFM_ERR_CAPACITY, "neighbors_out is too small (see m_out)");
        if (m > 0) {
            KLAUNCH(c, "export_nb", k_export_nb, grid_for(m, 256, 4096), 256, 0, c->d_m, c->d_nb, (int64_t)m,
                    c->pi, c->pj, c->pdx, c->pdy, c->pdz, c->ps, c->pt);
            HIPCHK(hipMemcpy2DAsync(out, sizeof(double) * ld_out, c->d_nb, sizeof(double) * m,
                                    sizeof(double) * m, 7, hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int bfm_update_neighbor_distances(bfm_ctx* c, const double* cells, int n, int ld, double* nb, int64_t m,
                                  int64_t ld_nb)
{
    return guarded(c, [&] {
        need_params(c);
        check_cells(c, n, ld);
        upload_cells(c, cells, n, ld);
        import_host_neighbors(c, nb, m, ld_nb, n);
        if (m > 0) {
            update_distances(c, c->d_cells, c->max_cells, n);
            KLAUNCH(c, "export_nb", k_export_nb, grid_for(m, 256, 4096), 256, 0, c->d_m, c->d_nb, (int64_t)m,
                    c->pi, c->pj, c->pdx, c->pdy, c->pdz, c->ps, c->pt);
            HIPCHK(hipMemcpy2DAsync(nb + 2 * ld_nb, sizeof(double) * ld_nb, c->d_nb + 2 * m, sizeof(double) * m,
                                    sizeof(double) * m, 5, hipMemcpyDeviceToHost, c->stream));
        }
        check_status(c);
    });
}

int bfm_get_conservative_forces(bfm_ctx* c, const double* cells, int n, int ld, const double* nb, int64_t m,
                                int64_t ld_nb, const int* assume_2d, double* forces)
{
    return guarded(c, [&] {
        need_params(c);
        check_cells(c, n, ld);
        require(assume_2d && forces, BFM_ERR_ARG, "NULL argument");
        upload_cells(c, cells, n, ld);
        import_host_neighbors(c, nb, m, ld_nb, n);
        if (n > 0)
            HIPCHK(hipMemcpyAsync(c->d_assume2d, assume_2d, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
        compute_forces(c, c->d_cells, n, c->max_cells);
        launch_velocities(c, c->d_cells, n, c->max_cells, c->d_assume2d, nullptr, nullptr, c->d_cc);
        VelParams vp = vel_params(c);
        FricGather fg = fric_gather(c);
        RowLists L = row_lists(c);
        if (fg.frec) launch_friction(c, c->d_cells, c->max_cells, nullptr, 0);
        if (c->list_csr)
            KLAUNCH(c, "conservative", k_conservative<true>, grid_for(n, 256), 256, 0, n, c->d_cells, c->max_cells,
                    c->d_assume2d, vp, c->d_cc, L, fg, c->d_out6);
        else
            KLAUNCH(c, "conservative", k_conservative<false>, grid_for(n, 256), 256, 0, n, c->d_cells,
                    c->max_cells, c->d_assume2d, vp, c->d_cc, L, fg, c->d_out6);
        download6(c, c->d_out6, forces, n);
        check_status(c);
    });
}

int bfm_get_velocities(bfm_ctx* c, const double* cells, int n, int ld, const double* nb, int64_t m,
                       int64_t ld_nb, const int* assume_2d, const double* noise, double* vel)
{
    return guarded(c, [&] {
        need_params(c);
        check_cells(c, n, ld);
        require(assume_2d && vel, BFM_ERR_ARG, "NULL argument");
        upload_cells(c, cells, n, ld);
        import_host_neighbors(c, nb, m, ld_nb, n);
        if (n > 0) {
            HIPCHK(hipMemcpyAsync(c->d_assume2d, assume_2d, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
            if (noise)
                HIPCHK(hipMemcpyAsync(c->d_noise, noise, sizeof(double) * 6 * n, hipMemcpyHostToDevice, c->stream));
        }
        compute_forces(c, c->d_cells, n, c->max_cells);
        launch_velocities(c, c->d_cells, n, c->max_cells, c->d_assume2d, noise ? c->d_noise : nullptr,
                          c->d_out6, nullptr);
        download6(c, c->d_out6, vel, n);
        check_status(c);
    });
}

int bfm_step_rk(bfm_ctx* c, const double* A, const double* b, const double* bs, int s, double dt,
                const double* cells, int n, int ld, const double* nb, int64_t m, int64_t ld_nb,
                const double* noise, double* cells_new, double* errors)
{
    return guarded(c, [&] {
        need_params(c);
        check_cells(c, n, ld);
        require(A && b && bs && cells_new && errors, BFM_ERR_ARG, "NULL argument");
        Tableau tb = make_tableau(A, b, bs, s);
        upload_cells(c, cells, n, ld);
        import_host_neighbors(c, nb, m, ld_nb, n);
        if (noise && n > 0)
            HIPCHK(hipMemcpyAsync(c->d_noise, noise, sizeof(double) * 6 * n, hipMemcpyHostToDevice, c->stream));
        step_rk_device(c, tb, dt, c->d_cells, n, c->max_cells, noise ? c->d_noise : nullptr, c->d_cells_new,
                       c->max_cells, c->d_err6, 0.0, nullptr);
        if (n > 0) {
            HIPCHK(hipMemcpy2DAsync(cells_new, sizeof(double) * n, c->d_cells_new, sizeof(double) * c->max_cells,
                                    sizeof(double) * n, c->p.ncols, hipMemcpyDeviceToHost, c->stream));
            download6(c, c->d_err6, errors, n);
        }
        check_status(c);
    });
}

int bfm_step_verlet(bfm_ctx* c, double dt, double rho, const double* cells, int n, int ld, const double* nb,
                    int64_t m, int64_t ld_nb, const double* noise, double* cells_new)
{
    return guarded(c, [&] {
        need_params(c);
        check_cells(c, n, ld);
        require(cells_new != nullptr, BFM_ERR_ARG, "NULL argument");
        upload_cells(c, cells, n, ld);
        import_host_neighbors(c, nb, m, ld_nb, n);
        if (noise && n > 0)
            HIPCHK(hipMemcpyAsync(c->d_noise, noise, sizeof(double) * 6 * n, hipMemcpyHostToDevice, c->stream));
        step_verlet_device(c, dt, rho, c->d_cells, n, c->max_cells, noise ? c->d_noise : nullptr, c->d_cells_new,
                           c->max_cells);
        if (n > 0)
            HIPCHK(hipMemcpy2DAsync(cells_new, sizeof(double) * n, c->d_cells_new, sizeof(double) * c->max_cells,
                                    sizeof(double) * n, c->p.ncols, hipMemcpyDeviceToHost, c->stream));
        check_status(c);
    });
}

int bfm_dev_step_verlet(bfm_ctx* c, double dt, double rho, const double* d_cells, int n, int ld,
                        const double* d_noise, double* d_cells_new, int ld_new)
{
    return guarded(c, [&] {
        check_cells(c, n, ld);
        require(d_cells_new != nullptr && ld_new >= n, BFM_ERR_ARG, "need d_cells_new with ld_new >= n");
        step_verlet_device(c, dt, rho, d_cells, n, ld, d_noise, d_cells_new, ld_new);
    });
}

int bfm_dev_build_neighbors(bfm_ctx* c, const double* d_cells, int n, int ld, double thr, double R, double Ldiv,
                            int64_t* m_out)
{
    return guarded(c, [&] {
        need_params(c);
        check_cells(c, n, ld);
        if (m_out) {
            build_neighbors_sync(c, d_cells, n, ld, thr, R, Ldiv);
            int m = 0;
            HIPCHK(hipMemcpy(&m, c->d_m, sizeof(int), hipMemcpyDeviceToHost));
            c->h_m = m;
            *m_out = m;
        } else {
            build_neighbors(c, d_cells, n, ld, thr, R, Ldiv);
        }
    });
}

int bfm_dev_update_neighbor_distances(bfm_ctx* c, const double* d_cells, int n, int ld)
{
    return guarded(c, [&] {
        check_cells(c, n, ld);
        require(c->list_n == n, BFM_ERR_ARG, "neighbor list was built for a different number of cells");
        update_distances(c, d_cells, ld, n);
    });
}

int bfm_dev_export_neighbors(bfm_ctx* c, double* d_nb, int64_t ld_out)
{
    return guarded(c, [&] {
        int m = 0;
        HIPCHK(hipMemcpyAsync(&m, c->d_m, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        require(m <= ld_out, BFM_ERR_CAPACITY, "output too small for the neighbor list");
        if (m > 0)
            KLAUNCH(c, "export_nb", k_export_nb, grid_for(m, 256, 4096), 256, 0, c->d_m, d_nb, ld_out, c->pi,
                    c->pj, c->pdx, c->pdy, c->pdz, c->ps, c->pt);
    });
}

int bfm_dev_import_neighbors(bfm_ctx* c, const double* d_nb, int64_t m, int64_t ld_nb, int n)
{
    return guarded(c, [&] {
        require(m >= 0 && m <= c->max_pairs, BFM_ERR_CAPACITY, "m exceeds max_pairs");
        require(n >= 0 && n <= c->max_cells, BFM_ERR_CAPACITY, "n exceeds max_cells");
        if (m > 0)
            KLAUNCH(c, "import_nb", k_import_nb, grid_for(m, 256, 4096), 256, 0, d_nb, m, ld_nb, c->pi, c->pj,
                    c->pdx, c->pdy, c->pdz, c->ps, c->pt);
        int mi = (int)m;
        HIPCHK(hipMemcpyAsync(c->d_m, &mi, sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->h_m = m;
        c->list_n = n;
        c->list_sorted_i = false;   // unknown order: use the general transpose
        c->transposed = false;
        c->list_csr = false;
    });
}

int bfm_dev_get_velocities(bfm_ctx* c, const double* d_cells, int n, int ld, const int* d_a2d,
                           const double* d_noise, double* d_vel)
{
    return guarded(c, [&] {
        need_params(c);
        check_cells(c, n, ld);
        require(c->list_n == n, BFM_ERR_ARG, "neighbor list was built for a different number of cells");
        compute_forces(c, d_cells, n, ld);
        launch_velocities(c, d_cells, n, ld, d_a2d, d_noise, d_vel, nullptr);
    });
}

int bfm_dev_stage(bfm_ctx* c, const double* d_cells, int n, int ld, double thr, double R, double Ldiv,
                  double* d_vel)
{
    return guarded(c, [&] {
        need_params(c);
        check_cells(c, n, ld);
        const bool fuse = n > 1 && fused_forces_ok(c);
        if (fuse && c->cand_variant == 8 && c->p.friction_mode == 0) {
            // an isotropic interaction without friction: velocities straight
            // from the search's per-cell lists (no row compaction)
            build_neighbors(c, d_cells, n, ld, thr, R, Ldiv, false, false);
            const FusedForces ff{force_params(c), d_cells, ld, c->d_tab_ov, c->d_tab_g, c->d_tab_r, c->tab_nd,
                                 c->tab_ng, nullptr};
            KLAUNCH(c, "pair_forces", k_slot_order<kSlotLanes>, grid_for(kSlotLanes * (int64_t)n, 256), 256, 0, n, ff,
                    (int64_t)c->max_cells, c->slot_cap, c->d_cnt64, c->d_own64, c->d_slot_j, c->d_slot_rec,
                    c->d_jslot, c->d_jslotk, c->d_jrow, c->d_ordj, c->d_status);
            KLAUNCH(c, "velocities", k_velocities_slots, grid_for(n, 256), 256, 0, n, d_cells, ld, vel_params(c),
                    (int64_t)c->max_cells, c->slot_cap, c->d_cnt64, c->d_own64, c->d_slot_rec, c->d_jrow,
                    c->d_ordj, d_vel, c->d_status);
            return;
        }
        // an isotropic interaction's pair forces come out of the compaction
        build_neighbors(c, d_cells, n, ld, thr, R, Ldiv, fuse);
        if (!fuse) compute_forces(c, d_cells, n, ld);
        launch_velocities(c, d_cells, n, ld, nullptr, nullptr, d_vel, nullptr);
    });
}

int bfm_dev_step_rk(bfm_ctx* c, const double* A, const double* b, const double* bs, int s, double dt,
                    const double* d_cells, int n, int ld, const double* d_noise, double* d_cells_new,
                    double* d_errors, double tol, double* d_err_scaled)
{
    return guarded(c, [&] {
        check_cells(c, n, ld);
        Tableau tb = make_tableau(A, b, bs, s);
        step_rk_device(c, tb, dt, d_cells, n, ld, d_noise, d_cells_new, ld, d_errors, tol, d_err_scaled);
    });
}

int bfm_autotune_neighbors(bfm_ctx* c, const double* d_cells, int n, int ld, double thr, double R, double Ldiv,
                           double* best_scale, int* best_variant, double* best_ms)
{
    return guarded(c, [&] {
        need_params(c);
        check_cells(c, n, ld);
        static const double scales[] = {0.5, 0.75, 1.0, 1.25};
        static const int variants[] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
        double best = INFINITY, bs = c->bin_scale;
        int bv = c->cand_variant;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        const bool was_timing = c->timing;
        c->timing = false;
        for (double sc : scales)
            for (int v : variants) {
                c->bin_scale = sc;
                c->cand_variant = v;
                build_neighbors_sync(c, d_cells, n, ld, thr, R, Ldiv);   // warm-up, grows slots if needed
                HIPCHK(hipEventRecord(e0, c->stream));
                for (int k = 0; k < 3; ++k) build_neighbors(c, d_cells, n, ld, thr, R, Ldiv);
                HIPCHK(hipEventRecord(e1, c->stream));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                raise_status(c, take_status(c));
                if (ms / 3 < best) { best = ms / 3; bs = sc; bv = v; }
            }
        c->timing = was_timing;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c->bin_scale = bs;
        c->cand_variant = bv;
        build_neighbors_sync(c, d_cells, n, ld, thr, R, Ldiv);
        if (best_scale) *best_scale = bs;
        if (best_variant) *best_variant = bv;
        if (best_ms) *best_ms = best;
    });
}

int bfm_enable_timing(bfm_ctx* c, int enable)
{
    return guarded(c, [&] { c->timing = enable != 0; });
}

int bfm_reset_timing(bfm_ctx* c)
{
    return guarded(c, [&] {
        resolve_timing(c);
        c->ktimes.clear();
        c->korder.clear();
    });
}

int bfm_kernel_times(bfm_ctx* c, char* names, int names_cap, double* ms, int64_t* counts, int cap, int* nk)
{
    return guarded(c, [&] {
        resolve_timing(c);
        std::string all;
        int k = 0;
        for (auto& name : c->korder) {
            if (k >= cap) break;
            ms[k] = c->ktimes[name].first;
            counts[k] = c->ktimes[name].second;
            all += name;
            all += '\n';
            ++k;
        }
        *nk = k;
        if (names && names_cap > 0) {
            std::strncpy(names, all.c_str(), names_cap - 1);
            names[names_cap - 1] = 0;
        }
    });
}

int bfm_stats(bfm_ctx* c, int64_t* m, int64_t* screened, int64_t* tests, double* edge, int* nbins)
{
    return guarded(c, [&] {
        HIPCHK(hipStreamSynchronize(c->stream));
        int mm = 0;
        HIPCHK(hipMemcpy(&mm, c->d_m, sizeof(int), hipMemcpyDeviceToHost));
        std::vector<unsigned long long> slots((size_t)kCounterSlots * kCounterStride);
        HIPCHK(hipMemcpy(slots.data(), c->d_counters, sizeof(unsigned long long) * slots.size(),
                         hipMemcpyDeviceToHost));
        unsigned long long ctr[2] = {0, 0};
        for (int k = 0; k < kCounterSlots; ++k) {
            ctr[0] += slots[(size_t)k * kCounterStride];
            ctr[1] += slots[(size_t)k * kCounterStride + 8];
        }
        Plan P;
        HIPCHK(hipMemcpy(&P, c->d_plan, sizeof(Plan), hipMemcpyDeviceToHost));
        if (m) *m = mm;
        if (screened) *screened = (int64_t)ctr[0];
        if (tests) *tests = (int64_t)ctr[1];
        if (edge) *edge = P.edge;
        if (nbins) *nbins = P.nbins;
        HIPCHK(hipMemset(c->d_counters, 0, sizeof(unsigned long long) * slots.size()));
    });
}

int bfm_set_tuning(bfm_ctx* c, double bin_scale, int block_threads, int kernel_variant)
{
    return guarded(c, [&] {
        if (kernel_variant >= 0) {
            require(kernel_variant <= 8, BFM_ERR_ARG, "kernel_variant must be 0..8");
            require(kernel_variant != 7 || c->max_cells < (1 << 25), BFM_ERR_ARG,
                    "kernel_variant 7 needs max_cells < 2^25");
            require(kernel_variant != 8 || c->max_cells <= (1 << 26), BFM_ERR_ARG,
                    "kernel_variant 8 needs max_cells <= 2^26");
            c->cand_variant = kernel_variant;
        }
        if (bin_scale > 0) {
            require(bin_scale >= 1.0 / 3.0 && bin_scale <= 4.0, BFM_ERR_ARG, "bin_scale out of range");
            c->bin_scale = bin_scale;
        }
        if (block_threads > 0) {
            require(block_threads == 64 || block_threads == 128 || block_threads == 256 || block_threads == 512,
                    BFM_ERR_ARG, "block_threads must be 64, 128, 256 or 512");
            c->block_threads = block_threads;
        }
    });
}

// ---- the 2-D model ----------------------------------------------------------

int bfm2d_set_params(bfm_ctx* c, const bfm2d_params* p)
{
    return guarded(c, [&] {
        require(p != nullptr, BFM_ERR_ARG, "params is NULL");
        require(p->R > 0 && p->Rcell > 0 && p->Rcell <= p->R, BFM_ERR_ARG, "invalid radii");
        require(p->ncols >= 17 && p->ncols <= kMaxCols, BFM_ERR_ARG, "ncols must be in 17..32");
        require(p->adhesion_mode >= BFM2D_ADHESION_NONE && p->adhesion_mode <= BFM2D_ADHESION_JKR_ISOTROPIC,
                BFM_ERR_ARG, "invalid adhesion_mode");
        require(p->adhesion_mode != 3, BFM_ERR_UNSUPPORTED, "GBK adhesion is not supported");
        require(p->adhesion_mode != BFM2D_ADHESION_JKR_ISOTROPIC || (p->E0 > 0 && p->gamma >= 0), BFM_ERR_ARG,
                "isotropic JKR needs E0 > 0 and gamma >= 0");
        if (p->adhesion_mode != BFM2D_ADHESION_NONE && !c->d_adhere) {
            HIPCHK(hipStreamSynchronize(c->stream));
            dalloc(&c->d_adhere, (size_t)c->max_pairs);
        }
        c->p2d = *p;
        c->p2d_set = true;
    });
}

int bfm2d_get_cell_neighbors(bfm_ctx* c, const double* cells, int n, int ld, double thr, double R, double Ldiv,
                             double* out, int64_t ld_out, int64_t* m_out)
{
    return guarded(c, [&] {
        require(c->p2d_set, BFM_ERR_ARG, "bfm2d_set_params has not been called");
        require(n >= 0 && n <= c->max_cells, BFM_ERR_CAPACITY, "n exceeds max_cells of the context");
        require(ld >= n, BFM_ERR_ARG, "ld < n");
        require(m_out != nullptr, BFM_ERR_ARG, "m_out is NULL");
        if (n > 0)
            HIPCHK(hipMemcpy2DAsync(c->d_cells, sizeof(double) * c->max_cells, cells, sizeof(double) * ld,
                                    sizeof(double) * n, c->p2d.ncols, hipMemcpyHostToDevice, c->stream));
        KLAUNCH(c, "embed_2d", k2d_embed, grid_for(n, 256), 256, 0, n, c->d_cells, c->max_cells, c->d_cells_new,
                c->max_cells);
        build_neighbors_sync(c, c->d_cells_new, n, c->max_cells, thr, R, Ldiv);
        int m = 0;
        HIPCHK(hipMemcpy(&m, c->d_m, sizeof(int), hipMemcpyDeviceToHost));
        c->h_m = m;
        *m_out = m;
        require(m <= ld_out, BFM_ERR_CAPACITY, "neighbors_out is too small (see m_out)");
        if (m > 0) {
            KLAUNCH(c, "export_nb", k_export_nb, grid_for(m, 256, 4096), 256, 0, c->d_m, c->d_nb, (int64_t)m, c->pi,
                    c->pj, c->pdx, c->pdy, c->pdz, c->ps, c->pt);
            std::vector<double> nb7((size_t)7 * m);
            HIPCHK(hipMemcpyAsync(nb7.data(), c->d_nb, sizeof(double) * 7 * m, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            const int src[6] = {0, 1, 2, 3, 5, 6};
            for (int q = 0; q < 6; ++q)
                std::memcpy(out + (size_t)q * ld_out, nb7.data() + (size_t)src[q] * m, sizeof(double) * m);
        }
        check_status(c);
    });
}

int bfm2d_get_velocities(bfm_ctx* c, const double* cells, int n, int ld, const double* nb, int64_t m,
                         int64_t ld_nb, const int* to_adhere, const double* noise, double* vel)
{
    return guarded(c, [&] {
        require(c->p2d_set, BFM_ERR_ARG, "bfm2d_set_params has not been called");
        require(vel != nullptr, BFM_ERR_ARG, "velocities_out is NULL");
        prepare_2d(c, cells, n, ld, nb, m, ld_nb, to_adhere, noise);
        Tableau tb;
        std::memset(&tb, 0, sizeof(tb));
        tb.s = 1;
        stages_2d(c, tb, 0.0, c->d_cells, c->max_cells, n, noise ? c->d_noise : nullptr);
        if (n > 0)
            HIPCHK(hipMemcpyAsync(vel, c->d_k, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, c->stream));
        check_status(c);
    });
}

int bfm2d_step_rk(bfm_ctx* c, const double* A, const double* b, const double* bs, int s, double dt,
                  const double* cells, int n, int ld, const double* nb, int64_t m, int64_t ld_nb,
                  const int* to_adhere, const double* noise, double* cells_out, int ld_out, double* errors)
{
    return guarded(c, [&] {
        require(c->p2d_set, BFM_ERR_ARG, "bfm2d_set_params has not been called");
        require(A && b && bs && cells_out && errors, BFM_ERR_ARG, "NULL argument");
        require(ld_out >= n, BFM_ERR_ARG, "ld_out < n");
        const Tableau tb = make_tableau(A, b, bs, s);
        prepare_2d(c, cells, n, ld, nb, m, ld_nb, to_adhere, noise);
        stages_2d(c, tb, dt, c->d_cells, c->max_cells, n, noise ? c->d_noise : nullptr);
        KLAUNCH(c, "final_2d", k2d_final, grid_for(n, 256), 256, 0, n, c->p2d.ncols, c->d_cells, c->max_cells, tb,
                dt, c->d_k, c->d_cells_new, c->max_cells, c->d_err6, 0.0, (unsigned long long*)nullptr, Grow2D{});
        if (n > 0) {
            HIPCHK(hipMemcpy2DAsync(cells_out, sizeof(double) * ld_out, c->d_cells_new,
                                    sizeof(double) * c->max_cells, sizeof(double) * n, c->p2d.ncols,
                                    hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(errors, c->d_err6, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, c->stream));
        }
        check_status(c);
    });
}

int bfm2d_dev_build_neighbors(bfm_ctx* c, const double* d_cells, int n, int ld, double thr, double R, double Ldiv,
                              const int* adhesion_map, int n_groups, int64_t* m_out)
{
    return guarded(c, [&] {
        require(c->p2d_set, BFM_ERR_ARG, "bfm2d_set_params has not been called");
        require(n >= 0 && n <= c->max_cells, BFM_ERR_CAPACITY, "n exceeds max_cells of the context");
        require(ld >= n, BFM_ERR_ARG, "ld < n");
        require(n_groups >= 0 && n_groups <= 64, BFM_ERR_ARG, "n_groups must be 0..64");
        KLAUNCH(c, "embed_2d", k2d_embed, grid_for(n, 256), 256, 0, n, d_cells, ld, c->d_cells_new, c->max_cells);
        if (m_out) build_neighbors_sync(c, c->d_cells_new, n, c->max_cells, thr, R, Ldiv);
        else build_neighbors(c, c->d_cells_new, n, c->max_cells, thr, R, Ldiv);
        if (!c->d_adhere) {
            HIPCHK(hipStreamSynchronize(c->stream));
            dalloc(&c->d_adhere, (size_t)c->max_pairs);
        }
        const int* amap = nullptr;
        if (adhesion_map && n_groups > 0) {
            if (!c->d_amap) dalloc(&c->d_amap, (size_t)64 * 64);
            HIPCHK(hipMemcpyAsync(c->d_amap, adhesion_map, sizeof(int) * n_groups * n_groups,
                                  hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            amap = c->d_amap;
        }
        KLAUNCH(c, "adhere_2d", k2d_adhere, kPairGrid, 256, 0, c->d_m, d_cells, ld, c->pi, c->pj, c->pdx, c->pdy,
                amap, n_groups, c->p2d.R, c->p2d.Rcell, c->d_adhere);
        if (m_out) {
            int m = 0;
            HIPCHK(hipMemcpy(&m, c->d_m, sizeof(int), hipMemcpyDeviceToHost));
            c->h_m = m;
            *m_out = m;
        }
    });
}

int bfm2d_dev_update_neighbor_distances(bfm_ctx* c, const double* d_cells, int n, int ld)
{
    return guarded(c, [&] {
        require(n >= 0 && n <= c->max_cells && ld >= n, BFM_ERR_ARG, "invalid n or ld");
        require(c->list_n == n, BFM_ERR_ARG, "neighbor list was built for a different number of cells");
        KLAUNCH(c, "embed_2d", k2d_embed, grid_for(n, 256), 256, 0, n, d_cells, ld, c->d_cells_new, c->max_cells);
        update_distances(c, c->d_cells_new, c->max_cells, n);
    });
}

int bfm2d_dev_step_rk(bfm_ctx* c, const double* A, const double* b, const double* bs, int s, double dt,
                      const double* d_cells, int n, int ld, const double* d_noise, double* d_cells_out, int ld_out,
                      double* d_errors)
{
    return guarded(c, [&] {
        require(c->p2d_set, BFM_ERR_ARG, "bfm2d_set_params has not been called");
        require(A && b && bs && d_cells_out, BFM_ERR_ARG, "NULL argument");
        require(n >= 0 && n <= c->max_cells && ld >= n && ld_out >= n, BFM_ERR_ARG, "invalid n, ld or ld_out");
        require(c->list_n == n && c->list_csr, BFM_ERR_ARG,
                "no device neighbor list for these cells (bfm2d_dev_build_neighbors)");
        const Tableau tb = make_tableau(A, b, bs, s);
        const Params2D P = params_2d(c);
        KLAUNCH(c, "pair_2d", k2d_pair, kPairGrid, 256, 0, c->d_m, n, c->pi, c->pj, c->pdx, c->pdy, c->ps, c->pt,
                c->d_adhere, P, c->d_rec, (int*)nullptr);
        stages_2d(c, tb, dt, d_cells, ld, n, d_noise);
        KLAUNCH(c, "final_2d", k2d_final, grid_for(n, 256), 256, 0, n, c->p2d.ncols, d_cells, ld, tb, dt, c->d_k,
                d_cells_out, ld_out, d_errors ? d_errors : c->d_err6, 0.0, (unsigned long long*)nullptr,
                Grow2D{});
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// getVelocities without allocation, for callers that cannot allocate device
// memory (code under stream capture, serving loops). Every buffer beyond the
// caller's inputs and outputs lives in `scratch`, a fixed-size device array
// of this library; each GPU holds its own copy until the process ends.
// ---------------------------------------------------------------------------
namespace {

constexpr int kNaMaxCells = 65536;
constexpr int64_t kNaMaxRows = 262144;
constexpr int kNaMaxDevices = 64;
// int region after the pair records: pi, pj, perm_i, perm_j (kNaMaxRows each),
// icnt, jcnt, ioff, joff (kNaMaxCells + 1 each), m, status.
constexpr int64_t kNaInts = 4 * kNaMaxRows + 4 * ((int64_t)kNaMaxCells + 1) + 2;
constexpr int64_t kNaDoubles = 8 * kNaMaxRows + (kNaInts + 1) / 2;   // ~21 MB
constexpr int kNaBadIndex = 1 << 16;     // status bit: a row index out of range

static __device__ __align__(256) double scratch[kNaDoubles];

std::mutex g_na_mu;
double* g_na_ptr[kNaMaxDevices] = {};

// The device address of `scratch` on device dev (current), looked up once per
// device; nullptr on failure.
double* na_scratch(int dev)
{
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_na_ptr[dev] = static_cast<double*>(p);
    }
    return g_na_ptr[dev];
}

__global__ void k_na_init(int* __restrict__ d_m, int m, int* __restrict__ status)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *d_m = m;
        *status = 0;
    }
}

// Row indices of an m x 7 neighbor array as int32; a row whose i or j is not
// an integer in [0, n) is mapped to cell n (counted in the spare last entry of
// the per-cell counts, gathered by no cell) and flagged.
__global__ void k_na_ij(const double* __restrict__ nb, int64_t m, int64_t ld, int n, int* __restrict__ pi,
                        int* __restrict__ pj, int* __restrict__ status)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
        const double vi = nb[k], vj = nb[ld + k];
        const bool ok = vi >= 0 && vi < n && vj >= 0 && vj < n && vi == (double)(int)vi && vj == (double)(int)vj;
        pi[k] = ok ? (int)vi : n;
        pj[k] = ok ? (int)vj : n;
        if (!ok) atomicOr(status, kNaBadIndex);
    }
}

// Exclusive prefix sum of len <= kNaMaxCells + 1 ints in one work-group:
// each thread sums a contiguous chunk, the 1024 chunk sums are scanned in LDS.
__global__ void __launch_bounds__(1024) k_na_scan(const int* __restrict__ in, int* __restrict__ out, int len)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int chunk = (len + 1023) / 1024;
    const int a = min(t * chunk, len), b = min(a + chunk, len);
    int s = 0;
    for (int q = a; q < b; ++q) s += in[q];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int q = a; q < b; ++q) {
        const int v = in[q];
        out[q] = run;
        run += v;
    }
}

int na_status_code(int st)
{
    if (st & kNaBadIndex) return BFM_ERR_ARG;
    if (st & ST_JKR_ROOT) return BFM_ERR_JKR_ROOT;
    if (st & ST_NAN) return BFM_ERR_NAN;
    return st ? BFM_ERR_HIP : BFM_OK;
}

}  // namespace

extern "C" int bfm_dev_get_velocities_noalloc(const bfm_params* p, const double* d_cells, int n, int ld,
                                              const double* d_neighbors, int64_t m, int64_t ld_nb,
                                              const int* d_assume_2d, const double* d_noise, double* d_vel,
                                              void* stream)
{
    if (!p) return BFM_ERR_ARG;
    if (n < 0 || m < 0 || ld < n || (m > 0 && ld_nb < m)) return BFM_ERR_ARG;
    if (n > kNaMaxCells || m > kNaMaxRows) return BFM_ERR_CAPACITY;
    if (p->adhesion_mode != BFM_ADHESION_NONE &&
        !(p->adhesion_mode == BFM_ADHESION_JKR_ISOTROPIC && p->jkr_mode == BFM_JKR_EXACT))
        return BFM_ERR_UNSUPPORTED;
    if (p->friction_mode != 0) return BFM_ERR_UNSUPPORTED;
    if (p->adhesion_mode != BFM_ADHESION_NONE && (p->colidx_gamma < 0 || p->colidx_gamma >= p->ncols))
        return BFM_ERR_ARG;
    if (n == 0) return BFM_OK;
    if (!d_cells || !d_vel || (m > 0 && !d_neighbors)) return BFM_ERR_ARG;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaMaxDevices) return BFM_ERR_HIP;
    double* base = na_scratch(dev);
    if (!base) return BFM_ERR_HIP;
    hipStream_t s = static_cast<hipStream_t>(stream);

    double* rec = base;
    int* ip = reinterpret_cast<int*>(base + 8 * kNaMaxRows);
    int* pi = ip;
    int* pj = pi + kNaMaxRows;
    int* perm_i = pj + kNaMaxRows;
    int* perm_j = perm_i + kNaMaxRows;
    int* icnt = perm_j + kNaMaxRows;
    int* jcnt = icnt + (kNaMaxCells + 1);
    int* ioff = jcnt + (kNaMaxCells + 1);
    int* joff = ioff + (kNaMaxCells + 1);
    int* d_m = joff + (kNaMaxCells + 1);
    int* status = d_m + 1;

    const int64_t nc = (int64_t)n + 1;
    const int gc = grid_for(nc, 256, 4096);
    hipLaunchKernelGGL(k_na_init, dim3(1), dim3(64), 0, s, d_m, (int)m, status);
    hipLaunchKernelGGL(k_fill_int, dim3(gc), dim3(256), 0, s, jcnt, 0, nc);
    hipLaunchKernelGGL(k_fill_int, dim3(gc), dim3(256), 0, s, icnt, 0, nc);
    if (m > 0)
        hipLaunchKernelGGL(k_na_ij, dim3(grid_for(m, 256, 4096)), dim3(256), 0, s, d_neighbors, m, ld_nb, n, pi,
                           pj, status);
    // pair records and the j-role counts (the imported-list path of
    // bfm_dev_get_velocities: the same kernels in the same order)
    const double* pdx = m > 0 ? d_neighbors + 2 * ld_nb : d_cells;   // not read when m == 0
    hipLaunchKernelGGL(k_pair_forces<false>, dim3(kPairGrid), dim3(256), 0, s, (const int*)d_m, d_cells, ld,
                       (const int*)pi, (const int*)pj, pdx, pdx + ld_nb, pdx + 2 * ld_nb, force_params_of(*p),
                       (const double*)nullptr, (const double*)nullptr, (const double*)nullptr, 0, 0, AnisoJKR{},
                       (const double*)nullptr, 0, pdx + 3 * ld_nb, pdx + 4 * ld_nb, (const double*)nullptr, rec,
                       jcnt, status, DistRefresh{});
    hipLaunchKernelGGL(k_na_scan, dim3(1), dim3(1024), 0, s, (const int*)jcnt, joff, (int)nc);
    hipLaunchKernelGGL(k_fill_int, dim3(gc), dim3(256), 0, s, jcnt, 0, nc);
    hipLaunchKernelGGL(k_fill_perm, dim3(kPairGrid), dim3(256), 0, s, (const int*)d_m, (const int*)pj,
                       (const int*)joff, jcnt, perm_j);
    hipLaunchKernelGGL(k_sort_segments, dim3(grid_for(n, 256)), dim3(256), 0, s, n, (const int*)joff, perm_j);
    hipLaunchKernelGGL(k_count_key, dim3(kPairGrid), dim3(256), 0, s, (const int*)d_m, (const int*)pi, icnt);
    hipLaunchKernelGGL(k_na_scan, dim3(1), dim3(1024), 0, s, (const int*)icnt, ioff, (int)nc);
    hipLaunchKernelGGL(k_fill_int, dim3(gc), dim3(256), 0, s, icnt, 0, nc);
    hipLaunchKernelGGL(k_fill_perm, dim3(kPairGrid), dim3(256), 0, s, (const int*)d_m, (const int*)pi,
                       (const int*)ioff, icnt, perm_i);
    hipLaunchKernelGGL(k_sort_segments, dim3(grid_for(n, 256)), dim3(256), 0, s, n, (const int*)ioff, perm_i);
    RowLists L;
    L.off64 = nullptr;
    L.csr = nullptr;
    L.cap = kNaMaxRows;
    L.ioff = ioff;
    L.perm_i = perm_i;
    L.joff = joff;
    L.perm_j = perm_j;
    FricGather fg;
    fg.frec = nullptr;
    fg.ps = nullptr;
    fg.pt = nullptr;
    hipLaunchKernelGGL(k_velocities<false>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, d_cells, ld, d_assume_2d,
                       d_noise, vel_params_of(*p), L, (const double*)rec, fg, d_vel, (double*)nullptr, status);
    if (hipGetLastError() != hipSuccess) return BFM_ERR_HIP;
    // Outside a stream capture the call is synchronous and reports the
    // kernels' status (bad index, JKR root, non-finite velocity).
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) return BFM_ERR_HIP;
    if (cs != hipStreamCaptureStatusNone) return BFM_OK;
    int st = 0;
    if (hipMemcpyAsync(&st, status, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return BFM_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return BFM_ERR_HIP;
    return na_status_code(st);
}

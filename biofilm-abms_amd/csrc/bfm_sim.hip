// bfm_sim.hip -- the growth / division loop of runSimulation
// (ABM-3d/ABM-3d-cpp/include/simulation.hpp:1106-2081) over the device
// mechanical step.
//
// State lives on the device in the reference's column-major cell layout with
// a fixed leading dimension (the capacity), so rows appended by division do
// not move the existing ones. Per iteration the host reads back a handful of
// scalars: the number of cells past the division length and the controller's
// scaled error. Division itself follows divideCells (growth.hpp:456-720) on
// the host, with the reference's random draws in the reference's order; the
// minimum-distance checks of a division are split into the part against the
// unchanged population (a device kernel, exact-safe screening) and the part
// among the rewritten rows (host, the same segment arithmetic).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "bfm_ctrl.hpp"
#include "bfm_ctx.hpp"
#include "bfm_device.hpp"
#include "bfm_internal.hpp"

using namespace bfm;

namespace {

struct SimError {
    int code;
    std::string msg;
};

[[noreturn]] void sim_fail(int code, const std::string& msg) { throw SimError{code, msg}; }

void hchk(hipError_t e, const char* what)
{
    if (e != hipSuccess) sim_fail(BFM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define SCHK(x) hchk((x), #x)

void abi(bfm_ctx* c, int rc)
{
    if (rc != BFM_OK) sim_fail(rc, bfm_last_error(c));
}

// ---------------------------------------------------------------------------
// random stream: MT19937 (Matsumoto & Nishimura 1998; boost::random::mt19937
// is this generator), uniform_01 on a 32-bit engine = x * 2^-32, and the
// Box-Muller draw of standardNormal (utils.hpp:648-655).
// ---------------------------------------------------------------------------
struct Mt19937 {
    uint32_t mt[624];
    int idx;
    explicit Mt19937(uint32_t seed)
    {
        mt[0] = seed;
        for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
        idx = 624;
    }
    void twist()
    {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            mt[i] = v;
        }
        idx = 0;
    }
    uint32_t next()
    {
        if (idx >= 624) twist();
        uint32_t y = mt[idx++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
};

constexpr double kTwoPi = 6.283185307179586476925286766559;

double uniform01(Mt19937& g) { return (double)g.next() * (1.0 / 4294967296.0); }

double standard_normal(Mt19937& g)
{
    double u = uniform01(g);
    double v = uniform01(g);
    double c = sqrt(-2 * log(u));
    return c * cos(kTwoPi * v);
}

// The distributions runSimulation builds (simulation.hpp:1294-1339).
struct Draws {
    const bfm_sim_config* cfg;
    double growth(Mt19937& g, int group) const
    {
        require_group(group);
        return cfg->growth_mean[group - 1] + cfg->growth_std[group - 1] * standard_normal(g);
    }
    double daughter_length(Mt19937& g) const
    {
        double r = 0.5 + cfg->daughter_length_std * standard_normal(g);
        if (r < 0) return 0.0;
        else if (r > 1) return 1.0;
        else return r;
    }
    double angle_xy(Mt19937& g) const
    {
        double r = uniform01(g);
        return -cfg->daughter_angle_xy_bound + 2 * cfg->daughter_angle_xy_bound * r;
    }
    double angle_z(Mt19937& g) const
    {
        double r = uniform01(g);
        return -cfg->daughter_angle_z_bound + 2 * cfg->daughter_angle_z_bound * r;
    }
    static void require_group(int group)
    {
        if (group < 1 || group > 2) sim_fail(BFM_ERR_ARG, "cell group must be 1 or 2");
    }
};

// rotateXY and rotateOutOfXY (utils.hpp:791-829).
void rotate_xy(const double n[3], double theta, double out[3])
{
    double s = sin(theta), c = cos(theta);
    out[0] = (c * n[0] + (-s) * n[1]) + 0.0 * n[2];
    out[1] = (s * n[0] + c * n[1]) + 0.0 * n[2];
    out[2] = (0.0 * n[0] + 0.0 * n[1]) + 1.0 * n[2];
}

void rotate_out_of_xy(const double n[3], double theta, double out[3])
{
    double v0 = n[1] * 1.0 - n[2] * 0.0;       // v = n x e_z
    double v1 = n[2] * 0.0 - n[0] * 1.0;
    double v2 = n[0] * 0.0 - n[1] * 0.0;
    double nrm = sqrt((v0 * v0 + v1 * v1) + v2 * v2);
    v0 /= nrm; v1 /= nrm; v2 /= nrm;
    double s = sin(theta), c = cos(theta);
    double w0 = v1 * n[2] - v2 * n[1];         // w = v x n
    double w1 = v2 * n[0] - v0 * n[2];
    double w2 = v0 * n[1] - v1 * n[0];
    out[0] = n[0] * c + w0 * s;
    out[1] = n[1] * c + w1 * s;
    out[2] = n[2] * c + w2 * s;
}

// Query segment for the minimum-distance checks: distances to the cells in
// [lo, hi) (one replicate's rows), skipping `self` and, when skip_dividing,
// the cells being divided (their rows are rewritten).
struct Query {
    double r[3], n[3], hl;
    int self;       // index in the population to skip (-1: none)
    int lo, hi;
    int skip_dividing;
};

Seg seg_of_row(const double* row)
{
    Seg s;
    s.r[0] = row[C_RX]; s.r[1] = row[C_RY]; s.r[2] = row[C_RZ];
    s.n[0] = row[C_NX]; s.n[1] = row[C_NY]; s.n[2] = row[C_NZ];
    s.hl = row[C_HALF_L];
    return s;
}

Query query_of_row(const double* row, int self, int lo, int hi, int skip_dividing)
{
    Query q;
    Seg s = seg_of_row(row);
    std::memcpy(q.r, s.r, sizeof(s.r));
    std::memcpy(q.n, s.n, sizeof(s.n));
    q.hl = s.hl;
    q.self = self;
    q.lo = lo;
    q.hi = hi;
    q.skip_dividing = skip_dividing;
    return q;
}

// |d| of distBetweenCells(query, other), as minDistToCell (growth.hpp:128-162).
__host__ __device__ inline double seg_dist(const Seg& a, const Seg& b)
{
    double d[3], s, t;
    dist_between_cells(a, b, d, s, t);
    return norm3(d[0], d[1], d[2]);
}

// A pair whose centers are farther apart than cut + both half lengths (with a
// relative and absolute margin far above the rounding of |d|) has |d| >= cut,
// so it cannot change a comparison of the minimum with cut.
__host__ __device__ inline bool beyond_cut(const Seg& a, const Seg& b, double cut)
{
    double dx = b.r[0] - a.r[0], dy = b.r[1] - a.r[1], dz = b.r[2] - a.r[2];
    double bound = (cut + a.hl + b.hl) * (1.0 + 1e-9) + 1e-9;
    return (dx * dx + dy * dy) + dz * dz > bound * bound;
}

// Minimum distances from queries to the unchanged part of the population.
// Only comparisons of the results with `cut` are meaningful (screened pairs
// are left out).
using BaseMinFn = std::function<void(const std::vector<Query>&, double cut, std::vector<double>&)>;

// One replicate's divideCells call (growth.hpp:456-720) as a state machine, so
// that the placement rounds of many replicates share one device query.
struct DivisionJob {
    // inputs
    Mt19937* rng = nullptr;
    std::vector<int>* parents = nullptr;
    double t = 0;
    int lo = 0, hi = 0;                 // the replicate's rows [lo, hi) before division
    std::vector<int> idx;               // rows of the dividing cells, ascending
    std::vector<double> mrows;          // their original values, nd x ncols row-major
    // state
    int nd = 0, ntries = 0;
    std::vector<int> check, satisfies;
    std::vector<double> mo, da;         // proposal: rewritten mothers and appended daughters
    std::vector<int> qcell;             // this round's queries: position among the 2 nd new rows
    size_t q0 = 0;                      // first query of this job in the round
    bool active() const
    {
        int sat = 0;
        for (int v : satisfies) sat += v;
        return ntries == 0 || (ntries < 10 && sat < nd);
    }
};

// One placement attempt of divideCells: re-orientation angles (four draws per
// cell with room, none on the last try), length ratios, the mothers' then the
// daughters' growth rates, in the reference's order.
void propose(const bfm_sim_config* cfg, int ncols, DivisionJob& jb)
{
    const Draws dr{cfg};
    const int nd = jb.nd;
    const double R = cfg->R;
    Mt19937& rng = *jb.rng;
    jb.mo = jb.mrows;
    jb.da = jb.mrows;
    std::vector<double> txy1(nd, 0.0), tz1(nd, 0.0), txy2(nd, 0.0), tz2(nd, 0.0);
    for (int i = 0; i < nd; ++i) {
        if (jb.ntries < 10 - 1 && jb.check[i] == 1) {
            txy1[i] = dr.angle_xy(rng);
            tz1[i] = dr.angle_z(rng);
            txy2[i] = dr.angle_xy(rng);
            tz2[i] = dr.angle_z(rng);
        }
    }
    for (int i = 0; i < nd; ++i) {
        const double* row = &jb.mrows[(size_t)i * ncols];
        double u[3] = {row[C_NX], row[C_NY], row[C_NZ]}, w[3], v[3];
        rotate_xy(u, txy1[i], w);
        rotate_out_of_xy(w, tz1[i], v);
        double* m = &jb.mo[(size_t)i * ncols];
        m[C_NX] = v[0]; m[C_NY] = v[1]; m[C_NZ] = v[2];
        rotate_xy(u, txy2[i], w);
        rotate_out_of_xy(w, tz2[i], v);
        double* d = &jb.da[(size_t)i * ncols];
        d[C_NX] = v[0]; d[C_NY] = v[1]; d[C_NZ] = v[2];
    }
    std::vector<double> M(nd);
    for (int i = 0; i < nd; ++i) M[i] = dr.daughter_length(rng);
    for (int i = 0; i < nd; ++i) {
        double* m = &jb.mo[(size_t)i * ncols];
        double* d = &jb.da[(size_t)i * ncols];
        double Ld = m[C_L] - 2 * R;
        double L1 = M[i] * Ld;
        double L2 = (1 - M[i]) * Ld;
        double div = -m[C_HALF_L] + L1 + R;
        double delta1 = L1 / 2 + R;
        double delta2 = L2 / 2 + R;
        m[C_L] = L1;
        m[C_HALF_L] = L1 / 2;
        d[C_L] = L2;
        d[C_HALF_L] = L2 / 2;
        m[C_RX] = m[C_RX] + (div - delta1) * m[C_NX];
        d[C_RX] = d[C_RX] + (div + delta2) * d[C_NX];
        m[C_RY] = m[C_RY] + (div - delta1) * m[C_NY];
        d[C_RY] = d[C_RY] + (div + delta2) * d[C_NY];
        m[C_RZ] = m[C_RZ] + (div - delta1) * m[C_NZ];
        d[C_RZ] = d[C_RZ] + (div + delta2) * d[C_NZ];
        m[C_T0] = jb.t;
        d[C_T0] = jb.t;
    }
    for (int i = 0; i < nd; ++i) {
        double* m = &jb.mo[(size_t)i * ncols];
        m[C_GROWTH] = dr.growth(rng, (int)m[C_GROUP]);
    }
    for (int i = 0; i < nd; ++i) {
        double* d = &jb.da[(size_t)i * ncols];
        d[C_GROWTH] = dr.growth(rng, (int)d[C_GROUP]);
    }
}

// divideCells for every job: the no-room check, then placement rounds until
// each job has its criterion met or ten tries, then ids and lineage
// (growth.hpp:700-712). Each round issues one base_min call for all jobs.
void divide_batch(const bfm_sim_config* cfg, int ncols, std::vector<DivisionJob>& jobs, const BaseMinFn& base_min)
{
    const double cut = 2 * cfg->Rcell;
    {
        std::vector<Query> q;
        for (auto& jb : jobs) {
            jb.nd = (int)jb.idx.size();
            jb.check.assign(jb.nd, 1);
            jb.satisfies.assign(jb.nd, 0);
            jb.ntries = 0;
            for (int i = 0; i < jb.nd; ++i)
                q.push_back(query_of_row(&jb.mrows[(size_t)i * ncols], jb.idx[i], jb.lo, jb.hi, 0));
        }
        std::vector<double> md;
        if (!q.empty()) base_min(q, cut, md);
        size_t k = 0;
        for (auto& jb : jobs)
            for (int i = 0; i < jb.nd; ++i, ++k) {
                if (md[k] < cut) jb.check[i] = 0;
                if (jb.check[i] == 0) jb.satisfies[i] = 1;
            }
    }
    for (;;) {
        std::vector<Query> q;
        std::vector<DivisionJob*> round;
        for (auto& jb : jobs) {
            if (!jb.active()) continue;
            round.push_back(&jb);
            propose(cfg, ncols, jb);
            jb.qcell.clear();
            jb.q0 = q.size();
            for (int i = 0; i < jb.nd; ++i) {
                if (jb.check[i] != 1) continue;
                q.push_back(query_of_row(&jb.mo[(size_t)i * ncols], -1, jb.lo, jb.hi, 1));
                jb.qcell.push_back(i);
                q.push_back(query_of_row(&jb.da[(size_t)i * ncols], -1, jb.lo, jb.hi, 1));
                jb.qcell.push_back(jb.nd + i);
            }
        }
        if (round.empty()) break;
        std::vector<double> md;
        if (!q.empty()) base_min(q, cut, md);
        for (DivisionJob* jp : round) {
            DivisionJob& jb = *jp;
            const int nd = jb.nd;
            for (size_t k = 0; k < jb.qcell.size(); ++k) {
                const Query& qq = q[jb.q0 + k];
                Seg a;
                std::memcpy(a.r, qq.r, sizeof(a.r));
                std::memcpy(a.n, qq.n, sizeof(a.n));
                a.hl = qq.hl;
                double& mk = md[jb.q0 + k];
                for (int r = 0; r < 2 * nd; ++r) {
                    if (r == jb.qcell[k]) continue;
                    const double* row = r < nd ? &jb.mo[(size_t)r * ncols] : &jb.da[(size_t)(r - nd) * ncols];
                    Seg b = seg_of_row(row);
                    if (beyond_cut(a, b, cut)) continue;
                    double dist = seg_dist(a, b);
                    if (dist < mk) mk = dist;
                }
            }
            for (size_t k = 0; k + 1 < jb.qcell.size(); k += 2) {
                const int i = jb.qcell[k];
                if (md[jb.q0 + k] >= cut && md[jb.q0 + k + 1] >= cut) jb.satisfies[i] = 1;
            }
            jb.ntries++;
        }
    }
    for (auto& jb : jobs) {
        std::vector<int>& parents = *jb.parents;
        int max_id = (int)parents.size() - 1;
        for (int i = 0; i < jb.nd; ++i) {
            int id1 = max_id + (2 * i + 1);
            int id2 = max_id + (2 * i + 2);
            int parent_id = (int)jb.mo[(size_t)i * ncols + C_ID];
            jb.mo[(size_t)i * ncols + C_ID] = id1;
            jb.da[(size_t)i * ncols + C_ID] = id2;
            parents.push_back(parent_id);
            parents.push_back(parent_id);
        }
    }
}

// ---------------------------------------------------------------------------
// device helpers of the loop
// ---------------------------------------------------------------------------

// growCells (growth.hpp:51-58).
__global__ void k_grow(double* __restrict__ cells, int ld, int n, double R, double dt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double l = cells[(size_t)C_L * ld + i];
    double g = cells[(size_t)C_GROWTH * ld + i];
    l = l + (g * (4 * R / 3 + l)) * dt;
    cells[(size_t)C_L * ld + i] = l;
    cells[(size_t)C_HALF_L * ld + i] = l / 2;
}

// divideMaxLength (growth.hpp:68-73): indices of cells with l > Ldiv, in
// arbitrary order (the host sorts them).
__global__ void k_find_dividing(const double* __restrict__ cells, int ld, int n, double Ldiv,
                                int* __restrict__ count, int* __restrict__ idx)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (cells[(size_t)C_L * ld + i] > Ldiv) idx[atomicAdd(count, 1)] = i;
}

__global__ void k_zero_int(int* p) { *p = 0; }

__global__ void k_copy_int(const int* __restrict__ src, int* __restrict__ dst) { *dst = *src; }

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_set_flags(int* __restrict__ flags, const int* __restrict__ idx, int k, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) flags[idx[i]] = v;
}

// rows (k x ncols, column-major, ld k) <- cells[idx[r], :]
__global__ void k_gather_rows(const double* __restrict__ cells, int ld, int ncols, const int* __restrict__ idx,
                              int k, double* __restrict__ rows)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k * ncols) return;
    int r = t % k, c = t / k;
    rows[(size_t)c * k + r] = cells[(size_t)c * ld + idx[r]];
}

// cells[idx[r], :] <- rows (k x ncols, column-major, ld k)
__global__ void k_scatter_rows(double* __restrict__ cells, int ld, int ncols, const int* __restrict__ idx, int k,
                               const double* __restrict__ rows)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k * ncols) return;
    int r = t % k, c = t / k;
    cells[(size_t)c * ld + idx[r]] = rows[(size_t)c * k + r];
}

// Minimum |d| from each query segment (7 doubles: r, n, half_l) to the cells
// j in [lo, hi) of its replicate with j != self and, when the query asks for
// it, skip[j] == 0 (qi4 = {self, lo, hi, skip?}). Pairs beyond `cut` are
// screened out, so only comparisons with cut are exact. Out holds the bit
// patterns of non-negative doubles (ordered like the values).
__global__ void __launch_bounds__(256)
k_min_dist(const double* __restrict__ cells, int ld, const double* __restrict__ q, const int4* __restrict__ qi4,
           const int* __restrict__ skip, double cut, unsigned long long* __restrict__ out)
{
    const int qi = blockIdx.y;
    Seg a;
    a.r[0] = q[qi * 7 + 0]; a.r[1] = q[qi * 7 + 1]; a.r[2] = q[qi * 7 + 2];
    a.n[0] = q[qi * 7 + 3]; a.n[1] = q[qi * 7 + 4]; a.n[2] = q[qi * 7 + 5];
    a.hl = q[qi * 7 + 6];
    const int4 meta = qi4[qi];
    const int self = meta.x;
    const bool use_skip = meta.w != 0;
    double best = INFINITY;
    for (int j = meta.y + blockIdx.x * blockDim.x + threadIdx.x; j < meta.z; j += gridDim.x * blockDim.x) {
        if (j == self || (use_skip && skip[j])) continue;
        Seg b;
        b.r[0] = cells[(size_t)C_RX * ld + j]; b.r[1] = cells[(size_t)C_RY * ld + j];
        b.r[2] = cells[(size_t)C_RZ * ld + j];
        b.n[0] = cells[(size_t)C_NX * ld + j]; b.n[1] = cells[(size_t)C_NY * ld + j];
        b.n[2] = cells[(size_t)C_NZ * ld + j];
        b.hl = cells[(size_t)C_HALF_L * ld + j];
        if (beyond_cut(a, b, cut)) continue;
        double d[3], s, t;
        dist_between_cells(a, b, d, s, t);
        double dist = norm3(d[0], d[1], d[2]);
        if (dist < best) best = dist;
    }
    for (int off = 32; off > 0; off >>= 1) best = fmin(best, __shfl_xor(best, off));
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
        atomicMin(&out[qi], (unsigned long long)__double_as_longlong(m));
    }
}

// d_rep[i] = r for the rows [starts[r], starts[r + 1]) of replicate r.
__global__ void k_fill_rep(int* __restrict__ rep, const int* __restrict__ starts)
{
    const int r = blockIdx.x;
    for (int i = starts[r] + threadIdx.x; i < starts[r + 1]; i += blockDim.x) rep[i] = r;
}

// Moves every row i (replicate rep[i]) to i + shift[rep[i]]: replicates stay
// contiguous and in order after each gains its daughters at its end.
__global__ void k_repack(const double* __restrict__ src, double* __restrict__ dst, int ld, int ncols, int n,
                         const int* __restrict__ rep, const int* __restrict__ shift)
{
    const int64_t total = (int64_t)n * ncols;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % n), col = (int)(t / n);
        dst[(int64_t)col * ld + i + shift[rep[i]]] = src[(int64_t)col * ld + i];
    }
}

int blocks_for(int64_t n, int b) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

// ---------------------------------------------------------------------------
// the loop
// ---------------------------------------------------------------------------

// Device side of the division placement checks: one k_min_dist launch per
// batch of queries (every replicate's queries of one placement round).
struct DeviceMinDist {
    hipStream_t st;
    int cap;
    double* d_q = nullptr;
    int4* d_meta = nullptr;
    unsigned long long* d_out = nullptr;
    DeviceMinDist(int capacity, hipStream_t stream) : st(stream), cap(std::max(capacity, 16))
    {
        SCHK(hipMalloc((void**)&d_q, sizeof(double) * 7 * (size_t)cap));
        SCHK(hipMalloc((void**)&d_meta, sizeof(int4) * (size_t)cap));
        SCHK(hipMalloc((void**)&d_out, sizeof(unsigned long long) * (size_t)cap));
    }
    ~DeviceMinDist()
    {
        (void)hipFree(d_q);
        (void)hipFree(d_meta);
        (void)hipFree(d_out);
    }
    void run(const std::vector<Query>& q, const double* cells, int ld, const int* skip, double cut,
             std::vector<double>& out)
    {
        out.resize(q.size());
        for (size_t k0 = 0; k0 < q.size(); k0 += (size_t)cap) {
            const int nq = (int)std::min(q.size() - k0, (size_t)cap);
            std::vector<double> hq((size_t)nq * 7);
            std::vector<int4> hm(nq);
            int span = 1;
            for (int k = 0; k < nq; ++k) {
                const Query& qq = q[k0 + k];
                for (int a = 0; a < 3; ++a) {
                    hq[(size_t)k * 7 + a] = qq.r[a];
                    hq[(size_t)k * 7 + 3 + a] = qq.n[a];
                }
                hq[(size_t)k * 7 + 6] = qq.hl;
                hm[k] = make_int4(qq.self, qq.lo, qq.hi, qq.skip_dividing);
                span = std::max(span, qq.hi - qq.lo);
            }
            SCHK(hipMemcpyAsync(d_q, hq.data(), sizeof(double) * hq.size(), hipMemcpyHostToDevice, st));
            SCHK(hipMemcpyAsync(d_meta, hm.data(), sizeof(int4) * nq, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(nq, 256)), dim3(256), 0, st, d_out, nq,
                               (unsigned long long)0x7FF0000000000000ull);
            const int gx = std::min(blocks_for(span, 256), 64);
            hipLaunchKernelGGL(k_min_dist, dim3(gx, nq), dim3(256), 0, st, cells, ld, d_q, d_meta, skip, cut, d_out);
            SCHK(hipGetLastError());
            std::vector<unsigned long long> bits(nq);
            SCHK(hipMemcpyAsync(bits.data(), d_out, sizeof(unsigned long long) * nq, hipMemcpyDeviceToHost, st));
            SCHK(hipStreamSynchronize(st));
            for (int k = 0; k < nq; ++k) std::memcpy(&out[k0 + k], &bits[k], sizeof(double));
        }
    }
};
struct TableauDef {
    int s;
    double A[49], b[7], bs[7];
    double order;
};

// Butcher tableaux of runSimulation (simulation.hpp:1207-1268).
TableauDef tableau(int which)
{
    TableauDef t;
    std::memset(&t, 0, sizeof(t));
    auto set = [&](int s, std::initializer_list<double> A, std::initializer_list<double> b,
                   std::initializer_list<double> bs, double order) {
        t.s = s;
        int k = 0;
        for (double v : A) t.A[k++] = v;
        k = 0;
        for (double v : b) t.b[k++] = v;
        k = 0;
        for (double v : bs) t.bs[k++] = v;
        t.order = order;
    };
    switch (which) {
    case 0:
        set(2, {0, 0, 1, 0}, {0.5, 0.5}, {1, 0}, 1);
        break;
    case 1:
        set(4, {0, 0, 0, 0, 1. / 2., 0, 0, 0, 0, 3. / 4., 0, 0, 2. / 9., 1. / 3., 4. / 9., 0},
            {2. / 9., 1. / 3., 4. / 9., 0}, {7. / 24., 1. / 4., 1. / 3., 1. / 8.}, 2);
        break;
    case 2:
        set(6,
            {0, 0, 0, 0, 0, 0,
             1. / 4., 0, 0, 0, 0, 0,
             3. / 32., 9. / 32., 0, 0, 0, 0,
             1932. / 2197., -7200. / 2197., 7296. / 2197., 0, 0, 0,
             439. / 216., -8.0, 3680. / 513., -845. / 4104., 0, 0,
             -8. / 27., 2.0, -3544. / 2565., 1859. / 4104., -11. / 40., 0},
            {16. / 135., 0, 6656. / 12825., 28561. / 56430., -9. / 50., 2. / 55.},
            {25. / 216., 0, 1408. / 2565., 2197. / 4104., -1. / 5., 0}, 4);
        break;
    case 3:
        set(7,
            {0, 0, 0, 0, 0, 0, 0,
             1. / 5., 0, 0, 0, 0, 0, 0,
             3. / 40., 9. / 40., 0, 0, 0, 0, 0,
             44. / 45., -56. / 15., 32. / 9., 0, 0, 0, 0,
             19372. / 6561., -25360. / 2187., 64448. / 6561., -212. / 729, 0, 0, 0,
             9017. / 3168., -355. / 33., 46732. / 5247., 49. / 176., -5103. / 18656., 0, 0,
             35. / 384., 0, 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84., 0},
            {35. / 384., 0, 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84., 0},
            {5179. / 57600., 0, 7571. / 16695., 393. / 640., -92097. / 339200., 187. / 2100., 1. / 40.}, 4);
        break;
    default:
        sim_fail(BFM_ERR_ARG, "tableau must be 0..3");
    }
    return t;
}

// Termination rule of runSimulation (simulation.hpp:1574-1614): any of the
// positive criteria; none given terminates at once.
bool terminate_now(const bfm_sim_config* cfg, int n, int64_t iter, double t)
{
    bool has_n = cfg->n_cells > 0, has_i = cfg->max_iter > 0, has_t = cfg->max_time > 0;
    if (!has_n && !has_i && !has_t) return true;
    return (has_n && n >= cfg->n_cells) || (has_i && iter >= cfg->max_iter) || (has_t && t >= cfg->max_time);
}

void check_config(const bfm_sim_config* cfg, int ncols)
{
    if (!cfg) sim_fail(BFM_ERR_ARG, "NULL config");
    if (cfg->iter_update_neighbors < 1 || cfg->iter_update_stepsize < 1)
        sim_fail(BFM_ERR_ARG, "iter_update_neighbors and iter_update_stepsize must be >= 1");
    if (cfg->max_tries_update_stepsize < 0) sim_fail(BFM_ERR_ARG, "max_tries_update_stepsize < 0");
    if (!(cfg->max_stepsize > 0) || !(cfg->min_stepsize > 0) || cfg->min_stepsize > cfg->max_stepsize)
        sim_fail(BFM_ERR_ARG, "need 0 < min_stepsize <= max_stepsize");
    if (!(cfg->max_error_allowed > 0)) sim_fail(BFM_ERR_ARG, "max_error_allowed must be > 0");
    if (ncols < 21 || ncols > 32) sim_fail(BFM_ERR_ARG, "ncols must be in [21, 32]");
}

struct PinnedBuf {
    void* p = nullptr;
    explicit PinnedBuf(size_t bytes) { SCHK(hipHostMalloc(&p, bytes, hipHostMallocDefault)); }
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    explicit DevBuf(size_t count) { SCHK(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T))); }
    ~DevBuf() { if (p) (void)hipFree(p); }
};

// Host memory the device writes the per-attempt readback into.
struct MappedBuf {
    void* h = nullptr;
    void* d = nullptr;
    explicit MappedBuf(size_t bytes)
    {
        SCHK(hipHostMalloc(&h, bytes, hipHostMallocMapped));
        std::memset(h, 0, bytes);
        SCHK(hipHostGetDevicePointer(&d, h, 0));
    }
    ~MappedBuf() { if (h) (void)hipHostFree(h); }
};

// ---------------------------------------------------------------------------
// per-replicate step-size controller on the device
// ---------------------------------------------------------------------------

// The controller's steps as kernels of their own (one work-group each), for
// batches whose attempt and tail kernels do not run them in their last
// work-group.

// The iteration's first attempt: replicates with go[r] (not finished on the
// host's counts) and not failed take dt; control selects step control; the
// tail's division count is zeroed.
__global__ void __launch_bounds__(256)
k_ctrl_begin(CtrlState S, CtrlConsts K, const int* __restrict__ go, int control, const int* __restrict__ starts,
             double* __restrict__ trial, int* __restrict__ active, unsigned long long* __restrict__ errmax,
             CtrlPub* pub, int* __restrict__ div_count)
{
    int pending = 0;
    long long cells = 0;
    if (threadIdx.x == 0) *div_count = 0;
    for (int r = threadIdx.x; r < K.n_rep; r += blockDim.x) {
        errmax[r] = kKeyNegInf;
        const bool on = go[r] && !S.failed[r];
        S.stepped[r] = on;
        active[r] = on;
        S.state[r] = on ? (control ? CS_FIRST : CS_FINAL) : CS_DONE;
        S.factor[r] = 1;
        S.tries[r] = 0;
        trial[r] = S.dt[r];
        if (on) { ++pending; cells += starts[r + 1] - starts[r]; }
    }
    ctrl_publish(pending, cells, 0, 0, pub, S.pending, S.totals);
}

__global__ void __launch_bounds__(256) k_ctrl_after(CtrlAfter A) { ctrl_after_block(A); }

// ---------------------------------------------------------------------------
// divideCells on the device (growth.hpp:456-720), all replicates at once
// ---------------------------------------------------------------------------

// Each replicate's MT19937 stream lives in device memory (624 words and the
// index) and is advanced by one wavefront: every lane computes the same draws
// (wave-uniform control flow), the 624-word twist is split over the lanes in
// the three phases its recurrence allows (words 0-226 read only old words,
// 227-453 read words 0-226 of this twist, 454-622 read words 227-395, and
// word 623 reads words 0 and 396), so the stream equals the sequential
// generator's. uniform_01 = x 2^-32 and the Box-Muller normal follow the host
// restatement (utils.hpp:648-655).
constexpr int kMtWords = 625;   // 624 state words + the index

__device__ __forceinline__ void wave_sync_sim()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct WaveMt {
    uint32_t* mt;   // 624 words in LDS
    int idx;        // uniform across the wavefront
    __device__ void twist(int lane)
    {
        auto step = [&](int i) __attribute__((always_inline)) {
            const uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            return v;
        };
        for (int b = 0; b < 227; b += 64) {
            const int i = b + lane;
            uint32_t v = 0;
            if (i < 227) v = step(i);
            wave_sync_sim();
            if (i < 227) mt[i] = v;
            wave_sync_sim();
        }
        for (int b = 227; b < 454; b += 64) {
            const int i = b + lane;
            uint32_t v = 0;
            if (i < 454) v = step(i);
            wave_sync_sim();
            if (i < 454) mt[i] = v;
            wave_sync_sim();
        }
        for (int b = 454; b < 623; b += 64) {
            const int i = b + lane;
            uint32_t v = 0;
            if (i < 623) v = step(i);
            wave_sync_sim();
            if (i < 623) mt[i] = v;
            wave_sync_sim();
        }
        uint32_t v = step(623);
        wave_sync_sim();
        if (lane == 0) mt[623] = v;
        wave_sync_sim();
        idx = 0;
    }
    __device__ uint32_t next(int lane)
    {
        if (idx >= 624) twist(lane);
        uint32_t y = mt[idx++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
    __device__ double uniform01(int lane) { return (double)next(lane) * (1.0 / 4294967296.0); }
    __device__ double normal(int lane)
    {
        const double u = uniform01(lane);
        const double v = uniform01(lane);
        const double c = sqrt(-2 * log(u));
        return c * cos(6.283185307179586476925286766559 * v);
    }
};

// The distributions of runSimulation (simulation.hpp:1294-1339) and the
// geometry of divideCells.
struct DivParams {
    double R, cut;                       // cut = 2 Rcell
    double gmean[2], gstd[2];
    double dl_std, axy, az;
    int ncols;
};

// The dividing rows of every replicate that steps, ascending, from the
// sorted list of rows past the division length: job r covers
// dlist[doff[r] .. doff[r] + nd[r]); rows of replicates that do not step are
// left out (nd = 0). One block.
__global__ void __launch_bounds__(256)
k_div_jobs(int R, const int* __restrict__ starts, const int* __restrict__ go, const int* __restrict__ sorted_rows,
           int nd_all, int* __restrict__ doff, int* __restrict__ nd, int* __restrict__ ntries,
           int* __restrict__ jobact)
{
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        int lo = 0, hi = nd_all;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sorted_rows[mid] < starts[r]) lo = mid + 1; else hi = mid;
        }
        int e = lo;
        hi = nd_all;
        while (e < hi) {
            const int mid = (e + hi) >> 1;
            if (sorted_rows[mid] < starts[r + 1]) e = mid + 1; else hi = mid;
        }
        doff[r] = lo;
        nd[r] = go[r] ? e - lo : 0;
        ntries[r] = 0;
        jobact[r] = nd[r] > 0;
    }
}

__device__ __forceinline__ int replicate_of(const int* __restrict__ starts, int R, int row)
{
    int lo = 0, hi = R - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= row) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// The no-room check's queries (growth.hpp:480-520): each dividing cell's
// segment against every other cell of its replicate; flags mark the rows
// being divided (left out of the placement checks).
__global__ void k_div_start(int R, int nd_all, const int* __restrict__ starts, const int* __restrict__ sorted_rows,
                            const int* __restrict__ doff, const int* __restrict__ nd, const double* __restrict__ cells,
                            int ld, int* __restrict__ flags, int* __restrict__ check, int* __restrict__ sat,
                            double* __restrict__ q, int4* __restrict__ meta)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nd_all) return;
    const int row = sorted_rows[k];
    const int r = replicate_of(starts, R, row);
    const bool on = k - doff[r] < nd[r];
    check[k] = 1;
    sat[k] = 0;
    if (on) flags[row] = 1;
    for (int a = 0; a < 3; ++a) {
        q[(int64_t)k * 7 + a] = cells[(int64_t)(C_RX + a) * ld + row];
        q[(int64_t)k * 7 + 3 + a] = cells[(int64_t)(C_NX + a) * ld + row];
    }
    q[(int64_t)k * 7 + 6] = cells[(int64_t)C_HALF_L * ld + row];
    meta[k] = on ? make_int4(row, starts[r], starts[r + 1], 0) : make_int4(-1, 0, 0, 0);
}

__device__ __forceinline__ double dist_key_value(unsigned long long b) { return __longlong_as_double((long long)b); }

__global__ void k_div_noroom(int nd_all, const unsigned long long* __restrict__ md, double cut,
                             int* __restrict__ check, int* __restrict__ sat)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nd_all) return;
    if (dist_key_value(md[k]) < cut) check[k] = 0;
    if (check[k] == 0) sat[k] = 1;
}

// rotateXY and rotateOutOfXY (utils.hpp:791-829), as the host restatement.
__device__ __forceinline__ void dev_rotate_xy(const double n[3], double theta, double out[3])
{
    double s = sin(theta), c = cos(theta);
    out[0] = (c * n[0] + (-s) * n[1]) + 0.0 * n[2];
    out[1] = (s * n[0] + c * n[1]) + 0.0 * n[2];
    out[2] = (0.0 * n[0] + 0.0 * n[1]) + 1.0 * n[2];
}

__device__ __forceinline__ void dev_rotate_out_of_xy(const double n[3], double theta, double out[3])
{
    double v0 = n[1] * 1.0 - n[2] * 0.0;
    double v1 = n[2] * 0.0 - n[0] * 1.0;
    double v2 = n[0] * 0.0 - n[1] * 0.0;
    double nrm = sqrt((v0 * v0 + v1 * v1) + v2 * v2);
    v0 /= nrm; v1 /= nrm; v2 /= nrm;
    double s = sin(theta), c = cos(theta);
    double w0 = v1 * n[2] - v2 * n[1];
    double w1 = v2 * n[0] - v0 * n[2];
    double w2 = v0 * n[1] - v1 * n[0];
    out[0] = n[0] * c + w0 * s;
    out[1] = n[1] * c + w1 * s;
    out[2] = n[2] * c + w2 * s;
}

// One placement attempt per active job (one wavefront per replicate): the
// re-orientation draws (four per cell with room, none on the last try), the
// length ratios, the mothers' then the daughters' growth rates, in the
// reference's order; the proposal rows mo / da (column-major, leading
// dimension ldp) and the queries of the cells with room (2k: mother, 2k + 1:
// daughter) against the unchanged rows of the replicate.
__global__ void __launch_bounds__(64)
k_div_propose(int R, DivParams P, double t_unused, const int* __restrict__ starts,
              const int* __restrict__ sorted_rows, const int* __restrict__ doff, const int* __restrict__ nd,
              const int* __restrict__ ntries, const int* __restrict__ jobact, const int* __restrict__ check,
              const double* __restrict__ t_rep, const double* __restrict__ cells, int ld,
              uint32_t* __restrict__ mt_state, double* __restrict__ mo, double* __restrict__ da, int ldp,
              double* __restrict__ q, int4* __restrict__ meta, int* __restrict__ status)
{
    __shared__ uint32_t s_mt[624];
    const int r = blockIdx.x, lane = threadIdx.x;
    if (r >= R) return;
    const int n0 = doff[r], m = nd[r], tries = ntries[r];
    if (!jobact[r]) {   // a finished job asks nothing
        for (int i = lane; i < m; i += 64) {
            meta[2 * (int64_t)(n0 + i)] = make_int4(-1, 0, 0, 0);
            meta[2 * (int64_t)(n0 + i) + 1] = make_int4(-1, 0, 0, 0);
        }
        return;
    }
    uint32_t* gst = mt_state + (int64_t)r * kMtWords;
    for (int i = lane; i < 624; i += 64) s_mt[i] = gst[i];
    WaveMt g{s_mt, (int)gst[624]};
    wave_sync_sim();
    const double t = t_rep[r];
    // rows start as the mothers' originals
    for (int col = 0; col < P.ncols; ++col)
        for (int i = lane; i < m; i += 64) {
            const double v = cells[(int64_t)col * ld + sorted_rows[n0 + i]];
            mo[(int64_t)col * ldp + n0 + i] = v;
            da[(int64_t)col * ldp + n0 + i] = v;
        }
    wave_sync_sim();
    // re-orientation draws, then the rotations (lane i % 64 writes cell i)
    for (int i = 0; i < m; ++i) {
        double txy1 = 0, tz1 = 0, txy2 = 0, tz2 = 0;
        if (tries < 10 - 1 && check[n0 + i] == 1) {
            txy1 = -P.axy + 2 * P.axy * g.uniform01(lane);
            tz1 = -P.az + 2 * P.az * g.uniform01(lane);
            txy2 = -P.axy + 2 * P.axy * g.uniform01(lane);
            tz2 = -P.az + 2 * P.az * g.uniform01(lane);
        }
        if (lane == (i & 63)) {
            const int64_t k = n0 + i;
            const double u[3] = {mo[(int64_t)C_NX * ldp + k], mo[(int64_t)C_NY * ldp + k], mo[(int64_t)C_NZ * ldp + k]};
            double w[3], v[3];
            dev_rotate_xy(u, txy1, w);
            dev_rotate_out_of_xy(w, tz1, v);
            mo[(int64_t)C_NX * ldp + k] = v[0]; mo[(int64_t)C_NY * ldp + k] = v[1]; mo[(int64_t)C_NZ * ldp + k] = v[2];
            dev_rotate_xy(u, txy2, w);
            dev_rotate_out_of_xy(w, tz2, v);
            da[(int64_t)C_NX * ldp + k] = v[0]; da[(int64_t)C_NY * ldp + k] = v[1]; da[(int64_t)C_NZ * ldp + k] = v[2];
        }
    }
    // length ratios (clipped N(0.5, std)), positions and lengths
    for (int i = 0; i < m; ++i) {
        double M = 0.5 + P.dl_std * g.normal(lane);
        M = M < 0 ? 0.0 : (M > 1 ? 1.0 : M);
        if (lane == (i & 63)) {
            const int64_t k = n0 + i;
            const double l = mo[(int64_t)C_L * ldp + k], hl = mo[(int64_t)C_HALF_L * ldp + k];
            const double Ld = l - 2 * P.R;
            const double L1 = M * Ld, L2 = (1 - M) * Ld;
            const double div = -hl + L1 + P.R;
            const double delta1 = L1 / 2 + P.R, delta2 = L2 / 2 + P.R;
            mo[(int64_t)C_L * ldp + k] = L1;
            mo[(int64_t)C_HALF_L * ldp + k] = L1 / 2;
            da[(int64_t)C_L * ldp + k] = L2;
            da[(int64_t)C_HALF_L * ldp + k] = L2 / 2;
            for (int a = 0; a < 3; ++a) {
                const double mn = mo[(int64_t)(C_NX + a) * ldp + k], dn = da[(int64_t)(C_NX + a) * ldp + k];
                mo[(int64_t)(C_RX + a) * ldp + k] = mo[(int64_t)(C_RX + a) * ldp + k] + (div - delta1) * mn;
                da[(int64_t)(C_RX + a) * ldp + k] = da[(int64_t)(C_RX + a) * ldp + k] + (div + delta2) * dn;
            }
            mo[(int64_t)C_T0 * ldp + k] = t;
            da[(int64_t)C_T0 * ldp + k] = t;
        }
    }
    // growth rates: the mothers', then the daughters'
    for (int pass = 0; pass < 2; ++pass) {
        double* rows = pass == 0 ? mo : da;
        for (int i = 0; i < m; ++i) {
            const int64_t k = n0 + i;
            const int group = (int)rows[(int64_t)C_GROUP * ldp + k];
            const bool okg = group == 1 || group == 2;
            const double z = g.normal(lane);
            if (lane == (i & 63)) {
                if (okg) rows[(int64_t)C_GROWTH * ldp + k] = P.gmean[group - 1] + P.gstd[group - 1] * z;
                else atomicOr(status, 1);
            }
        }
    }
    wave_sync_sim();
    for (int i = lane; i < 624; i += 64) gst[i] = s_mt[i];
    if (lane == 0) gst[624] = (uint32_t)g.idx;
    // queries of the cells with room
    for (int i = lane; i < m; i += 64) {
        const int64_t k = n0 + i;
        for (int h = 0; h < 2; ++h) {
            const double* rows = h == 0 ? mo : da;
            const int64_t qk = 2 * k + h;
            for (int a = 0; a < 3; ++a) {
                q[qk * 7 + a] = rows[(int64_t)(C_RX + a) * ldp + k];
                q[qk * 7 + 3 + a] = rows[(int64_t)(C_NX + a) * ldp + k];
            }
            q[qk * 7 + 6] = rows[(int64_t)C_HALF_L * ldp + k];
            meta[qk] = check[k] == 1 ? make_int4(-1, starts[r], starts[r + 1], 1) : make_int4(-1, 0, 0, 0);
        }
    }
}

__device__ __forceinline__ Seg seg_of_prop(const double* __restrict__ rows, int ldp, int64_t k)
{
    Seg s;
    s.r[0] = rows[(int64_t)C_RX * ldp + k]; s.r[1] = rows[(int64_t)C_RY * ldp + k];
    s.r[2] = rows[(int64_t)C_RZ * ldp + k];
    s.n[0] = rows[(int64_t)C_NX * ldp + k]; s.n[1] = rows[(int64_t)C_NY * ldp + k];
    s.n[2] = rows[(int64_t)C_NZ * ldp + k];
    s.hl = rows[(int64_t)C_HALF_L * ldp + k];
    return s;
}

// Completes an attempt (one wavefront per active job): each cell with room
// takes the minimum over the other new rows of its replicate as well, and
// is satisfied when its mother and daughter are both at least cut from
// everything; the job goes on while a cell is unsatisfied and fewer than ten
// attempts were made. any_active counts the jobs that go on.
__global__ void __launch_bounds__(64)
k_div_eval(int R, double cut, const int* __restrict__ doff, const int* __restrict__ nd, int* __restrict__ ntries,
           int* __restrict__ jobact, const int* __restrict__ check, int* __restrict__ sat,
           const double* __restrict__ mo, const double* __restrict__ da, int ldp,
           const unsigned long long* __restrict__ md, int* __restrict__ any_active)
{
    const int r = blockIdx.x, lane = threadIdx.x;
    if (r >= R || !jobact[r]) return;
    const int n0 = doff[r], m = nd[r];
    int unsat = 0;
    for (int i = lane; i < m; i += 64) {
        const int64_t k = n0 + i;
        if (check[k] == 1) {
            double best[2];
            for (int h = 0; h < 2; ++h) {
                const Seg a = seg_of_prop(h == 0 ? mo : da, ldp, k);
                double b = dist_key_value(md[2 * k + h]);
                for (int x = 0; x < 2 * m; ++x) {
                    if (x == i + h * m) continue;
                    const Seg o = seg_of_prop(x < m ? mo : da, ldp, n0 + (x < m ? x : x - m));
                    if (beyond_cut(a, o, cut)) continue;
                    const double dd = seg_dist(a, o);
                    if (dd < b) b = dd;
                }
                best[h] = b;
            }
            if (best[0] >= cut && best[1] >= cut) sat[k] = 1;
        }
        unsat += sat[k] == 0;
    }
    for (int off = 32; off > 0; off >>= 1) unsat += __shfl_xor(unsat, off);
    if (lane == 0) {
        const int tr = ntries[r] + 1;
        ntries[r] = tr;
        const int on = tr < 10 && unsat > 0;
        jobact[r] = on;
        if (on) atomicAdd(any_active, 1);
    }
}

// New ids (growth.hpp:700-712: the mother of pair i gets max_id + 2i + 1,
// the daughter max_id + 2i + 2, both the parent's id as parent), appended to
// the lineage log as (replicate, id, parent); the new layout (each replicate
// gains its daughters at its end) and the row shifts. One block.
__global__ void __launch_bounds__(256)
k_div_layout(int R, const int* __restrict__ doff, const int* __restrict__ nd, double* __restrict__ mo,
             double* __restrict__ da, int ldp, long long* __restrict__ next_id, int* __restrict__ log_n,
             int4* __restrict__ lineage, int log_cap, int* __restrict__ starts, int* __restrict__ shift,
             int* __restrict__ old_end, int* __restrict__ d_n, int* __restrict__ sizes_out, int* __restrict__ status)
{
    __shared__ int s_tot[256];
    // ids and lineage
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        const int m = nd[r], n0 = doff[r];
        if (m == 0) continue;
        const long long base = next_id[r];
        const int pos = atomicAdd(log_n, 2 * m);
        for (int i = 0; i < m; ++i) {
            const int64_t k = n0 + i;
            const int parent = (int)mo[(int64_t)C_ID * ldp + k];
            const long long id1 = base + 2 * i, id2 = base + 2 * i + 1;
            mo[(int64_t)C_ID * ldp + k] = (double)id1;
            da[(int64_t)C_ID * ldp + k] = (double)id2;
            if (pos + 2 * i + 1 < log_cap) {
                lineage[pos + 2 * i] = make_int4(r, (int)id1, parent, 0);
                lineage[pos + 2 * i + 1] = make_int4(r, (int)id2, parent, 0);
            } else {
                atomicOr(status, 2);
            }
        }
        next_id[r] = base + 2 * m;
    }
    // new starts: an exclusive scan of the new sizes, in chunks of the block
    __syncthreads();
    int carry = 0;
    for (int r0 = 0; r0 < R; r0 += blockDim.x) {
        const int r = r0 + threadIdx.x;
        const int sz = r < R ? (starts[r + 1] - starts[r]) + nd[r] : 0;
        s_tot[threadIdx.x] = sz;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {
            const int v = threadIdx.x >= (unsigned)o ? s_tot[threadIdx.x - o] : 0;
            __syncthreads();
            s_tot[threadIdx.x] += v;
            __syncthreads();
        }
        const int incl = s_tot[threadIdx.x];
        const int chunk = s_tot[blockDim.x - 1];
        __syncthreads();
        if (r < R) {
            const int ns = carry + incl - sz;
            shift[r] = ns - starts[r];
            old_end[r] = starts[r + 1];
            sizes_out[r] = sz;
        }
        carry += chunk;
        __syncthreads();
    }
    __syncthreads();
    // starts <- new starts (after every thread has read the old ones)
    for (int r = threadIdx.x; r < R; r += blockDim.x) starts[r] = starts[r] + shift[r];
    __syncthreads();
    if (threadIdx.x == 0) {
        starts[R] = carry;
        *d_n = carry;
    }
}

// Writes the rewritten mothers at their shifted rows and the daughters at
// the end of their replicate (old end + shift + i), and clears the dividing
// flags of the old rows.
__global__ void k_div_place(int R, int nd_all, int ncols, const int* __restrict__ sorted_rows,
                            const int* __restrict__ starts_new, const int* __restrict__ doff,
                            const int* __restrict__ nd, const int* __restrict__ shift, const int* __restrict__ old_end,
                            const double* __restrict__ mo, const double* __restrict__ da, int ldp,
                            double* __restrict__ dst, int ld, int* __restrict__ flags)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nd_all) return;
    const int row = sorted_rows[k];
    flags[row] = 0;
    // the replicate of the old row: shifted starts are new starts, so search
    // the old ends (ascending)
    int lo = 0, hi = R - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (old_end[mid] <= row) lo = mid + 1; else hi = mid;
    }
    const int r = lo;
    const int i = k - doff[r];
    if (i < 0 || i >= nd[r]) return;
    const int64_t mrow = row + shift[r], drow = old_end[r] + shift[r] + i;
    for (int col = 0; col < ncols; ++col) {
        dst[(int64_t)col * ld + mrow] = mo[(int64_t)col * ldp + k];
        dst[(int64_t)col * ld + drow] = da[(int64_t)col * ldp + k];
    }
    (void)starts_new;
}

// runSimulation for a batch of independent replicates in lockstep: every
// iteration advances each unfinished replicate by its own dt with its own
// controller, division stream and lineage. Replicates occupy contiguous row
// ranges (re-packed after divisions, which keeps each replicate's row order),
// the neighbor build keeps pairs inside a replicate, and one captured graph
// (stages + update for all replicates) serves every RK attempt.
// Extra inputs and outputs of a one-replicate run (bfm_simulate).
struct SingleRunIO {
    const int* parents_in = nullptr;
    int n_parents_in = 0;
    int* parents_out = nullptr;
    int* n_parents_out = nullptr;
    double* dt_trace = nullptr;     // accepted dt of replicate 0 per iteration
    int cap_trace = 0;
    const FrameHook* hook = nullptr;
};

void ensemble(bfm_ctx* c, const bfm_sim_config* cfg, int R, const int* sizes, const double* cells_in, int ncols,
              const uint64_t* seeds, int autotune, double* cells_out, int cap, int* sizes_out, double* t_out,
              int64_t* iters_out, int* failed_out, bfm_ensemble_stats* stats, const SingleRunIO* io = nullptr)
{
    check_config(cfg, ncols);
    if (!c->params_set) sim_fail(BFM_ERR_ARG, "bfm_set_params has not been called");
    if (c->p.ncols != ncols) sim_fail(BFM_ERR_ARG, "ncols differs from the parameters' ncols");
    if (R < 1 || !sizes || !cells_in || !seeds || !cells_out || !sizes_out)
        sim_fail(BFM_ERR_ARG, "need n_rep >= 1 and non-NULL arrays");
    int64_t total64 = 0;
    for (int r = 0; r < R; ++r) {
        if (sizes[r] < 1) sim_fail(BFM_ERR_ARG, "every replicate needs at least one cell");
        total64 += sizes[r];
    }
    if (total64 > cap || cap > c->max_cells) sim_fail(BFM_ERR_ARG, "need sum(sizes) <= cap <= max_cells");
    const auto wall0 = std::chrono::steady_clock::now();
    auto wall = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count(); };
    // tableau -1: velocity Verlet at the fixed dt = max_stepsize (simulation.hpp:1787-1798)
    const bool verlet = cfg->tableau == -1;
    if (verlet && (R != 1 || !(cfg->rho > 0)))
        sim_fail(BFM_ERR_ARG, "velocity Verlet runs one colony and needs rho > 0");
    const TableauDef tdef = tableau(verlet ? 0 : cfg->tableau);
    Tableau tb;
    std::memset(&tb, 0, sizeof(tb));
    tb.s = tdef.s;
    for (int i = 0; i < 49; ++i) tb.A[i] = tdef.A[i];
    for (int i = 0; i < 7; ++i) { tb.b[i] = tdef.b[i]; tb.bs[i] = tdef.bs[i]; }
    hipStream_t st = c->stream;
    const int ld = cap;
    int n = (int)total64;

    DevBuf<double> bufA((size_t)ld * ncols), bufB((size_t)ld * ncols), d_dt(R);
    DevBuf<int> d_n(1), d_active(R), d_flags(cap), d_idx(cap), d_count(1), d_starts(R + 1), d_shift(R);
    DevBuf<unsigned long long> d_errmax(R);
    DevBuf<int> d_bad(R);
    // the device controller's state and its host-mapped readback
    DevBuf<double> d_cdt(R), d_cfactor(R), d_ct(R), d_cerr(R);
    DevBuf<int> d_ctries(R), d_cstate(R), d_cstepped(R), d_cfailed(R), d_go(R);
    DevBuf<long long> d_citers(R);
    DevBuf<int> d_cpending(1);
    DevBuf<unsigned> d_tick(2);   // last-work-group tickets of the attempt and tail kernels
    MappedBuf pub(sizeof(CtrlPub));
    CtrlPub* h_pub = (CtrlPub*)pub.h;
    CtrlPub* d_pub = (CtrlPub*)pub.d;
    // the end-of-iteration values, one block per parity of the iteration
    // (IterSnap, then per replicate t, dt, err and flags)
    const size_t snap_stride = (sizeof(IterSnap) + (size_t)R * (3 * sizeof(double) + sizeof(int)) + 255) / 256 * 256;
    MappedBuf snapbuf(2 * snap_stride);
    IterSnap* snap_h[2];
    IterSnap* snap_d[2];
    double *t_h[2], *dt_h[2], *err_h[2];
    int* flags_h[2];
    double *t_d[2], *dt_d[2], *err_d[2];
    int* flags_d[2];
    for (int q = 0; q < 2; ++q) {
        snap_h[q] = (IterSnap*)((char*)snapbuf.h + q * snap_stride);
        snap_d[q] = (IterSnap*)((char*)snapbuf.d + q * snap_stride);
        t_h[q] = (double*)(snap_h[q] + 1);
        dt_h[q] = t_h[q] + R;
        err_h[q] = dt_h[q] + R;
        flags_h[q] = (int*)(err_h[q] + R);
        t_d[q] = (double*)(snap_d[q] + 1);
        dt_d[q] = t_d[q] + R;
        err_d[q] = dt_d[q] + R;
        flags_d[q] = (int*)(err_d[q] + R);
    }
    DevBuf<long long> d_totals(3);
    DevBuf<int> d_spec(1);
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } evs;
    for (int q = 0; q < 2; ++q) SCHK(hipEventCreateWithFlags(&evs.e[q], hipEventDisableTiming));
    hipEvent_t* ev_end = evs.e;
    const CtrlState cs{d_cdt.p, d_cfactor.p, d_ct.p, d_cerr.p, d_ctries.p, d_cstate.p, d_cstepped.p, d_cfailed.p,
                       d_citers.p, d_cpending.p, d_totals.p};
    PinnedBuf h_dt(sizeof(double) * R), h_act(sizeof(int) * R), h_bad(sizeof(int) * R);
    int* p_bad = (int*)h_bad.p;
    SCHK(hipMemsetAsync(d_bad.p, 0, sizeof(int) * R, st));
    double* p_dt = (double*)h_dt.p;
    int* p_act = (int*)h_act.p;
    double* d_cur = bufA.p;
    double* d_new = bufB.p;
    SCHK(hipMemsetAsync(d_flags.p, 0, sizeof(int) * cap, st));
    SCHK(hipMemcpy2DAsync(d_cur, sizeof(double) * ld, cells_in, sizeof(double) * n, sizeof(double) * n, ncols,
                          hipMemcpyHostToDevice, st));

    std::vector<int> start(R + 1, 0), nr(sizes, sizes + R);
    for (int r = 0; r < R; ++r) start[r + 1] = start[r] + nr[r];
    std::vector<double> t(R, 0.0), dt(R, cfg->max_stepsize);
    std::vector<int64_t> iters(R, 0);
    std::vector<char> done(R, 0);
    std::vector<Mt19937> rng;
    rng.reserve(R);
    std::vector<std::vector<int>> parents(R);
    for (int r = 0; r < R; ++r) {
        rng.emplace_back((uint32_t)seeds[r]);
        if (io && r == 0 && io->parents_in && io->n_parents_in > 0) {
            parents[r].assign(io->parents_in, io->parents_in + io->n_parents_in);
            continue;
        }
        int max_id = 0;
        for (int i = start[r]; i < start[r + 1]; ++i) max_id = std::max(max_id, (int)cells_in[(size_t)C_ID * n + i]);
        parents[r].assign(max_id + 1, -1);
    }
    abi(c, set_replicate_count(c, R));
    // division state on the device: each replicate's random stream, its next
    // cell id and the lineage log (replicate, id, parent)
    DevBuf<uint32_t> d_mt((size_t)R * kMtWords);
    DevBuf<long long> d_next_id(R);
    const int log_cap = 2 * cap + 16;
    DevBuf<int4> d_lineage(log_cap);
    DevBuf<int> d_log_n(1);
    int log_read = 0;
    {
        std::vector<uint32_t> mt((size_t)R * kMtWords);
        std::vector<long long> nid(R);
        for (int r = 0; r < R; ++r) {
            std::memcpy(&mt[(size_t)r * kMtWords], rng[r].mt, sizeof(uint32_t) * 624);
            mt[(size_t)r * kMtWords + 624] = (uint32_t)rng[r].idx;
            nid[r] = (long long)parents[r].size();
        }
        SCHK(hipMemcpyAsync(d_mt.p, mt.data(), sizeof(uint32_t) * mt.size(), hipMemcpyHostToDevice, st));
        SCHK(hipMemcpyAsync(d_next_id.p, nid.data(), sizeof(long long) * R, hipMemcpyHostToDevice, st));
        SCHK(hipMemsetAsync(d_log_n.p, 0, sizeof(int), st));
        SCHK(hipStreamSynchronize(st));
    }
    // appends the lineage the device logged since the last call to parents
    auto take_lineage = [&] {
        int nlog = 0;
        SCHK(hipMemcpyAsync(&nlog, d_log_n.p, sizeof(int), hipMemcpyDeviceToHost, st));
        SCHK(hipStreamSynchronize(st));
        if (nlog > log_cap) sim_fail(BFM_ERR_CAPACITY, "lineage log overflow");
        if (nlog <= log_read) return;
        std::vector<int4> ent(nlog - log_read);
        SCHK(hipMemcpy(ent.data(), d_lineage.p + log_read, sizeof(int4) * ent.size(), hipMemcpyDeviceToHost));
        for (const int4& e : ent) {
            std::vector<int>& pr = parents[e.x];
            if ((int)pr.size() <= e.y) pr.resize(e.y + 1, -1);
            pr[e.y] = e.z;
        }
        log_read = nlog;
    };
    {
        std::vector<long long> zl(R, 0);
        std::vector<int> zi(R, 0);
        std::vector<double> zd(R, 0.0);
        SCHK(hipMemcpyAsync(d_cdt.p, dt.data(), sizeof(double) * R, hipMemcpyHostToDevice, st));
        SCHK(hipMemcpyAsync(d_ct.p, zd.data(), sizeof(double) * R, hipMemcpyHostToDevice, st));
        SCHK(hipMemcpyAsync(d_cerr.p, zd.data(), sizeof(double) * R, hipMemcpyHostToDevice, st));
        SCHK(hipMemcpyAsync(d_cfailed.p, zi.data(), sizeof(int) * R, hipMemcpyHostToDevice, st));
        SCHK(hipMemcpyAsync(d_citers.p, zl.data(), sizeof(long long) * R, hipMemcpyHostToDevice, st));
        SCHK(hipMemsetAsync(d_cpending.p, 0, sizeof(int), st));
        SCHK(hipMemsetAsync(d_totals.p, 0, sizeof(long long) * 3, st));
        SCHK(hipMemsetAsync(d_spec.p, 0, sizeof(int), st));
        SCHK(hipMemsetAsync(d_tick.p, 0, sizeof(unsigned) * 2, st));
        // every attempt starts from the key of -inf (k_ctrl_after resets it)
        hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(R, 256)), dim3(256), 0, st, d_errmax.p, R, kKeyNegInf);
        SCHK(hipGetLastError());
        SCHK(hipStreamSynchronize(st));
    }
    auto upload_layout = [&] {
        SCHK(hipMemcpyAsync(d_starts.p, start.data(), sizeof(int) * (R + 1), hipMemcpyHostToDevice, st));
        SCHK(hipMemcpyAsync(d_n.p, &n, sizeof(int), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_fill_rep, dim3(R), dim3(256), 0, st, c->d_rep, d_starts.p);
        SCHK(hipGetLastError());
        SCHK(hipStreamSynchronize(st));
    };
    upload_layout();

    const double thr = cfg->neighbor_threshold, Rr = cfg->R, Ldiv = cfg->Ldiv;
    int64_t builds = 0, attempts = 0, divisions = 0, iter = 0, stage_updates = 0;
    double div_seconds = 0;
    bool list_fresh = false;   // the list's distances are those of d_cur
    auto build = [&] {
        int64_t m = 0;
        abi(c, bfm_dev_build_neighbors(c, d_cur, n, ld, thr, Rr, Ldiv, &m));
        ++builds;
        list_fresh = true;
    };
    build();
    double tuned_scale = c->bin_scale, tuned_ms = 0;
    int tuned_variant = c->cand_variant;
    if (autotune) {
        abi(c, bfm_autotune_neighbors(c, d_cur, n, ld, thr, Rr, Ldiv, &tuned_scale, &tuned_variant, &tuned_ms));
        ++builds;
    }

    // one captured RK attempt per buffer orientation
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    struct GraphGuard {
        hipGraphExec_t* e;
        hipGraph_t* g;
        ~GraphGuard()
        {
            for (int k = 0; k < 2; ++k) {
                if (e[k]) (void)hipGraphExecDestroy(e[k]);
                if (g[k]) (void)hipGraphDestroy(g[k]);
            }
        }
    } guard{exec, graph};
    // one-launch attempts (an isotropic interaction without friction) go
    // straight to the stream; the staged ones are captured once per buffer
    // orientation
    const bool direct = !verlet && attempt_fused(c);
    const bool was_timing = c->timing;
    c->timing = false;
    for (int k = 0; k < 2 && !verlet && !direct; ++k) {
        const double* src = k == 0 ? bufA.p : bufB.p;
        double* dst = k == 0 ? bufB.p : bufA.p;
        SCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = step_attempt_batched(c, tb, src, ld, d_n.p, c->d_rep, d_dt.p, d_active.p, R, dst,
                                      cfg->max_error_allowed, d_errmax.p, d_bad.p);
        hipError_t ec = hipStreamEndCapture(st, &graph[k]);
        abi(c, rc);
        SCHK(ec);
        SCHK(hipGraphInstantiate(&exec[k], graph[k], nullptr, nullptr, 0));
    }
    c->timing = was_timing;

    // device division buffers, grown to the largest batch of dividing cells
    struct DivBufs {
        int cap = 0;
        double *mo = nullptr, *da = nullptr, *q = nullptr;
        int4* meta = nullptr;
        unsigned long long* md = nullptr;
        int *sorted = nullptr, *check = nullptr, *sat = nullptr;
        void* temp = nullptr;
        size_t temp_bytes = 0;
        ~DivBufs() { release(); }
        void release()
        {
            for (void* p : {(void*)mo, (void*)da, (void*)q, (void*)meta, (void*)md, (void*)sorted, (void*)check,
                            (void*)sat, temp})
                if (p) (void)hipFree(p);
            mo = da = q = nullptr;
            meta = nullptr;
            md = nullptr;
            sorted = check = sat = nullptr;
            temp = nullptr;
        }
    } dv;
    DevBuf<int> d_doff(R), d_nd(R), d_ntries(R), d_jobact(R), d_old_end(R), d_divstat(1);
    MappedBuf divpub(sizeof(int) * (4 + (size_t)R));
    int* h_div = (int*)divpub.h;    // [0] jobs still placing, [1] status, [2..] new sizes
    int* d_divp = (int*)divpub.d;
    const int key_bits = std::max(1, (int)std::ceil(std::log2((double)cap + 1)));
    auto ensure_div = [&](int nd_all) {
        if (nd_all <= dv.cap) return;
        SCHK(hipStreamSynchronize(st));
        dv.release();
        const int k = std::max(nd_all, 64);
        SCHK(hipMalloc((void**)&dv.mo, sizeof(double) * (size_t)k * ncols));
        SCHK(hipMalloc((void**)&dv.da, sizeof(double) * (size_t)k * ncols));
        SCHK(hipMalloc((void**)&dv.q, sizeof(double) * 7 * 2 * (size_t)k));
        SCHK(hipMalloc((void**)&dv.meta, sizeof(int4) * 2 * (size_t)k));
        SCHK(hipMalloc((void**)&dv.md, sizeof(unsigned long long) * 2 * (size_t)k));
        SCHK(hipMalloc((void**)&dv.sorted, sizeof(int) * (size_t)k));
        SCHK(hipMalloc((void**)&dv.check, sizeof(int) * (size_t)k));
        SCHK(hipMalloc((void**)&dv.sat, sizeof(int) * (size_t)k));
        dv.temp_bytes = 0;
        SCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, dv.temp_bytes, d_idx.p, dv.sorted, k, 0, key_bits, st));
        SCHK(hipMalloc(&dv.temp, std::max<size_t>(dv.temp_bytes, 16)));
        dv.cap = k;
    };
    const DivParams dpar{cfg->R, 2 * cfg->Rcell, {cfg->growth_mean[0], cfg->growth_mean[1]},
                         {cfg->growth_std[0], cfg->growth_std[1]}, cfg->daughter_length_std,
                         cfg->daughter_angle_xy_bound, cfg->daughter_angle_z_bound, ncols};
    std::vector<char> act(R, 0);
    std::vector<char> failed(R, 0);
    int n_failed = 0;
    double last_err0 = 0;
    double last_frame_wall = 0;
    std::vector<double> frame_buf;
    const CtrlConsts kc{cfg->min_error, 1.0 / (tdef.order + 1), cfg->min_stepsize, cfg->max_stepsize,
                        cfg->max_tries_update_stepsize, R, tb.s, cfg->max_time};
    std::vector<int> go_host(R, -1);
    // the replicates that take this iteration (the host's termination rules)
    auto upload_go = [&] {
        bool changed = false;
        for (int r = 0; r < R; ++r)
            if (go_host[r] != (act[r] ? 1 : 0)) { go_host[r] = act[r] ? 1 : 0; changed = true; }
        if (changed) {
            SCHK(hipStreamSynchronize(st));   // p_act may still feed an earlier copy
            for (int r = 0; r < R; ++r) p_act[r] = go_host[r];
            SCHK(hipMemcpyAsync(d_go.p, p_act, sizeof(int) * R, hipMemcpyHostToDevice, st));
        }
    };
    // the per-replicate values k_ctrl_end published for the iteration just
    // finished (read after a synchronisation): time, step counts, dt trace,
    // failures (a non-finite result freezes its replicate at the state the
    // iteration started from; the reference stops such a run,
    // simulation.hpp:1803-1822; the others go on)
    unsigned iter_par = 0;   // parity of the iteration being processed (its IterSnap block)
    bool expect_retries = false;   // a single colony's next iterations are likely to retry
    auto take_iteration = [&] {
        const IterSnap* sn = snap_h[iter_par];
        const double* pt = t_h[iter_par];
        const double* pdt = dt_h[iter_par];
        const int* pflags = flags_h[iter_par];
        for (int r = 0; r < R; ++r) {
            const int fl = pflags[r];
            if (fl & 1) {
                if (io && io->dt_trace && r == 0 && iters[0] < io->cap_trace) io->dt_trace[iters[0]] = pdt[0];
                iters[r]++;
            }
            t[r] = pt[r];
            dt[r] = pdt[r];
            if ((fl & 2) && !failed[r]) {
                failed[r] = 1;
                done[r] = 1;
                ++n_failed;
            }
        }
        last_err0 = err_h[iter_par][0];
        stage_updates = sn->cells_total * tb.s;
        // below the largest step the controller sets dt from the error and
        // retries are frequent
        expect_retries = dt[0] < cfg->max_stepsize;
        attempts = sn->attempts;
    };
    int ndiv_next = -1;        // the next iteration's dividing cells (-1: not known)
    bool ahead_ran = false;    // the iteration to process was enqueued ahead and ran
    const bool pipelined = !verlet && direct && R == 1;
    int cur_control = 1;
    const int* cur_gate = nullptr;
    // the controller's step after an attempt (first_go: after the
    // iteration's first attempt, which set nothing up before it)
    auto after = [&](const int* first_go) {
        return CtrlAfter{cs, kc, d_errmax.p, d_bad.p, d_starts.p, d_dt.p, d_active.p, d_pub, first_go, cur_control,
                         d_count.p, cur_gate};
    };
    // one attempt after the first and the controller's step: a direct
    // launch (one kernel, the step in its last work-group for small
    // batches) or the captured graph of the staged attempt
    auto next_attempt = [&] {
        if (direct) {
            abi(c, attempt_direct(c, tb, d_cur, ld, d_n.p, n, c->d_rep, false, d_new, cfg->max_error_allowed,
                                  after(nullptr), d_tick.p));
        } else {
            SCHK(hipGraphLaunch(exec[d_cur == bufA.p ? 0 : 1], st));
            hipLaunchKernelGGL(k_ctrl_after, dim3(1), dim3(256), 0, st, after(nullptr));
        }
    };
    // the iteration's end (one kernel and, for large batches, the
    // publication): the controller's end-of-iteration step, the columns the
    // attempts leave alone (all columns of replicates that did not step),
    // growth, the next iteration's dividing cells, the new state's packed
    // segments (the next step_prepare brings the list's distances to them),
    // then the iteration's values to the host (IterSnap of parity q). It does
    // nothing while a replicate still retries; with the one-launch attempts
    // (which read only the step's frozen sums) it is enqueued before the host
    // knows that and enqueued again after the retries
    auto enqueue_end = [&](unsigned q) {
        const CtrlEnd ce{cs, kc, d_dt.p, d_active.p, t_d[q], dt_d[q], err_d[q], flags_d[q]};
        abi(c, iteration_tail_batched(c, d_cur, ld, d_n.p, n, c->d_rep, ce, Rr, Ldiv, d_new, d_count.p, d_idx.p,
                                      d_cpending.p, true, cur_gate, snap_d[q], d_spec.p, d_tick.p + 1));
        SCHK(hipEventRecord(ev_end[q], st));
    };
    // iteration `it` (parity q) from d_cur into d_new; gated: enqueued
    // before the host has read the previous iteration, its kernels gated on
    // the previous tail's d_spec
    auto enqueue_iteration = [&](int64_t it, unsigned q, bool gated) {
        cur_control = it % cfg->iter_update_stepsize == 0;
        cur_gate = gated ? d_spec.p : nullptr;
        bool cc_in_attempt = false;
        abi(c, step_prepare(c, d_cur, n, ld, !list_fresh, &cc_in_attempt, cur_gate));
        if (!gated) upload_go();
        if (direct) {   // the first attempt's controller step sets the iteration up
            abi(c, attempt_direct(c, tb, d_cur, ld, d_n.p, n, c->d_rep, cc_in_attempt, d_new,
                                  cfg->max_error_allowed, after(d_go.p), d_tick.p));
        } else {
            hipLaunchKernelGGL(k_ctrl_begin, dim3(1), dim3(256), 0, st, cs, kc, d_go.p, cur_control, d_starts.p,
                               d_dt.p, d_active.p, d_errmax.p, d_pub, d_count.p);
            next_attempt();
        }
        if (cur_control) next_attempt();
        // a batch of replicates nearly always has one still retrying after
        // two attempts, and a single colony whose step the controller holds
        // below the largest one retries often (an iteration enqueued ahead
        // would then wait for the host's retries): every attempt the controller may ask for
        // (max_tries retries and the final re-run) is enqueued at once; those
        // after the last replicate finished return at once
        if (direct && (R > 1 || (pipelined && expect_retries)) && cur_control)
            for (int k = 0; k < cfg->max_tries_update_stepsize; ++k) next_attempt();
        SCHK(hipGetLastError());
        if (direct) enqueue_end(q);   // ahead of the retry check
    };
    double next_report = cfg->report_seconds;
    std::vector<int> idx;
    for (;;) {
        // an iteration enqueued ahead that ran is processed first (the
        // termination rules are checked after it)
        if (!ahead_ran) {
        bool any = false;
        for (int r = 0; r < R; ++r) {
            if (!done[r] && terminate_now(cfg, nr[r], iter, t[r])) done[r] = 1;
            act[r] = !done[r];
            any |= act[r] != 0;
        }
        if (!any) break;
        if (cfg->max_wall_seconds > 0 && wall() >= cfg->max_wall_seconds) break;
        if (cfg->report_seconds > 0 && wall() >= next_report) {
            int n_act = 0;
            double tmin = 1e300;
            for (int r = 0; r < R; ++r)
                if (act[r]) { ++n_act; tmin = std::min(tmin, t[r]); }
            if (R == 1)
                std::fprintf(stderr, "[bfm_simulate] wall %.1f s  iter %lld  n %d  t %.6f h  dt %.3e h  rk attempts %lld\n",
                             wall(), (long long)iter, n, t[0], dt[0], (long long)attempts);
            else
                std::fprintf(stderr,
                             "[bfm_ensemble] wall %.1f s  iter %lld  cells %d  active %d  min t %.6f h  attempts %lld\n",
                             wall(), (long long)iter, n, n_act, tmin, (long long)attempts);
            std::fflush(stderr);
            next_report += cfg->report_seconds;
        }

        // divideMaxLength + divideCells per active replicate; the count of
        // the current cells comes with the previous iteration's readback
        if (ndiv_next < 0) {
            hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, d_count.p);
            hipLaunchKernelGGL(k_find_dividing, dim3(blocks_for(n, 256)), dim3(256), 0, st, d_cur, ld, n, Ldiv,
                               d_count.p, d_idx.p);
            hipLaunchKernelGGL(k_copy_int, dim3(1), dim3(1), 0, st, d_count.p, &d_pub->ndiv);
            SCHK(hipStreamSynchronize(st));
            ndiv_next = h_pub->ndiv;
        }
        const int nd_all = ndiv_next;
        ndiv_next = -1;
        if (nd_all > 0) {
            // divideCells of every replicate that steps, on the device
            auto td = std::chrono::steady_clock::now();
            upload_go();
            ensure_div(nd_all);
            size_t tb_bytes = dv.temp_bytes;
            SCHK(hipcub::DeviceRadixSort::SortKeys(dv.temp, tb_bytes, d_idx.p, dv.sorted, nd_all, 0, key_bits, st));
            hipLaunchKernelGGL(k_div_jobs, dim3(1), dim3(256), 0, st, R, d_starts.p, d_go.p, dv.sorted, nd_all,
                               d_doff.p, d_nd.p, d_ntries.p, d_jobact.p);
            const int nb = blocks_for(nd_all, 256);
            hipLaunchKernelGGL(k_div_start, dim3(nb), dim3(256), 0, st, R, nd_all, d_starts.p, dv.sorted, d_doff.p,
                               d_nd.p, d_cur, ld, d_flags.p, dv.check, dv.sat, dv.q, dv.meta);
            int span = 1;
            for (int r = 0; r < R; ++r) span = std::max(span, nr[r]);
            const int gx = std::min(blocks_for(span, 256), 64);
            hipLaunchKernelGGL(k_fill_u64, dim3(nb), dim3(256), 0, st, dv.md, nd_all,
                               (unsigned long long)0x7FF0000000000000ull);
            // minimum distances of nq queries, in launches of at most 65,535 (the grid's y limit)
            auto min_dist = [&](int nq) {
                for (int q0 = 0; q0 < nq; q0 += 65535)
                    hipLaunchKernelGGL(k_min_dist, dim3(gx, std::min(65535, nq - q0)), dim3(256), 0, st, d_cur, ld,
                                       dv.q + 7 * (size_t)q0, dv.meta + q0, d_flags.p, dpar.cut, dv.md + q0);
            };
            min_dist(nd_all);
            hipLaunchKernelGGL(k_div_noroom, dim3(nb), dim3(256), 0, st, nd_all, dv.md, dpar.cut, dv.check, dv.sat);
            SCHK(hipMemsetAsync(d_divstat.p, 0, sizeof(int), st));
            for (int round = 0; round < 10; ++round) {   // placement attempts until every job is done
                hipLaunchKernelGGL(k_div_propose, dim3(R), dim3(64), 0, st, R, dpar, 0.0, d_starts.p, dv.sorted,
                                   d_doff.p, d_nd.p, d_ntries.p, d_jobact.p, dv.check, d_ct.p, d_cur, ld, d_mt.p,
                                   dv.mo, dv.da, dv.cap, dv.q, dv.meta, d_divstat.p);
                hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(2 * nd_all, 256)), dim3(256), 0, st, dv.md,
                                   2 * nd_all, (unsigned long long)0x7FF0000000000000ull);
                min_dist(2 * nd_all);
                hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, d_divp);
                hipLaunchKernelGGL(k_div_eval, dim3(R), dim3(64), 0, st, R, dpar.cut, d_doff.p, d_nd.p, d_ntries.p,
                                   d_jobact.p, dv.check, dv.sat, dv.mo, dv.da, dv.cap, dv.md, d_divp);
                SCHK(hipGetLastError());
                SCHK(hipStreamSynchronize(st));
                if (h_div[0] == 0) break;
            }
            // dividing cells of the stepping replicates (the capacity check
            // comes before anything is written)
            std::vector<int> ndr(R, 0);
            SCHK(hipMemcpyAsync(ndr.data(), d_nd.p, sizeof(int) * R, hipMemcpyDeviceToHost, st));
            int dstat = 0;
            SCHK(hipMemcpyAsync(&dstat, d_divstat.p, sizeof(int), hipMemcpyDeviceToHost, st));
            SCHK(hipStreamSynchronize(st));
            if (dstat & 1) sim_fail(BFM_ERR_ARG, "cell group must be 1 or 2");
            int nd = 0;
            for (int r = 0; r < R; ++r) nd += ndr[r];
            if (nd > 0) {
                if (n + nd > cap) sim_fail(BFM_ERR_CAPACITY, "division would exceed the cell capacity");
                hipLaunchKernelGGL(k_div_layout, dim3(1), dim3(256), 0, st, R, d_doff.p, d_nd.p, dv.mo, dv.da,
                                   dv.cap, d_next_id.p, d_log_n.p, d_lineage.p, log_cap, d_starts.p, d_shift.p,
                                   d_old_end.p, d_n.p, d_divp + 2, d_divstat.p);
                hipLaunchKernelGGL(k_repack, dim3(std::min(blocks_for((int64_t)n * ncols, 256), 8192)), dim3(256), 0,
                                   st, d_cur, d_new, ld, ncols, n, c->d_rep, d_shift.p);
                hipLaunchKernelGGL(k_div_place, dim3(nb), dim3(256), 0, st, R, nd_all, ncols, dv.sorted, d_starts.p,
                                   d_doff.p, d_nd.p, d_shift.p, d_old_end.p, dv.mo, dv.da, dv.cap, d_new, ld,
                                   d_flags.p);
                hipLaunchKernelGGL(k_fill_rep, dim3(R), dim3(256), 0, st, c->d_rep, d_starts.p);
                SCHK(hipGetLastError());
                SCHK(hipStreamSynchronize(st));
                std::swap(d_cur, d_new);
                for (int r = 0; r < R; ++r) nr[r] = h_div[2 + r];
                for (int r = 0; r < R; ++r) start[r + 1] = start[r] + nr[r];
                n = start[R];
                divisions += nd;
                div_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - td).count();
                build();
            }
        }
        }   // !ahead_ran
        if (verlet) {
            // stepVelocityVerlet (simulation.hpp:1787-1798); it leaves the list's
            // distances at the new positions, refreshed again after growth below
            p_dt[0] = dt[0];
            p_act[0] = 1;
            SCHK(hipMemcpyAsync(d_dt.p, p_dt, sizeof(double), hipMemcpyHostToDevice, st));
            SCHK(hipMemcpyAsync(d_active.p, p_act, sizeof(int), hipMemcpyHostToDevice, st));
            abi(c, step_verlet(c, dt[0], cfg->rho, d_cur, n, ld, d_new, ld, d_bad.p));
            ++attempts;
            stage_updates += 2 * (int64_t)n;
            SCHK(hipMemcpyAsync(p_bad, d_bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
            SCHK(hipStreamSynchronize(st));
            if (p_bad[0]) {   // the run stops at the last finite state (simulation.hpp:1803-1822)
                failed[0] = 1;
                ++n_failed;
                break;
            }
            std::swap(d_cur, d_new);
            abi(c, grow_batched(c, d_cur, ld, d_n.p, Rr, c->d_rep, d_dt.p, d_active.p));
            abi(c, bfm_dev_update_neighbor_distances(c, d_cur, n, ld));
            abi(c, bfm_sync(c));
        } else {
        // adaptive RK step of every active replicate (simulation.hpp:1677-1769)
        // with the controller on the device: the first attempt, the re-run
        // or first retry (every stepping replicate takes a second attempt when
        // step control is on), then the iteration's end and one
        // synchronisation; more attempts only for replicates still retrying.
        // A single colony on the one-kernel path also enqueues the next
        // iteration before that synchronisation when the host knows it needs
        // no rebuild and no termination; its kernels are gated on the tail
        // (d_spec: no division, no flag, time left) and do nothing otherwise.
        const unsigned par = iter_par;
        if (!ahead_ran) enqueue_iteration(iter, par, false);
        bool ahead = false;
        if (pipelined && iteration_ahead_ok(c, n) && (iter + 1) % cfg->iter_update_neighbors != 0 &&
            !(cfg->max_iter > 0 && iter + 1 >= cfg->max_iter) && !(cfg->n_cells > 0 && n >= cfg->n_cells)) {
            std::swap(d_cur, d_new);   // the next iteration starts from this one's result
            list_fresh = false;
            enqueue_iteration(iter + 1, par ^ 1u, true);
            std::swap(d_cur, d_new);
            ahead = true;
        }
        if (direct) SCHK(hipEventSynchronize(ev_end[par]));
        else SCHK(hipStreamSynchronize(st));
        if (!direct || !snap_h[par]->tail_ran) {   // a replicate still retries: the iteration ahead does nothing
            ahead = false;
            cur_control = iter % cfg->iter_update_stepsize == 0;
            cur_gate = nullptr;
            SCHK(hipStreamSynchronize(st));
            while (h_pub->pending > 0) {
                next_attempt();
                SCHK(hipGetLastError());
                SCHK(hipStreamSynchronize(st));
            }
            enqueue_end(par);
            SCHK(hipStreamSynchronize(st));
        }
        const IterSnap* sn = snap_h[par];
        if (sn->status) abi(c, bfm_sync(c));   // reports and clears the flags
        ahead_ran = ahead && sn->spec;
        if (ahead && !ahead_ran) SCHK(hipStreamSynchronize(st));   // its kernels found the gate closed
        ndiv_next = sn->ndiv;
        std::swap(d_cur, d_new);
        list_fresh = false;
        }
        if (verlet) {
            if (io && io->dt_trace && act[0] && iters[0] < io->cap_trace) io->dt_trace[iters[0]] = dt[0];
            for (int r = 0; r < R; ++r)
                if (act[r]) { t[r] += dt[r]; iters[r]++; }
        } else {
            take_iteration();
            iter_par ^= 1u;
        }
        iter++;
        if (iter % cfg->iter_update_neighbors == 0) build();
        // a frame when t has just passed a multiple of dt_write (simulation.hpp:2023-2063)
        if (io && io->hook && io->hook->fn && io->hook->dt_write > 0 && act[0]) {
            const double dw = io->hook->dt_write;
            const double told = std::fmod(t[0] - dt[0] + 1e-12, dw), tnew = std::fmod(t[0] + 1e-12, dw);
            if (told > tnew) {
                frame_buf.resize((size_t)n * ncols);
                SCHK(hipMemcpy2DAsync(frame_buf.data(), sizeof(double) * n, d_cur, sizeof(double) * ld,
                                      sizeof(double) * n, ncols, hipMemcpyDeviceToHost, st));
                SCHK(hipStreamSynchronize(st));
                const double now = wall();
                take_lineage();
                io->hook->fn(frame_buf.data(), n, n, iters[0], t[0], dt[0], last_err0, now - last_frame_wall,
                             parents[0]);
                last_frame_wall = now;
            }
        }
    }
    // the context's list keeps distances at the final state, as after each
    // iteration of runSimulation
    if (!list_fresh && !verlet) abi(c, bfm_dev_update_neighbor_distances(c, d_cur, n, ld));
    abi(c, bfm_sync(c));
    take_lineage();
    SCHK(hipMemcpy2DAsync(cells_out, sizeof(double) * cap, d_cur, sizeof(double) * ld, sizeof(double) * n, ncols,
                          hipMemcpyDeviceToHost, st));
    SCHK(hipStreamSynchronize(st));
    for (int r = 0; r < R; ++r) {
        sizes_out[r] = nr[r];
        if (t_out) t_out[r] = t[r];
        if (iters_out) iters_out[r] = iters[r];
        if (failed_out) failed_out[r] = failed[r];
    }
    abi(c, set_replicate_count(c, 0));
    if (io && io->parents_out && io->n_parents_out) {
        std::copy(parents[0].begin(), parents[0].end(), io->parents_out);
        *io->n_parents_out = (int)parents[0].size();
    }
    if (stats) {
        stats->iterations = iter;
        stats->attempts = attempts;
        stats->divisions = divisions;
        stats->neighbor_builds = builds;
        stats->stage_cell_updates = stage_updates;
        stats->wall_seconds = wall();
        stats->host_division_seconds = div_seconds;
        stats->autotune_ms = tuned_ms;
        stats->tuned_bin_scale = tuned_scale;
        stats->tuned_variant = tuned_variant;
        stats->n_final = n;
        stats->failed = n_failed;
    }
}

// runSimulation for one colony: the batched loop with one replicate (one
// captured graph per RK attempt, forces computed once per iteration).
void simulate(bfm_ctx* c, const bfm_sim_config* cfg, const double* cells_in, int n_in, int ncols,
              const int* parents_in, int n_parents_in, double* cells_out, int cap, int* n_out, int* parents_out,
              int* n_parents_out, double* dt_trace, int cap_trace, bfm_sim_stats* stats,
              const FrameHook* hook = nullptr)
{
    if (!cells_in || !cells_out || !n_out || n_in < 1 || cap < n_in || cap > c->max_cells)
        sim_fail(BFM_ERR_ARG, "need 1 <= n_in <= cap <= max_cells and non-NULL arrays");
    SingleRunIO io;
    io.parents_in = parents_in;
    io.n_parents_in = n_parents_in;
    io.parents_out = parents_out;
    io.n_parents_out = n_parents_out;
    io.dt_trace = dt_trace;
    io.cap_trace = cap_trace;
    io.hook = hook;
    const uint64_t seed = cfg ? cfg->seed : 0;
    int n = 0, failed = 0;
    double t = 0;
    int64_t iters = 0;
    bfm_ensemble_stats es;
    std::memset(&es, 0, sizeof(es));
    ensemble(c, cfg, 1, &n_in, cells_in, ncols, &seed, 0, cells_out, cap, &n, &t, &iters, &failed, &es, &io);
    *n_out = n;
    if (stats) {
        stats->iterations = iters;
        stats->rk_steps = es.attempts;
        stats->divisions = es.divisions;
        stats->neighbor_builds = es.neighbor_builds;
        stats->t_end = t;
        stats->wall_seconds = es.wall_seconds;
        stats->host_division_seconds = es.host_division_seconds;
    }
    if (failed) sim_fail(BFM_ERR_NAN, "non-finite positions or orientations (the run stopped at the last finite state)");
}

// ---------------------------------------------------------------------------
// the 2-D tree's growth loop (runSimulationAdaptiveLagrangian of ABM-2d)
// ---------------------------------------------------------------------------
// 2-D cell columns (include/indices.hpp of ABM-2d).
enum Col2 { E2_ID = 0, E2_RX = 1, E2_RY = 2, E2_NX = 3, E2_NY = 4, E2_L = 9, E2_HALF_L = 10, E2_T0 = 11,
            E2_GROWTH = 12, E2_GROUP = 16 };

// rotate (utils.hpp:373-380 of ABM-2d): the 2x2 rotation by theta.
void rotate_2d(const double n[2], double theta, double out[2])
{
    const double s = sin(theta), c = cos(theta);
    out[0] = c * n[0] + (-s) * n[1];
    out[1] = s * n[0] + c * n[1];
}

Seg seg_2d(const double* cells, int ld, int i)
{
    Seg g;
    g.r[0] = cells[(size_t)E2_RX * ld + i]; g.r[1] = cells[(size_t)E2_RY * ld + i]; g.r[2] = 0.0;
    g.n[0] = cells[(size_t)E2_NX * ld + i]; g.n[1] = cells[(size_t)E2_NY * ld + i]; g.n[2] = 0.0;
    g.hl = cells[(size_t)E2_HALF_L * ld + i];
    return g;
}

// One division batch of the 2-D tree (divideMaxLength + divideCells,
// growth.hpp:67-71, 438-697 of ABM-2d) on host cells (column-major, leading
// dimension ld, capacity cap): the no-room test against the population, then
// placement tries (two re-orientation draws per cell with room except on the
// last try, the length ratios, the mothers' then the daughters' growth
// rates) until every cell with room has both daughters at least 2 Rcell from
// all other cells, or ten tries; ids and lineage last. minDistToCell (:82-117)
// is only compared with 2 Rcell, so pairs whose bounding distance exceeds
// that are skipped. Returns the number of tries (0: nothing divided).
int divide_2d(const bfm_sim_config* cfg, int ncols, Mt19937& rng, double t, double* cells, int ld, int& n, int cap,
              std::vector<int>& parents)
{
    std::vector<int> idx;
    for (int i = 0; i < n; ++i)
        if (cells[(size_t)E2_L * ld + i] > cfg->Ldiv) idx.push_back(i);
    const int nd = (int)idx.size();
    if (nd == 0) return 0;
    if (n + nd > cap) sim_fail(BFM_ERR_CAPACITY, "a division would exceed the capacity");
    const Draws dr{cfg};
    const double R = cfg->R, cut = 2 * cfg->Rcell;
    std::vector<char> dividing(n, 0);
    for (int i : idx) dividing[i] = 1;
    // minimum over the cells that do not divide (their rows stay as they are)
    auto min_to_rest = [&](const Seg& a, int self) {
        double best = INFINITY;
        for (int j = 0; j < n; ++j) {
            if (j == self || (self < 0 && dividing[j])) continue;
            const Seg b = seg_2d(cells, ld, j);
            if (beyond_cut(a, b, cut)) continue;
            best = std::min(best, seg_dist(a, b));
        }
        return best;
    };
    std::vector<int> check(nd, 1), satisfies(nd, 0);
    for (int k = 0; k < nd; ++k)
        if (min_to_rest(seg_2d(cells, ld, idx[k]), idx[k]) < cut) check[k] = 0, satisfies[k] = 1;
    // the dividing rows as they were, and this try's mothers and daughters
    std::vector<double> orig((size_t)nd * ncols), mo, da;
    for (int k = 0; k < nd; ++k)
        for (int col = 0; col < ncols; ++col) orig[(size_t)k * ncols + col] = cells[(size_t)col * ld + idx[k]];
    auto sat_all = [&] {
        int v = 0;
        for (int x : satisfies) v += x;
        return v == nd;
    };
    int ntries = 0;
    while (ntries == 0 || (ntries < 10 && !sat_all())) {
        mo = orig;
        da = orig;
        std::vector<double> th1(nd, 0.0), th2(nd, 0.0);
        for (int k = 0; k < nd; ++k)
            if (ntries < 10 - 1 && check[k] == 1) {
                th1[k] = dr.angle_xy(rng);
                th2[k] = dr.angle_xy(rng);
            }
        for (int k = 0; k < nd; ++k) {
            double* m = &mo[(size_t)k * ncols];
            double* d = &da[(size_t)k * ncols];
            const double u[2] = {m[E2_NX], m[E2_NY]};
            double v[2];
            rotate_2d(u, th1[k], v);
            m[E2_NX] = v[0]; m[E2_NY] = v[1];
            rotate_2d(u, th2[k], v);
            d[E2_NX] = v[0]; d[E2_NY] = v[1];
        }
        std::vector<double> M(nd);
        for (int k = 0; k < nd; ++k) M[k] = dr.daughter_length(rng);
        for (int k = 0; k < nd; ++k) {
            double* m = &mo[(size_t)k * ncols];
            double* d = &da[(size_t)k * ncols];
            const double Ld = m[E2_L] - 2 * R;
            const double L1 = M[k] * Ld, L2 = (1 - M[k]) * Ld;
            const double div = -m[E2_HALF_L] + L1 + R;
            const double delta1 = L1 / 2 + R, delta2 = L2 / 2 + R;
            m[E2_L] = L1; m[E2_HALF_L] = L1 / 2;
            d[E2_L] = L2; d[E2_HALF_L] = L2 / 2;
            m[E2_RX] = m[E2_RX] + (div - delta1) * m[E2_NX];
            d[E2_RX] = d[E2_RX] + (div + delta2) * d[E2_NX];
            m[E2_RY] = m[E2_RY] + (div - delta1) * m[E2_NY];
            d[E2_RY] = d[E2_RY] + (div + delta2) * d[E2_NY];
            m[E2_T0] = t;
            d[E2_T0] = t;
        }
        for (int k = 0; k < nd; ++k) mo[(size_t)k * ncols + E2_GROWTH] = dr.growth(rng, (int)mo[(size_t)k * ncols + E2_GROUP]);
        for (int k = 0; k < nd; ++k) da[(size_t)k * ncols + E2_GROWTH] = dr.growth(rng, (int)da[(size_t)k * ncols + E2_GROUP]);
        auto seg_row = [&](const double* row) {
            Seg g;
            g.r[0] = row[E2_RX]; g.r[1] = row[E2_RY]; g.r[2] = 0.0;
            g.n[0] = row[E2_NX]; g.n[1] = row[E2_NY]; g.n[2] = 0.0;
            g.hl = row[E2_HALF_L];
            return g;
        };
        // minimum over the whole new population: the unchanged cells, then the
        // 2 nd rewritten rows other than the query itself
        auto min_new = [&](int self_row) {
            const Seg a = seg_row(self_row < nd ? &mo[(size_t)self_row * ncols] : &da[(size_t)(self_row - nd) * ncols]);
            double best = min_to_rest(a, -1);
            for (int r = 0; r < 2 * nd; ++r) {
                if (r == self_row) continue;
                const Seg b = seg_row(r < nd ? &mo[(size_t)r * ncols] : &da[(size_t)(r - nd) * ncols]);
                if (beyond_cut(a, b, cut)) continue;
                best = std::min(best, seg_dist(a, b));
            }
            return best;
        };
        for (int k = 0; k < nd; ++k)
            if (check[k] == 1 && min_new(k) >= cut && min_new(nd + k) >= cut) satisfies[k] = 1;
        ntries++;
    }
    const int max_id = (int)parents.size() - 1;
    for (int k = 0; k < nd; ++k) {
        const int parent_id = (int)mo[(size_t)k * ncols + E2_ID];
        mo[(size_t)k * ncols + E2_ID] = max_id + (2 * k + 1);
        da[(size_t)k * ncols + E2_ID] = max_id + (2 * k + 2);
        parents.push_back(parent_id);
        parents.push_back(parent_id);
    }
    for (int k = 0; k < nd; ++k)
        for (int col = 0; col < ncols; ++col) {
            cells[(size_t)col * ld + idx[k]] = mo[(size_t)k * ncols + col];
            cells[(size_t)col * ld + n + k] = da[(size_t)k * ncols + col];
        }
    n += nd;
    return ntries;
}

// growCells of the 2-D tree (growth.hpp:51-57 of ABM-2d) for an accepted
// step that was not the controller's re-run, and divideMaxLength's count.
__global__ void __launch_bounds__(256)
k2d_grow_count(double* __restrict__ cells, int ld, int n, double R, double dt, double Ldiv, int grow,
               int* __restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int div = 0;
    if (i < n) {
        double l = cells[(size_t)E2_L * ld + i];
        if (grow) {
            const double g = cells[(size_t)E2_GROWTH * ld + i];
            l = l + (g * (4 * R / 3 + l)) * dt;
            cells[(size_t)E2_L * ld + i] = l;
            cells[(size_t)E2_HALF_L * ld + i] = l / 2;
        }
        div = l > Ldiv;
    }
    const unsigned long long ballot = __ballot(div);
    if ((threadIdx.x & 63) == 0 && ballot) atomicAdd(count, (int)__popcll(ballot));
}

__global__ void k_key_init(unsigned long long* key) { *key = 0x000FFFFFFFFFFFFFull; }   // ordered key of -inf

// runSimulationAdaptiveLagrangian's loop of the 2-D tree (simulation.hpp:
// 549-765 of ABM-2d) with the cells on the device (leading dimension cap).
// Each iteration enqueues its first RK attempt right after the previous
// iteration's work and synchronises once: a small kernel publishes the
// attempt's scaled error, the division count of the current cells (counted
// when they grew) and the status word into host-mapped memory. The rare
// iteration that divides discards that attempt: the population goes to the
// host, divides there with the reference's draws, comes back, and the list
// is rebuilt before the attempt is taken again. The controller's re-run also
// grows the cells and counts the next divisions.
void simulate2d(bfm_ctx* c, const bfm_sim_config* cfg, const int* amap, int ng, const double* cells_in, int n_in,
                const int* parents_in, int n_parents_in, double* cells_out, int cap, int* n_out, int* parents_out,
                int* n_parents_out, double* dt_trace, int cap_trace, bfm_sim_stats* stats)
{
    if (!c->p2d_set) sim_fail(BFM_ERR_ARG, "bfm2d_set_params has not been called");
    if (!cfg) sim_fail(BFM_ERR_ARG, "NULL config");
    const int ncols = c->p2d.ncols;
    if (!cells_in || !cells_out || !n_out || n_in < 1 || cap < n_in || cap > c->max_cells)
        sim_fail(BFM_ERR_ARG, "need 1 <= n_in <= cap <= max_cells and non-NULL arrays");
    if (cfg->iter_update_neighbors < 1 || cfg->iter_update_stepsize < 1 || cfg->max_tries_update_stepsize < 0)
        sim_fail(BFM_ERR_ARG, "iter_update_* must be >= 1 and max_tries_update_stepsize >= 0");
    if (!(cfg->max_stepsize > 0) || !(cfg->min_stepsize > 0) || cfg->min_stepsize > cfg->max_stepsize)
        sim_fail(BFM_ERR_ARG, "need 0 < min_stepsize <= max_stepsize");
    if (!(cfg->max_error_allowed > 0)) sim_fail(BFM_ERR_ARG, "max_error_allowed must be > 0");
    if (ng < 0 || ng > 64) sim_fail(BFM_ERR_ARG, "n_groups must be 0..64");
    const TableauDef tdef = tableau(cfg->tableau);
    Tableau tb;
    std::memset(&tb, 0, sizeof(tb));
    tb.s = tdef.s;
    for (int i = 0; i < 49; ++i) tb.A[i] = tdef.A[i];
    for (int i = 0; i < 7; ++i) { tb.b[i] = tdef.b[i]; tb.bs[i] = tdef.bs[i]; }
    const auto wall0 = std::chrono::steady_clock::now();
    auto wall = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count(); };
    hipStream_t st = c->stream;
    DevBuf<double> bufA((size_t)cap * ncols), bufB((size_t)cap * ncols), d_sums((size_t)cap * 8);
    DevBuf<unsigned long long> d_err(1);
    DevBuf<int> d_count(1);
    MappedBuf pub(64);
    volatile unsigned long long* h_pub = (volatile unsigned long long*)pub.h;
    PinnedBuf host(sizeof(double) * (size_t)cap * ncols);
    double* hc = (double*)host.p;
    SCHK(hipMemcpy2DAsync(bufA.p, sizeof(double) * cap, cells_in, sizeof(double) * n_in, sizeof(double) * n_in,
                          ncols, hipMemcpyHostToDevice, st));
    std::vector<int> parents;
    if (parents_in) {
        parents.assign(parents_in, parents_in + n_parents_in);
    } else {
        int max_id = 0;
        for (int i = 0; i < n_in; ++i) max_id = std::max(max_id, (int)cells_in[i]);
        parents.assign(max_id + 1, -1);
    }
    Mt19937 rng((uint32_t)cfg->seed);
    double* cur = bufA.p;
    double* nxt = bufB.p;
    int n = n_in;
    int64_t iter = 0, attempts = 0, divisions = 0, builds = 0;
    double t = 0, dt = cfg->max_stepsize, host_div_s = 0, next_report = cfg->report_seconds;
    const double thr = cfg->neighbor_threshold, R = cfg->R, Ldiv = cfg->Ldiv, tol = cfg->max_error_allowed;
    unsigned long long* d_pub = (unsigned long long*)pub.d;
    auto build = [&] {   // list, to_adhere and pair records (a synchronising build)
        int64_t m = 0;
        abi(c, bfm2d_dev_build_neighbors(c, cur, n, cap, thr, R, Ldiv, amap, ng, &m));
        abi(c, step2d_prepare(c, n, d_sums.p));
        ++builds;
    };
    auto enqueue_attempt = [&](double h) {
        ++attempts;
        abi(c, step2d_attempt(c, tb, h, cur, n, cap, d_sums.p, nxt, cap, tol, d_err.p, nullptr, R, Ldiv));
    };
    auto finish = [&] {   // outputs: the current cells, lineage and counters
        SCHK(hipMemcpy2DAsync(cells_out, sizeof(double) * cap, cur, sizeof(double) * cap, sizeof(double) * n, ncols,
                              hipMemcpyDeviceToHost, st));
        SCHK(hipStreamSynchronize(st));
        *n_out = n;
        if (parents_out) std::copy(parents.begin(), parents.end(), parents_out);
        if (n_parents_out) *n_parents_out = (int)parents.size();
        if (stats) {
            stats->iterations = iter;
            stats->rk_steps = attempts;
            stats->divisions = divisions;
            stats->neighbor_builds = builds;
            stats->t_end = t;
            stats->wall_seconds = wall();
            stats->host_division_seconds = host_div_s;
        }
    };
    // publishes and waits: returns the attempt's max(scaled error, min_error).
    // A non-finite stage velocity ends the run with the cells the attempt
    // started from in cells_out.
    auto sync_read = [&]() {
        abi(c, step2d_publish(c, d_err.p, d_count.p, d_pub));
        SCHK(hipStreamSynchronize(st));
        const int status = (int)h_pub[2];
        if (status) {
            const int zero = 0;
            SCHK(hipMemcpy(c->d_status, &zero, sizeof(int), hipMemcpyHostToDevice));
            finish();
            char msg[160];
            std::snprintf(msg, sizeof(msg), "non-finite stage velocity in the 2-D growth loop (iteration %lld, "
                          "t = %.9g h, n = %d; cells_out holds the attempt's starting state)", (long long)iter, t, n);
            if (status & ST_NAN) sim_fail(BFM_ERR_NAN, msg);
            sim_fail(BFM_ERR_HIP, "device status error in the 2-D growth loop");
        }
        return std::max(order_key_value(h_pub[0]), cfg->min_error);
    };
    hipLaunchKernelGGL(k_key_init, dim3(1), dim3(1), 0, st, d_err.p);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, d_count.p);
    hipLaunchKernelGGL(k2d_grow_count, dim3(blocks_for(n, 256)), dim3(256), 0, st, cur, cap, n, R, 0.0, Ldiv, 0,
                       d_count.p);
    SCHK(hipGetLastError());
    build();
    for (;;) {
        // the iteration's first attempt, taken before the division decision is known
        enqueue_attempt(dt);
        double max_error = sync_read();
        const int nd = (int)h_pub[1];
        if (terminate_now(cfg, n, iter, t)) break;
        if (cfg->max_wall_seconds > 0 && wall() > cfg->max_wall_seconds) break;
        if (cfg->report_seconds > 0 && wall() >= next_report) {
            std::fprintf(stderr, "[bfm2d] iter %lld  n %d  t %.6g h  dt %.3g h  wall %.1f s\n", (long long)iter, n, t,
                         dt, wall());
            next_report += cfg->report_seconds;
        }
        if (nd > 0) {   // divide, rebuild, take the first attempt again
            const auto h0 = std::chrono::steady_clock::now();
            SCHK(hipMemcpy2DAsync(hc, sizeof(double) * cap, cur, sizeof(double) * cap, sizeof(double) * n, ncols,
                                  hipMemcpyDeviceToHost, st));
            SCHK(hipStreamSynchronize(st));
            const int n0 = n;
            divide_2d(cfg, ncols, rng, t, hc, cap, n, cap, parents);
            divisions += n - n0;
            SCHK(hipMemcpy2DAsync(cur, sizeof(double) * cap, hc, sizeof(double) * cap, sizeof(double) * n, ncols,
                                  hipMemcpyHostToDevice, st));
            host_div_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - h0).count();
            build();
            enqueue_attempt(dt);
            max_error = sync_read();
        }
        if (iter % cfg->iter_update_stepsize == 0) {   // simulation.hpp:639-712 of ABM-2d
            const double inv = 1.0 / (tdef.order + 1);
            bool exceeded = max_error > 1.0;
            double factor = 0.9 * pow(1.0 / max_error, inv);
            if (factor >= 10) factor = 10;
            else if (factor < 0.2) factor = 0.2;
            int j = 0;
            while (exceeded && j < cfg->max_tries_update_stepsize) {
                enqueue_attempt(dt * factor);
                max_error = sync_read();
                exceeded = max_error > 1.0;
                factor *= 0.9 * pow(1.0 / max_error, inv);
                if (factor >= 10) { factor = 10; break; }
                else if (factor < 0.2) { factor = 0.2; break; }
                j++;
            }
            if (dt * factor < cfg->min_stepsize) factor = cfg->min_stepsize / dt;
            else if (dt * factor > cfg->max_stepsize) factor = cfg->max_stepsize / dt;
            dt *= factor;
            // the re-run, with growCells and the next division count fused in
            ++attempts;
            abi(c, step2d_attempt(c, tb, dt, cur, n, cap, d_sums.p, nxt, cap, 0.0, d_err.p, d_count.p, R, Ldiv));
            std::swap(cur, nxt);
        } else {
            std::swap(cur, nxt);
            hipLaunchKernelGGL(k2d_grow_count, dim3(blocks_for(n, 256)), dim3(256), 0, st, cur, cap, n, R, dt, Ldiv,
                               1, d_count.p);
            SCHK(hipGetLastError());
        }
        abi(c, step2d_rows_update(c, cur, n, cap, d_sums.p));
        if (dt_trace && iter < cap_trace) dt_trace[iter] = dt;
        t += dt;
        iter++;
        if (iter % cfg->iter_update_neighbors == 0) build();
    }
    finish();
}

}  // namespace

namespace bfm {

int simulate_with_frames(bfm_ctx* c, const bfm_sim_config* cfg, const double* cells_in, int n_in, int ncols,
                         const int* parents_in, int n_parents_in, double* cells_out, int cap, int* n_out,
                         int* parents_out, int* n_parents_out, bfm_sim_stats* stats, const FrameHook* hook)
{
    if (!c) return BFM_ERR_ARG;
    try {
        SCHK(hipSetDevice(c->device));
        simulate(c, cfg, cells_in, n_in, ncols, parents_in, n_parents_in, cells_out, cap, n_out, parents_out,
                 n_parents_out, nullptr, 0, stats, hook);
        return BFM_OK;
    } catch (const SimError& e) {
        c->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return BFM_ERR_ARG;
    }
}

}  // namespace bfm

extern "C" {

int bfm_ensemble_simulate(bfm_ctx* c, const bfm_sim_config* cfg, int n_rep, const int* sizes,
                          const double* cells_in, int ncols, const uint64_t* seeds, int autotune,
                          double* cells_out, int cap, int* sizes_out, double* t_out, int64_t* iters_out,
                          int* failed_out, bfm_ensemble_stats* stats)
{
    if (!c) return BFM_ERR_ARG;
    try {
        SCHK(hipSetDevice(c->device));
        ensemble(c, cfg, n_rep, sizes, cells_in, ncols, seeds, autotune, cells_out, cap, sizes_out, t_out,
                 iters_out, failed_out, stats);
        return BFM_OK;
    } catch (const SimError& e) {
        (void)set_replicate_count(c, 0);
        c->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        (void)set_replicate_count(c, 0);
        c->err = e.what();
        return BFM_ERR_ARG;
    }
}

int bfm_simulate(bfm_ctx* c, const bfm_sim_config* cfg, const double* cells_in, int n_in, int ncols,
                 const int* parents_in, int n_parents_in, double* cells_out, int cap, int* n_out, int* parents_out,
                 int* n_parents_out, double* dt_trace, int cap_trace, bfm_sim_stats* stats)
{
    if (!c) return BFM_ERR_ARG;
    try {
        SCHK(hipSetDevice(c->device));
        simulate(c, cfg, cells_in, n_in, ncols, parents_in, n_parents_in, cells_out, cap, n_out, parents_out,
                 n_parents_out, dt_trace, cap_trace, stats);
        return BFM_OK;
    } catch (const SimError& e) {
        c->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return BFM_ERR_ARG;
    }
}

int bfm_divide_cells(const bfm_sim_config* cfg, double* cells, int* n_io, int cap, int ncols, double t,
                     uint64_t seed, int* parents, int* n_parents, int parents_cap, int* tries_out)
{
    try {
        if (!cfg || !cells || !n_io || !parents || !n_parents) return BFM_ERR_ARG;
        const int n = *n_io;
        if (n < 0 || n > cap || ncols < 21 || ncols > 32) return BFM_ERR_ARG;
        std::vector<int> idx;
        for (int i = 0; i < n; ++i)
            if (cells[(size_t)C_L * cap + i] > cfg->Ldiv) idx.push_back(i);
        const int nd = (int)idx.size();
        if (n + nd > cap || *n_parents + 2 * nd > parents_cap) return BFM_ERR_CAPACITY;
        if (tries_out) *tries_out = 0;
        if (nd == 0) return BFM_OK;
        std::vector<double> mrows((size_t)nd * ncols);
        for (int r = 0; r < nd; ++r)
            for (int k = 0; k < ncols; ++k) mrows[(size_t)r * ncols + k] = cells[(size_t)k * cap + idx[r]];
        std::vector<char> dividing(n, 0);
        for (int i : idx) dividing[i] = 1;
        BaseMinFn base_min = [&](const std::vector<Query>& q, double cut, std::vector<double>& out) {
            out.assign(q.size(), INFINITY);
            for (size_t k = 0; k < q.size(); ++k) {
                Seg a;
                std::memcpy(a.r, q[k].r, sizeof(a.r));
                std::memcpy(a.n, q[k].n, sizeof(a.n));
                a.hl = q[k].hl;
                for (int j = q[k].lo; j < q[k].hi; ++j) {
                    if (j == q[k].self || (q[k].skip_dividing && dividing[j])) continue;
                    Seg b;
                    b.r[0] = cells[(size_t)C_RX * cap + j]; b.r[1] = cells[(size_t)C_RY * cap + j];
                    b.r[2] = cells[(size_t)C_RZ * cap + j];
                    b.n[0] = cells[(size_t)C_NX * cap + j]; b.n[1] = cells[(size_t)C_NY * cap + j];
                    b.n[2] = cells[(size_t)C_NZ * cap + j];
                    b.hl = cells[(size_t)C_HALF_L * cap + j];
                    if (beyond_cut(a, b, cut)) continue;
                    double dist = seg_dist(a, b);
                    if (dist < out[k]) out[k] = dist;
                }
            }
        };
        std::vector<int> par(parents, parents + *n_parents);
        Mt19937 rng((uint32_t)seed);
        std::vector<DivisionJob> jobs(1);
        DivisionJob& jb = jobs[0];
        jb.rng = &rng;
        jb.parents = &par;
        jb.t = t;
        jb.lo = 0;
        jb.hi = n;
        jb.idx = idx;
        jb.mrows = std::move(mrows);
        divide_batch(cfg, ncols, jobs, base_min);
        for (int r = 0; r < nd; ++r)
            for (int k = 0; k < ncols; ++k) {
                cells[(size_t)k * cap + idx[r]] = jb.mo[(size_t)r * ncols + k];
                cells[(size_t)k * cap + n + r] = jb.da[(size_t)r * ncols + k];
            }
        *n_io = n + nd;
        std::copy(par.begin(), par.end(), parents);
        *n_parents = (int)par.size();
        if (tries_out) *tries_out = jb.ntries;
        return BFM_OK;
    } catch (const SimError& e) {
        return e.code;
    } catch (const std::exception&) {
        return BFM_ERR_ARG;
    }
}

int bfm_rng_uniform(uint64_t seed, int count, double* out)
{
    if (count < 0 || (count > 0 && !out)) return BFM_ERR_ARG;
    Mt19937 g((uint32_t)seed);
    for (int i = 0; i < count; ++i) out[i] = uniform01(g);
    return BFM_OK;
}

namespace {
// count uniform_01 draws, then count standard normals, of one stream on
// the device generator (one wavefront)
__global__ void __launch_bounds__(64)
k_rng_draws(uint32_t* __restrict__ state, int count, double* __restrict__ out)
{
    __shared__ uint32_t s_mt[624];
    const int lane = threadIdx.x;
    for (int i = lane; i < 624; i += 64) s_mt[i] = state[i];
    WaveMt g{s_mt, (int)state[624]};
    wave_sync_sim();
    for (int k = 0; k < count; ++k) {
        const double u = g.uniform01(lane);
        if (lane == 0) out[k] = u;
    }
    for (int k = 0; k < count; ++k) {
        const double z = g.normal(lane);
        if (lane == 0) out[count + k] = z;
    }
}
}   // namespace

int bfm_rng_draws_device(int device, uint64_t seed, int count, double* out)
{
    if (count < 0 || (count > 0 && !out)) return BFM_ERR_ARG;
    if (count == 0) return BFM_OK;
    try {
        SCHK(hipSetDevice(device));
        Mt19937 g((uint32_t)seed);
        std::vector<uint32_t> st(kMtWords);
        std::memcpy(st.data(), g.mt, sizeof(uint32_t) * 624);
        st[624] = (uint32_t)g.idx;
        DevBuf<uint32_t> d_st(kMtWords);
        DevBuf<double> d_out(2 * (size_t)count);
        SCHK(hipMemcpy(d_st.p, st.data(), sizeof(uint32_t) * kMtWords, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_rng_draws, dim3(1), dim3(64), 0, 0, d_st.p, count, d_out.p);
        SCHK(hipGetLastError());
        SCHK(hipMemcpy(out, d_out.p, sizeof(double) * 2 * count, hipMemcpyDeviceToHost));
    } catch (const SimError& e) {
        return e.code;
    }
    return BFM_OK;
}

int bfm2d_simulate(bfm_ctx* c, const bfm_sim_config* cfg, const int* adhesion_map, int n_groups,
                   const double* cells_in, int n_in, const int* parents_in, int n_parents_in, double* cells_out,
                   int cap, int* n_out, int* parents_out, int* n_parents_out, double* dt_trace, int cap_trace,
                   bfm_sim_stats* stats)
{
    if (!c) return BFM_ERR_ARG;
    try {
        SCHK(hipSetDevice(c->device));
        simulate2d(c, cfg, adhesion_map, n_groups, cells_in, n_in, parents_in, n_parents_in, cells_out, cap, n_out,
                   parents_out, n_parents_out, dt_trace, cap_trace, stats);
        return BFM_OK;
    } catch (const SimError& e) {
        c->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return BFM_ERR_ARG;
    }
}

int bfm2d_divide_cells(const bfm_sim_config* cfg, double* cells, int* n_io, int cap, int ncols, double t,
                       uint64_t seed, int* parents, int* n_parents, int parents_cap, int* tries_out)
{
    try {
        if (!cfg || !cells || !n_io || !parents || !n_parents || ncols < 17 || *n_io < 0 || *n_io > cap)
            return BFM_ERR_ARG;
        int nd = 0;
        for (int i = 0; i < *n_io; ++i) nd += cells[(size_t)E2_L * cap + i] > cfg->Ldiv;
        if (*n_io + nd > cap || *n_parents + 2 * nd > parents_cap) return BFM_ERR_CAPACITY;
        Mt19937 rng((uint32_t)seed);
        std::vector<int> par(parents, parents + *n_parents);
        const int tries = divide_2d(cfg, ncols, rng, t, cells, cap, *n_io, cap, par);
        std::copy(par.begin(), par.end(), parents);
        *n_parents = (int)par.size();
        if (tries_out) *tries_out = tries;
        return BFM_OK;
    } catch (const SimError& e) {
        return e.code;
    } catch (const std::exception&) {
        return BFM_ERR_ARG;
    }
}

}  // extern "C"

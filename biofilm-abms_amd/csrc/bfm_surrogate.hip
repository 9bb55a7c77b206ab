// bfm_surrogate.hip -- learned JKR contact radius (BF16 MFMA MLP + FP64 polish).
//
// The exact JKR radius is the larger root of the quartic of jkrContactRadius
// (jkr.hpp:67-144). With u = sqrt(a), p = Req sqrt(4 pi gamma / E),
// q = Req delta it is the largest root of u^4 - p u - q, and with
// t = u / p^(1/3), xi = q / p^(4/3) that is t^4 - t = xi: a dimensionless
// function of the inputs. The network maps the configured inputs
// (normalised overlap delta / R, both cell radii, surface energy) through
// 4 -> 64 -> 64 -> 64 -> 1 (SiLU) to t. Weights and activations are BF16, the
// three 64-wide layers run on MFMA (v_mfma_f32_16x16x16_bf16 for the first,
// gfx950's v_mfma_f32_16x16x32_bf16 for the two 64 x 64 layers) with FP32
// accumulation; the last layer is an FP32 dot product. The guess
// u = t p^(1/3) is then polished by FP64 Newton steps on u^4 - p u - q.
// Pairs outside the training envelope use the exact FP64 solve.
//
// Training runs on the GPU with this file's kernels: the same MFMA forward
// (storing pre-activations), FP32 backward kernels for the rows and the
// weight gradients, and Adam on FP32 master weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "bfm_device.hpp"
#include "bfm_surrogate.hpp"

using namespace bfm;

namespace bfm {

constexpr int kH = 64;         // hidden width
constexpr int kIn = 4;         // inputs
// parameter layout (FP32 master / gradients / Adam moments), offsets in floats
constexpr int kW1 = 0;                    // [64][4]
constexpr int kB1 = kW1 + kH * kIn;       // [64]
constexpr int kW2 = kB1 + kH;             // [64][64]
constexpr int kB2 = kW2 + kH * kH;
constexpr int kW3 = kB2 + kH;             // [64][64]
constexpr int kB3 = kW3 + kH * kH;
constexpr int kW4 = kB3 + kH;             // [64]
constexpr int kB4 = kW4 + kH;             // [1]
constexpr int kNParams = kB4 + 1;

// Feature normalisation of the training envelope (overlap in [-0.2 R, 0.3 R],
// gamma in [10, 1000] Pa um, R1 = R2 = 0.8 um).
using Envelope = SurrogateEnvelope;

struct SurrogateWeights {
    // device: BF16 weights (row-major [out][in]; W1 padded to [64][16]) and FP32 biases
    uint16_t* w1 = nullptr;   // 64 x 16
    uint16_t* w2 = nullptr;   // 64 x 64
    uint16_t* w3 = nullptr;   // 64 x 64
    uint16_t* w4 = nullptr;   // 64
    float* b = nullptr;       // b1(64) b2(64) b3(64) b4(1)
    double* radius = nullptr; // per-pair output, max_pairs
    int* st_row = nullptr;    // staged rows inside the envelope (max_pairs) ...
    double* st_delta = nullptr;
    double* st_gamma = nullptr;
    int* st_count = nullptr;  // ... and each staging block's count
    int64_t cap = 0;
    Envelope env{};
    int polish = 1;
};

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b)
{
    uint32_t u = (uint32_t)b << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}
__device__ inline float bfround(float x) { return bf2f(f2bf(x)); }

__device__ inline float sigmoidf_(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigmoidf_(z); }
__device__ inline float dsilu(float z)
{
    const float s = sigmoidf_(z);
    return s * (1.0f + z * (1.0f - s));
}

// Inputs of the network (all four configured inputs), scaled to about [-1, 1].
__host__ __device__ inline void features(double delta, double r1, double r2, double gamma,
                                         const Envelope& e, float x[4])
{
    const double dn = delta / e.R;
    x[0] = (float)((2 * dn - (e.dhi + e.dlo)) / (e.dhi - e.dlo));
    x[1] = (float)(r1 / e.R - 1.0);
    x[2] = (float)(r2 / e.R - 1.0);
    const double lg = log(gamma), l0 = log(e.glo), l1 = log(e.ghi);
    x[3] = (float)((2 * lg - (l1 + l0)) / (l1 - l0));
}

// B-operand fragment of W (row-major [64][K]) for k-step s and n-tile t:
// lane l holds W[16 t + l % 16][16 s + 4 (l / 16) + i], i = 0..3.
__device__ inline bf16x4 load_bfrag(const uint16_t* __restrict__ W, int K, int s, int t, int lane)
{
    // four consecutive BF16 (K is a multiple of 4, so the 8-byte load is aligned)
    const uint16_t* p = W + (16 * t + (lane & 15)) * K + 16 * s + 4 * (lane >> 4);
    return *reinterpret_cast<const bf16x4*>(p);
}

struct WFrags {
    bf16x4 w1[4];
    bf16x4 w2[4][4];
    bf16x4 w3[4][4];
};

__device__ inline void load_frags(const SurrogateWeights& sw_unused, const uint16_t* w1, const uint16_t* w2,
                                  const uint16_t* w3, int lane, WFrags& F)
{
    (void)sw_unused;
#pragma unroll
    for (int t = 0; t < 4; ++t) F.w1[t] = load_bfrag(w1, 16, 0, t, lane);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            F.w2[s][t] = load_bfrag(w2, 64, s, t, lane);
            F.w3[s][t] = load_bfrag(w3, 64, s, t, lane);
        }
}

// Forward of one 16-sample tile on one wavefront. feat: LDS [16][4] floats
// (sample-major); hbuf: LDS scratch [16][68] uint16 of this wavefront.
// Returns, in lanes with lane % 16 == 0, y of samples 4 (lane/16) + i in y4[i].
// When zs is given (training), it receives the FP32 pre-activations Z1..Z3
// and the BF16-rounded activations H1..H3 of the tile, [6][16][64].
__device__ inline void mlp_tile(const WFrags& F, const float* __restrict__ b, const uint16_t* __restrict__ w4,
                                const float* feat, uint16_t* hbuf, int lane, float y4[4], float* zs)
{
    // layer 1: A = features (K padded to 16), lanes 0..15 hold sample lane, k = 0..3
    bf16x4 a;
    if (lane < 16) {
        a[0] = (short)f2bf(feat[lane * 4 + 0]); a[1] = (short)f2bf(feat[lane * 4 + 1]);
        a[2] = (short)f2bf(feat[lane * 4 + 2]); a[3] = (short)f2bf(feat[lane * 4 + 3]);
    } else {
        a[0] = a[1] = a[2] = a[3] = 0;
    }
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 z = {0.f, 0.f, 0.f, 0.f};
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, F.w1[t], z, 0, 0, 0);
    }
    const int col = lane & 15, rowb = 4 * (lane >> 4);
    for (int layer = 0; layer < 3; ++layer) {
        const float* bl = b + 64 * layer;
        // bias + SiLU, BF16 activations to LDS in [m][n] order
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = rowb + i, n = 16 * t + col;
                const float z = acc[t][i] + bl[n];
                const uint16_t h = f2bf(silu(z));
                hbuf[m * 68 + n] = h;
                if (zs) {
                    zs[(2 * layer) * 1024 + m * 64 + n] = z;
                    zs[(2 * layer + 1) * 1024 + m * 64 + n] = bf2f(h);
                }
            }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (layer == 2) break;
        // A fragments: lane l holds H[m = l % 16][k = 16 s + 4 (l / 16) + i]
        bf16x4 af[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint16_t* p = hbuf + (lane & 15) * 68 + 16 * s + 4 * (lane >> 4);
            af[s][0] = (short)p[0]; af[s][1] = (short)p[1]; af[s][2] = (short)p[2]; af[s][3] = (short)p[3];
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 4; ++s)
                z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[s], layer == 0 ? F.w2[s][t] : F.w3[s][t], z, 0,
                                                              0, 0);
            acc[t] = z;
        }
    }
    // layer 4: y[m] = b4 + sum_n W4[n] H3[m][n] (FP32), reduced over the 16 lanes of a row group
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const float w = bf2f(w4[16 * t + col]);
#pragma unroll
        for (int i = 0; i < 4; ++i) part[i] += w * bf2f(hbuf[(rowb + i) * 68 + 16 * t + col]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float v = part[i];
        v += __shfl_xor(v, 8, 16);
        v += __shfl_xor(v, 4, 16);
        v += __shfl_xor(v, 2, 16);
        v += __shfl_xor(v, 1, 16);
        y4[i] = v + b[192];
    }
    __builtin_amdgcn_wave_barrier();
}

// The features in FP32 for inference: they are rounded to BF16 before the
// first layer, so the FP32 logarithm and quotients change nothing there.
__device__ inline void features_f32(double delta, double r, double gamma, const Envelope& e, float x[4])
{
    const float dn = (float)(delta / e.R);
    x[0] = (2.0f * dn - (float)(e.dhi + e.dlo)) / (float)(e.dhi - e.dlo);
    x[1] = x[2] = (float)(r / e.R) - 1.0f;
    const float l0 = logf((float)e.glo), l1 = logf((float)e.ghi);
    x[3] = (2.0f * logf((float)gamma) - (l1 + l0)) / (l1 - l0);
}

// Inference form of polish_radius: the starting guess u = t p^(1/3) carries
// the network's error (about 1e-3), so p^(1/3) is taken in FP32; p, q and the
// Newton steps stay FP64.
__device__ inline double polish_radius_fast(float t, double delta, double Req, double E, double gamma, int polish)
{
    const double p = Req * sqrt(4 * kPi * gamma / E);
    const double q = Req * delta;
    double u = (double)(t * cbrtf((float)p));
    for (int it = 0; it < polish; ++it) {
        const double u2 = u * u;
        const double F = (u2 * u2 - p * u) - q;
        const double dF = (4.0 * u2) * u - p;
        if (!(dF > 0)) break;
        u = u - F / dF;
    }
    return u * u;
}

// Radius from the network output t: u = t p^(1/3), then `polish` Newton steps.
__device__ inline double polish_radius(double t, double delta, double Req, double E, double gamma, int polish)
{
    const double p = Req * sqrt(4 * kPi * gamma / E);
    const double q = Req * delta;
    double u = t * cbrt(p);
    for (int it = 0; it < polish; ++it) {
        const double u2 = u * u;
        const double F = (u2 * u2 - p * u) - q;
        const double dF = (4.0 * u2) * u - p;
        if (!(dF > 0)) break;
        u = u - F / dF;
    }
    return u * u;
}

struct RadiusArgs {
    double R, Rcell, E0, Req, max_overlap;
    int colidx_gamma;
    int polish;
};

// Inference forward of a wavefront's 64 samples (four 16-sample tiles) in the
// transposed form D = W H^T: neurons are the MFMA M index and samples the N
// index. A layer's output fragment (lane l: neurons 16 t + 4 (l / 16) + i of
// sample l % 16, i = 0..3) is then, after bias and SiLU, exactly the B
// fragment of k-step t of the next layer, so activations stay in registers
// (no LDS round trip between layers), and the weight A fragments are the
// WFrags of the training forward. x: the four features of sample `lane`
// (zero where unused); returns y of sample `lane`.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));

// Two FP32 values to packed BF16 (round to nearest even, one v_cvt_pk_bf16_f32).
__device__ inline uint32_t pack_bf16x2(float a, float b)
{
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2v));
}

// SiLU with the hardware exp2 and reciprocal (1 ulp) instead of an IEEE divide.
__device__ inline float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

typedef short short8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));

// One K = 32 step of gfx950's v_mfma_f32_16x16x32_bf16 from two K = 16
// fragments. Its operand lane l holds K positions 8 (l / 16) .. + 7; the two
// 16-wide fragments give lane l the K indices 4 (l / 16) + i of k-step 2 s'
// and of 2 s' + 1, so position 8 (l / 16) + j stands for input
// 32 s' + (j < 4 ? 4 (l / 16) + j : 16 + 4 (l / 16) + j - 4). The weight and
// activation operands use the same order, so the product is the sum of the
// two K = 16 products.
__device__ inline f32x4 mfma_k32(bf16x4 a0, bf16x4 a1, bf16x4 b0, bf16x4 b1, f32x4 c)
{
    const short8 a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
    const short8 b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8v, a), __builtin_bit_cast(bf16x8v, b),
                                                   c, 0, 0, 0);
}

// Forward of up to 64 samples (ntiles 16-sample tiles); x: the four features
// of sample `lane`. Returns y of sample `lane` (0 for lanes past the tiles).
__device__ inline float mlp_forward64(const bf16x4 (*wl)[64], const float* bs, const float x[4], int lane,
                                      int ntiles = 4)
{
    const int col = lane & 15, g = lane >> 4;
    float y_mine = 0.f;
#pragma unroll 1
    for (int T = 0; T < ntiles; ++T) {
        // the weight fragments are re-read from LDS for every tile rather than
        // held in 72 registers across the loop
        asm volatile("" ::: "memory");
        // layer 1: B = X^T (K = 4 features padded to 16; lanes 0..15 hold sample 16 T + lane)
        bf16x4 xb;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float v = __shfl(x[i], 16 * T + col);
            xb[i] = lane < 16 ? (short)f2bf(v) : (short)0;
        }
        f32x4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wl[t][lane], xb, *(const f32x4*)(bs + 16 * t + 4 * g),
                                                              0, 0, 0);
        bf16x4 h[4];
        u32x2 hw[4];
#pragma unroll
        for (int layer = 0; layer < 3; ++layer) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                hw[t][0] = pack_bf16x2(silu_fast(acc[t][0]), silu_fast(acc[t][1]));
                hw[t][1] = pack_bf16x2(silu_fast(acc[t][2]), silu_fast(acc[t][3]));
                h[t] = __builtin_bit_cast(bf16x4, hw[t]);
            }
            if (layer == 2) break;
            // the accumulator starts at the next layer's bias; two K = 32 steps
            const int fb = layer == 0 ? 4 : 20;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                f32x4 z = *(const f32x4*)(bs + 64 * (layer + 1) + 16 * t + 4 * g);
#pragma unroll
                for (int s2 = 0; s2 < 2; ++s2)
                    z = mfma_k32(wl[fb + 4 * (2 * s2) + t][lane], wl[fb + 4 * (2 * s2 + 1) + t][lane], h[2 * s2],
                                 h[2 * s2 + 1], z);
                acc[t] = z;
            }
        }
        // layer 4: y = b4 + sum_n W4[n] H3[n]; this lane holds 16 of the 64 neurons of sample col
        float part = 0.f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const f32x4 w = *(const f32x4*)(bs + 196 + 16 * t + 4 * g);
            part += w[0] * __builtin_bit_cast(float, hw[t][0] << 16);
            part += w[1] * __builtin_bit_cast(float, hw[t][0] & 0xffff0000u);
            part += w[2] * __builtin_bit_cast(float, hw[t][1] << 16);
            part += w[3] * __builtin_bit_cast(float, hw[t][1] & 0xffff0000u);
        }
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        if (T == g) y_mine = part + bs[192];
    }
    return y_mine;
}

// LDS copy of the weight A fragments in fragment-major order (conflict-free
// 8-byte reads): [0, 4) W1 n-tiles, [4, 20) W2 (k-step s, n-tile t) at
// 4 + 4 s + t, [20, 36) W3 likewise.
__device__ inline void stage_frags(bf16x4 (*wl)[64], const uint16_t* __restrict__ w1, const uint16_t* __restrict__ w2,
                                   const uint16_t* __restrict__ w3)
{
    for (int e = threadIdx.x; e < 36 * 64; e += blockDim.x) {
        const int f = e >> 6, l = e & 63;
        bf16x4 v;
        if (f < 4) v = load_bfrag(w1, 16, 0, f, l);
        else if (f < 20) v = load_bfrag(w2, 64, (f - 4) >> 2, (f - 4) & 3, l);
        else v = load_bfrag(w3, 64, (f - 20) >> 2, (f - 20) & 3, l);
        wl[f][l] = v;
    }
}

// LDS copy of the biases and of W4 (as FP32) for mlp_forward64:
// [0, 192) b1 b2 b3, [192] b4, [196, 260) W4.
__device__ inline void stage_bias(float* bs, const float* __restrict__ b, const uint16_t* __restrict__ w4)
{
    for (int i = threadIdx.x; i < 260; i += blockDim.x)
        bs[i] = i < 193 ? b[i] : (i >= 196 ? bf2f(w4[i - 196]) : 0.f);
    __syncthreads();
}

// Staging blocks of the learned radius (k_sur_stage's grid), samples per
// wavefront pass of the network (four 16-sample tiles) and the network
// kernel's grid (32-sample passes on twice the wavefronts measured slower).
constexpr int kSurStageBlocks = 1024;
constexpr int kSurSlice = 64;
constexpr int kSurNetBlocks = 640;

// First pass of the learned radius: block b owns the rows [b C, (b + 1) C),
// C = ceil(m / kSurStageBlocks), and copies those that overlap, adhere
// (gamma = min of the two cells' > 0) and lie inside the envelope into its
// staging slots [b C, b C + count) (row, capped overlap, gamma), in any order;
// counts[b] = count. The test is the pair force kernel's, expression for
// expression, so both kernels agree on which rows the network serves. rf.geom
// set: each row's distances are first refreshed from the packed segments
// (updateNeighborDistances) and written back for the pair force kernel.
__global__ void __launch_bounds__(256)
k_sur_stage(const int* __restrict__ d_m, const double* __restrict__ cells, int ld, const int* __restrict__ pi,
            const int* __restrict__ pj, const double* __restrict__ pdx, const double* __restrict__ pdy,
            const double* __restrict__ pdz, RadiusArgs ra, Envelope env, int* __restrict__ st_row,
            double* __restrict__ st_delta, double* __restrict__ st_gamma, int* __restrict__ counts, DistRefresh rf)
{
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int m = *d_m;
    const int C = (m + kSurStageBlocks - 1) / kSurStageBlocks;
    const int lo = blockIdx.x * C, hi = min(m, lo + C);
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int base = lo; base < hi; base += blockDim.x) {
        const int k = base + threadIdx.x;
        bool inside = false;
        double delta = 0, gamma = 0;
        // rows of a build that overflowed its capacity (reported by status) may be stale
        if (k < hi && (unsigned)pi[k] < (unsigned)ld && (unsigned)pj[k] < (unsigned)ld) {
            double d[3], sv, tv;
            row_geometry(rf, k, pi[k], pj[k], pdx, pdy, pdz, nullptr, nullptr, d, sv, tv);
            const double dist = norm3(d[0], d[1], d[2]);
            const double overlap = 2 * ra.R - dist;
            if (overlap > 0) {
                gamma = dmin(cells[(int64_t)ra.colidx_gamma * ld + pi[k]],
                             cells[(int64_t)ra.colidx_gamma * ld + pj[k]]);
                delta = overlap;
                if (ra.max_overlap >= 0 && delta > ra.max_overlap) delta = ra.max_overlap;
                inside = gamma > 0 && sur_inside(delta, gamma, env);
            }
        }
        const uint64_t mk = __ballot(inside);
        int wbase = 0;
        if (lane == 0 && mk) wbase = atomicAdd(&cnt, (int)__popcll(mk));
        wbase = __shfl(wbase, 0);
        if (inside) {
            const int pos = lo + wbase + (int)__popcll(mk & below);
            st_row[pos] = k;
            st_delta[pos] = delta;
            st_gamma[pos] = gamma;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = cnt;
}

// Second pass: the network on the staged rows as consecutive kSurSlice-sample
// slices, one per wavefront at a time (the last slice runs only its filled
// 16-sample tiles), then the FP64 polish; radius[row] for each staged row.
// Sample g of the concatenated staging lists is found from the blocks'
// exclusive prefix counts by a binary search in LDS.
__global__ void __launch_bounds__(256)
k_sur_network(const int* __restrict__ d_m, const int* __restrict__ counts, const int* __restrict__ st_row,
              const double* __restrict__ st_delta, const double* __restrict__ st_gamma, RadiusArgs ra, Envelope env,
              const uint16_t* __restrict__ w1, const uint16_t* __restrict__ w2, const uint16_t* __restrict__ w3,
              const uint16_t* __restrict__ w4, const float* __restrict__ bias, double* __restrict__ radius)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __shared__ __align__(16) float bs[260];
    __shared__ bf16x4 wl[36][64];
    __shared__ int pre[kSurStageBlocks + 1];
    __shared__ int wsum[4];
    // exclusive prefix of the counts: four per thread, then wavefront and block sums
    constexpr int per = kSurStageBlocks / 256;
    int v[per], tsum = 0;
#pragma unroll
    for (int i = 0; i < per; ++i) {
        v[i] = counts[threadIdx.x * per + i];
        tsum += v[i];
    }
    int incl = tsum;
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    stage_frags(wl, w1, w2, w3);
    stage_bias(bs, bias, w4);   // ends with a barrier
    int wofs = 0;
    for (int w = 0; w < wave; ++w) wofs += wsum[w];
    int run = wofs + incl - tsum;
#pragma unroll
    for (int i = 0; i < per; ++i) {
        pre[threadIdx.x * per + i] = run;
        run += v[i];
    }
    if (threadIdx.x == 255) pre[kSurStageBlocks] = run;
    __syncthreads();
    const int total = pre[kSurStageBlocks];
    const int m = *d_m;
    const int C = (m + kSurStageBlocks - 1) / kSurStageBlocks;
    const int nslices = (total + kSurSlice - 1) / kSurSlice;
    for (int sl = blockIdx.x * 4 + wave; sl < nslices; sl += gridDim.x * 4) {
        const int cnt = min(kSurSlice, total - kSurSlice * sl);
        const int g = kSurSlice * sl + lane;
        const bool on = lane < cnt;
        int kq = 0;
        double dq = 0, gq = 0;
        float x[4] = {0.f, 0.f, 0.f, 0.f};
        if (on) {
            int lo = 0, hi = kSurStageBlocks - 1;   // the largest b with pre[b] <= g
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pre[mid] <= g) lo = mid;
                else hi = mid - 1;
            }
            const int idx = lo * C + (g - pre[lo]);
            kq = st_row[idx];
            dq = st_delta[idx];
            gq = st_gamma[idx];
            features_f32(dq, ra.R, gq, env, x);
        }
        const float y = mlp_forward64(wl, bs, x, lane, (cnt + 15) >> 4);
        if (on) radius[kq] = polish_radius_fast(y, dq, ra.Req, ra.E0, gq, ra.polish);
    }
}

// Network evaluation for given (delta, gamma) samples (tests and reports):
// t (raw output) and the polished radius.
__global__ void __launch_bounds__(256)
k_sur_eval(int n, const double* __restrict__ delta, const double* __restrict__ gamma, double R, double Req,
           double E, Envelope env, int polish, const uint16_t* __restrict__ w1, const uint16_t* __restrict__ w2,
           const uint16_t* __restrict__ w3, const uint16_t* __restrict__ w4, const float* __restrict__ bias,
           float* __restrict__ t_out, double* __restrict__ r_out)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __shared__ __align__(16) float bs[260];
    __shared__ bf16x4 wl[36][64];
    stage_frags(wl, w1, w2, w3);
    stage_bias(bs, bias, w4);
    const int waves_total = gridDim.x * 4;
    for (int base = (blockIdx.x * 4 + wave) * 64; base < n; base += waves_total * 64) {
        const int k = base + lane;
        float x[4] = {0.f, 0.f, 0.f, 0.f};
        if (k < n) features(delta[k], R, R, gamma[k], env, x);
        const float y = mlp_forward64(wl, bs, x, lane);
        if (k < n) {
            t_out[k] = y;
            r_out[k] = polish_radius((double)y, delta[k], Req, E, gamma[k], polish);
        }
    }
}

// ---- training kernels ---------------------------------------------------
// Forward of a batch (sample indices idx[0..B)) storing, per sample, the
// pre-activations and activations: act[(6 * b + slot) * 64 + n], slot =
// Z1 H1 Z2 H2 Z3 H3; y[b]; and the BF16-rounded inputs xin[b * 4 + c].
__global__ void __launch_bounds__(256)
k_train_fwd(int B, const int* __restrict__ idx, const float* __restrict__ X, const uint16_t* __restrict__ w1,
            const uint16_t* __restrict__ w2, const uint16_t* __restrict__ w3, const uint16_t* __restrict__ w4,
            const float* __restrict__ bias, float* __restrict__ act, float* __restrict__ xin,
            float* __restrict__ y)
{
    __shared__ float feat[4][16 * 4];
    __shared__ uint16_t hbuf[4][16 * 68];
    __shared__ float zs[4][6 * 1024];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    WFrags F;
    load_frags(SurrogateWeights{}, w1, w2, w3, lane, F);
    const int tiles = (B + 15) / 16;
    for (int tile = blockIdx.x * 4 + wave; tile < tiles; tile += gridDim.x * 4) {
        const int b0 = tile * 16;
        if (lane < 16) {
            const int b = b0 + lane;
            for (int c = 0; c < 4; ++c) {
                const float v = b < B ? X[(int64_t)idx[b] * 4 + c] : 0.f;
                feat[wave][lane * 4 + c] = v;
                if (b < B) xin[(int64_t)b * 4 + c] = bfround(v);
            }
        }
        __builtin_amdgcn_wave_barrier();
        float y4[4];
        mlp_tile(F, bias, w4, feat[wave], hbuf[wave], lane, y4, zs[wave]);
        if ((lane & 15) == 0)
            for (int i = 0; i < 4; ++i) {
                const int b = b0 + 4 * (lane >> 4) + i;
                if (b < B) y[b] = y4[i];
            }
        for (int e = lane; e < 16 * 64; e += 64) {
            const int mrow = e >> 6, n = e & 63, b = b0 + mrow;
            if (b < B)
                for (int sl = 0; sl < 6; ++sl) act[((int64_t)6 * b + sl) * 64 + n] = zs[wave][sl * 1024 + e];
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// dY = 2 (y - target) / B and the layer-4 gradients' row part: dZ3 = dY W4 * silu'(Z3).
__global__ void k_train_out(int B, const int* __restrict__ idx, const float* __restrict__ T,
                            const float* __restrict__ y, const uint16_t* __restrict__ w4,
                            const float* __restrict__ act, float* __restrict__ dy, float* __restrict__ dz,
                            float* __restrict__ loss)
{
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
        const float e = y[b] - T[idx[b]];
        const float g = 2.0f * e / (float)B;
        dy[b] = g;
        atomicAdd(loss, e * e / (float)B);
        for (int n = 0; n < 64; ++n)
            dz[(int64_t)b * 64 + n] = g * bf2f(w4[n]) * dsilu(act[((int64_t)6 * b + 4) * 64 + n]);
    }
}

// dZ_prev = (dZ W) * silu'(Z_prev) for one hidden layer (W: BF16 [64][64]).
__global__ void __launch_bounds__(256)
k_train_back(int B, const float* __restrict__ dz, const uint16_t* __restrict__ W, const float* __restrict__ act,
             int zslot, float* __restrict__ dz_prev)
{
    __shared__ float Ws[64 * 64];
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) Ws[e] = bf2f(W[e]);
    __syncthreads();
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
        float g[64];
        for (int n = 0; n < 64; ++n) g[n] = dz[(int64_t)b * 64 + n];
        for (int k = 0; k < 64; ++k) {
            float s = 0.f;
            for (int n = 0; n < 64; ++n) s += g[n] * Ws[n * 64 + k];
            dz_prev[(int64_t)b * 64 + k] = s * dsilu(act[((int64_t)6 * b + zslot) * 64 + k]);
        }
    }
}

// Weight gradients dW[n][k] += sum_b dZ[b][n] H[b][k], db[n] += sum_b dZ[b][n]
// over a chunk of 64 samples per workgroup (atomics into the FP32 gradient).
__global__ void __launch_bounds__(256)
k_train_wgrad(int B, int K, const float* __restrict__ dz, int dz_stride, const float* __restrict__ hin,
              int64_t h_stride, int h_off, float* __restrict__ gW, float* __restrict__ gb, int nout)
{
    __shared__ float sdz[64][65];
    __shared__ float sh[64][65];
    const int b0 = blockIdx.x * 64;
    const int nb = min(64, B - b0);
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int r = e >> 6, c = e & 63;
        sdz[r][c] = (r < nb && c < nout) ? dz[(int64_t)(b0 + r) * dz_stride + c] : 0.f;
        sh[r][c] = (r < nb && c < K) ? hin[(int64_t)(b0 + r) * h_stride + h_off + c] : 0.f;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nout * K; e += 256) {
        const int n = e / K, k = e % K;
        float s = 0.f;
        for (int r = 0; r < nb; ++r) s += sdz[r][n] * sh[r][k];
        atomicAdd(&gW[n * K + k], s);
    }
    for (int n = threadIdx.x; n < nout; n += 256) {
        float s = 0.f;
        for (int r = 0; r < nb; ++r) s += sdz[r][n];
        atomicAdd(&gb[n], s);
    }
}

__global__ void k_zero_f(float* p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.f;
}

// Adam on the FP32 master parameters; writes the BF16 / FP32 copies the
// forward uses (W1 padded to [64][16]).
__global__ void k_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M1,
                       float* __restrict__ M2, float lr, float b1, float b2, float eps, float c1, float c2,
                       uint16_t* __restrict__ w1, uint16_t* __restrict__ w2, uint16_t* __restrict__ w3,
                       uint16_t* __restrict__ w4, float* __restrict__ bias, int update)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kNParams; i += gridDim.x * blockDim.x) {
        float p = P[i];
        if (update) {
            const float g = G[i];
            const float m = b1 * M1[i] + (1 - b1) * g;
            const float v = b2 * M2[i] + (1 - b2) * g * g;
            M1[i] = m;
            M2[i] = v;
            p -= lr * (m / c1) / (sqrtf(v / c2) + eps);
            P[i] = p;
        }
        if (i < kB1) { const int n = i / 4, k = i % 4; w1[n * 16 + k] = f2bf(p); }
        else if (i < kW2) bias[i - kB1] = p;
        else if (i < kB2) w2[i - kW2] = f2bf(p);
        else if (i < kW3) bias[64 + i - kB2] = p;
        else if (i < kB3) w3[i - kW3] = f2bf(p);
        else if (i < kW4) bias[128 + i - kB3] = p;
        else if (i < kB4) w4[i - kW4] = f2bf(p);
        else bias[192] = p;
    }
}

// Training samples: delta / R ~ U[dlo, dhi], gamma ~ U[glo, ghi] from a
// counter-based hash of (seed, index); target t = u / p^(1/3) of the exact
// FP64 root; samples below pull-off (no root) are flagged invalid.
__device__ inline double hash01(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    x ^= x >> 31;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void k_sur_gen(int n, uint64_t seed, Envelope env, double Req, double E, float* __restrict__ X,
                          float* __restrict__ T, int* __restrict__ ok)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double dn = env.dlo + (env.dhi - env.dlo) * hash01(seed * 0x100000001ull + 2 * (uint64_t)i);
        const double gamma = env.glo + (env.ghi - env.glo) * hash01(seed * 0x100000001ull + 2 * (uint64_t)i + 1);
        const double delta = dn * env.R;
        const double a = jkr_radius_newton(delta, Req, E, gamma, 0.0, 400, nullptr);
        const double p = Req * sqrt(4 * kPi * gamma / E);
        const double u = sqrt(a);
        const double q = Req * delta;
        const double resid = (a * a - p * u) - q;
        const bool valid = a > 0 && fabs(resid) <= 1e-9 * fmax(1.0, fabs(q));
        float x[4];
        features(delta, env.R, env.R, gamma, env, x);
        for (int c = 0; c < 4; ++c) X[(int64_t)i * 4 + c] = x[c];
        T[i] = valid ? (float)(u / cbrt(p)) : 0.f;
        ok[i] = valid;
    }
}

}  // namespace bfm

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

struct HipErr {
    hipError_t e;
    const char* what;
};
inline void hc(hipError_t e, const char* w)
{
    if (e != hipSuccess) throw HipErr{e, w};
}
#define SHIP(x) hc((x), #x)

template <typename T>
void dmal(T** p, size_t n)
{
    SHIP(hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1)));
}

Envelope default_env(double R)
{
    Envelope e;
    e.dlo = -0.2; e.dhi = 0.3; e.glo = 10.0; e.ghi = 1000.0; e.R = R;
    return e;
}

void upload_weights(bfm::SurrogateWeights* s, const uint16_t* w1, const float* b1, const uint16_t* w2,
                    const float* b2, const uint16_t* w3, const float* b3, const uint16_t* w4, const float* b4)
{
    std::vector<uint16_t> w1p(64 * 16, 0);
    for (int n = 0; n < 64; ++n)
        for (int k = 0; k < 4; ++k) w1p[n * 16 + k] = w1[n * 4 + k];
    SHIP(hipMemcpy(s->w1, w1p.data(), 64 * 16 * 2, hipMemcpyHostToDevice));
    SHIP(hipMemcpy(s->w2, w2, 64 * 64 * 2, hipMemcpyHostToDevice));
    SHIP(hipMemcpy(s->w3, w3, 64 * 64 * 2, hipMemcpyHostToDevice));
    SHIP(hipMemcpy(s->w4, w4, 64 * 2, hipMemcpyHostToDevice));
    std::vector<float> bb(193);
    std::memcpy(bb.data(), b1, 64 * 4);
    std::memcpy(bb.data() + 64, b2, 64 * 4);
    std::memcpy(bb.data() + 128, b3, 64 * 4);
    bb[192] = b4[0];
    SHIP(hipMemcpy(s->b, bb.data(), 193 * 4, hipMemcpyHostToDevice));
}

bfm::SurrogateWeights* ensure_sur(bfm_ctx* c)
{
    if (!c->sur) {
        auto* s = new bfm::SurrogateWeights();
        dmal(&s->w1, 64 * 16); dmal(&s->w2, 64 * 64); dmal(&s->w3, 64 * 64); dmal(&s->w4, 64);
        dmal(&s->b, 193);
        SHIP(hipMemset(s->w1, 0, 64 * 16 * 2));   // K padding of layer 1 must stay zero
        SHIP(hipMemset(s->b, 0, 193 * 4));
        s->env = default_env(0.8);
        c->sur = s;
    }
    return c->sur;
}

int fail(bfm_ctx* c, int code, const std::string& msg)
{
    c->err = msg;
    return code;
}

}  // namespace

namespace bfm {

const double* surrogate_pair_radii(bfm_ctx* c, const double* d_cells, int ld, int n, const DistRefresh& rf)
{
    (void)n;
    SurrogateWeights* s = c->sur;
    if (s->cap < c->max_pairs) {
        for (void* q : {(void*)s->radius, (void*)s->st_row, (void*)s->st_delta, (void*)s->st_gamma,
                        (void*)s->st_count})
            if (q) SHIP(hipFree(q));
        dmal(&s->radius, c->max_pairs);
        dmal(&s->st_row, c->max_pairs);
        dmal(&s->st_delta, c->max_pairs);
        dmal(&s->st_gamma, c->max_pairs);
        dmal(&s->st_count, kSurStageBlocks);
        s->cap = c->max_pairs;
    }
    RadiusArgs ra;
    ra.R = c->p.R; ra.Rcell = c->p.Rcell; ra.E0 = c->p.E0; ra.Req = c->p.R / 2;
    ra.max_overlap = 2 * (c->p.R - c->p.Rcell);
    ra.colidx_gamma = c->p.colidx_gamma;
    ra.polish = s->polish;
    bfm::launch_timed(c, "sur_stage", [&] {
        hipLaunchKernelGGL(k_sur_stage, dim3(kSurStageBlocks), dim3(256), 0, c->stream, c->d_m, d_cells, ld, c->pi,
                           c->pj, c->pdx, c->pdy, c->pdz, ra, s->env, s->st_row, s->st_delta, s->st_gamma,
                           s->st_count, rf);
    });
    bfm::launch_timed(c, "sur_network", [&] {
        hipLaunchKernelGGL(k_sur_network, dim3(kSurNetBlocks), dim3(256), 0, c->stream, c->d_m, s->st_count, s->st_row,
                           s->st_delta, s->st_gamma, ra, s->env, s->w1, s->w2, s->w3, s->w4, s->b, s->radius);
    });
    return s->radius;
}

SurrogateEnvelope surrogate_envelope(const bfm_ctx* c)
{
    if (c->sur) return c->sur->env;
    SurrogateEnvelope e;
    e.dlo = 1; e.dhi = 0; e.glo = 1; e.ghi = 0; e.R = 1;   // contains nothing
    return e;
}

void surrogate_free(bfm_ctx* c)
{
    if (!c->sur) return;
    for (void* q : {(void*)c->sur->w1, (void*)c->sur->w2, (void*)c->sur->w3, (void*)c->sur->w4,
                    (void*)c->sur->b, (void*)c->sur->radius, (void*)c->sur->st_row, (void*)c->sur->st_delta,
                    (void*)c->sur->st_gamma, (void*)c->sur->st_count})
        if (q) (void)hipFree(q);
    delete c->sur;
    c->sur = nullptr;
}

}  // namespace bfm

extern "C" {

int bfm_set_surrogate(bfm_ctx* c, const uint16_t* w1, const float* b1, const uint16_t* w2, const float* b2,
                      const uint16_t* w3, const float* b3, const uint16_t* w4, const float* b4,
                      const double* envelope)
{
    if (!c) return BFM_ERR_ARG;
    if (!w1 || !b1 || !w2 || !b2 || !w3 || !b3 || !w4 || !b4) return fail(c, BFM_ERR_ARG, "NULL weights");
    try {
        SHIP(hipSetDevice(c->device));
        auto* s = ensure_sur(c);
        upload_weights(s, w1, b1, w2, b2, w3, b3, w4, b4);
        if (envelope) {
            s->env.dlo = envelope[0] / s->env.R;
            s->env.dhi = envelope[1] / s->env.R;
            s->env.glo = envelope[2];
            s->env.ghi = envelope[3];
        }
        return BFM_OK;
    } catch (const HipErr& e) {
        return fail(c, BFM_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
    }
}

int bfm_surrogate_set_polish(bfm_ctx* c, int polish_steps)
{
    if (!c || polish_steps < 0 || polish_steps > 50) return BFM_ERR_ARG;
    ensure_sur(c)->polish = polish_steps;
    return BFM_OK;
}

int bfm_surrogate_eval(bfm_ctx* c, const double* delta, const double* gamma, int n, int polish, float* t_out,
                       double* radius_out)
{
    if (!c || !c->sur) return BFM_ERR_ARG;
    try {
        SHIP(hipSetDevice(c->device));
        auto* s = c->sur;
        double *dd, *dg, *dr;
        float* dt;
        dmal(&dd, n); dmal(&dg, n); dmal(&dr, n); dmal(&dt, n);
        SHIP(hipMemcpy(dd, delta, sizeof(double) * n, hipMemcpyHostToDevice));
        SHIP(hipMemcpy(dg, gamma, sizeof(double) * n, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_sur_eval, dim3(std::max(1, std::min(4096, (n + 255) / 256))), dim3(256), 0, c->stream,
                           n, dd, dg, s->env.R, c->p.R / 2, c->p.E0, s->env, polish, s->w1, s->w2, s->w3, s->w4,
                           s->b, dt, dr);
        SHIP(hipGetLastError());
        SHIP(hipStreamSynchronize(c->stream));
        if (t_out) SHIP(hipMemcpy(t_out, dt, sizeof(float) * n, hipMemcpyDeviceToHost));
        if (radius_out) SHIP(hipMemcpy(radius_out, dr, sizeof(double) * n, hipMemcpyDeviceToHost));
        for (void* q : {(void*)dd, (void*)dg, (void*)dr, (void*)dt}) SHIP(hipFree(q));
        return BFM_OK;
    } catch (const HipErr& e) {
        return fail(c, BFM_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
    }
}

// One forward + backward pass on a batch (FP32 master parameters params[],
// inputs X [B][4], targets T [B]); returns the loss, y [B] and the gradient
// grad[kNParams]. Used to check the GPU gradients against the CPU oracle.
int bfm_surrogate_grad(bfm_ctx* c, const float* params, const float* X, const float* T, int B, float* loss_out,
                       float* y_out, float* grad_out)
{
    if (!c || B < 1) return BFM_ERR_ARG;
    try {
        SHIP(hipSetDevice(c->device));
        auto* s = ensure_sur(c);
        const int NP = kNParams;
        float *P, *G, *dX, *dT, *act, *xin, *y, *dy, *dz, *dz2, *loss;
        int* idx;
        dmal(&P, NP); dmal(&G, NP); dmal(&dX, (size_t)B * 4); dmal(&dT, B); dmal(&act, (size_t)B * 6 * 64);
        dmal(&xin, (size_t)B * 4); dmal(&y, B); dmal(&dy, B); dmal(&dz, (size_t)B * 64); dmal(&dz2, (size_t)B * 64);
        dmal(&loss, 1); dmal(&idx, B);
        std::vector<int> hidx(B);
        for (int i = 0; i < B; ++i) hidx[i] = i;
        SHIP(hipMemcpy(P, params, sizeof(float) * NP, hipMemcpyHostToDevice));
        SHIP(hipMemcpy(dX, X, sizeof(float) * B * 4, hipMemcpyHostToDevice));
        SHIP(hipMemcpy(dT, T, sizeof(float) * B, hipMemcpyHostToDevice));
        SHIP(hipMemcpy(idx, hidx.data(), sizeof(int) * B, hipMemcpyHostToDevice));
        hipStream_t st = c->stream;
        hipLaunchKernelGGL(k_adam, dim3(64), dim3(256), 0, st, P, G, G, G, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, s->w1,
                           s->w2, s->w3, s->w4, s->b, 0);
        hipLaunchKernelGGL(k_zero_f, dim3(64), dim3(256), 0, st, G, NP);
        hipLaunchKernelGGL(k_zero_f, dim3(1), dim3(64), 0, st, loss, 1);
        hipLaunchKernelGGL(k_train_fwd, dim3(std::max(1, (B + 63) / 64)), dim3(256), 0, st, B, idx, dX, s->w1, s->w2,
                           s->w3, s->w4, s->b, act, xin, y);
        hipLaunchKernelGGL(k_train_out, dim3((B + 255) / 256), dim3(256), 0, st, B, idx, dT, y, s->w4, act, dy,
                           dz, loss);
        const int gw = (B + 63) / 64;
        // layer 4: dW4 = sum dY H3, db4 = sum dY
        hipLaunchKernelGGL(k_train_wgrad, dim3(gw), dim3(256), 0, st, B, 64, dy, 1, act, (int64_t)6 * 64, 5 * 64,
                           G + kW4, G + kB4, 1);
        // layer 3
        hipLaunchKernelGGL(k_train_wgrad, dim3(gw), dim3(256), 0, st, B, 64, dz, 64, act, (int64_t)6 * 64, 3 * 64,
                           G + kW3, G + kB3, 64);
        hipLaunchKernelGGL(k_train_back, dim3((B + 255) / 256), dim3(256), 0, st, B, dz, s->w3, act, 2, dz2);
        // layer 2
        hipLaunchKernelGGL(k_train_wgrad, dim3(gw), dim3(256), 0, st, B, 64, dz2, 64, act, (int64_t)6 * 64, 1 * 64,
                           G + kW2, G + kB2, 64);
        hipLaunchKernelGGL(k_train_back, dim3((B + 255) / 256), dim3(256), 0, st, B, dz2, s->w2, act, 0, dz);
        // layer 1
        hipLaunchKernelGGL(k_train_wgrad, dim3(gw), dim3(256), 0, st, B, 4, dz, 64, xin, (int64_t)4, 0, G + kW1,
                           G + kB1, 64);
        SHIP(hipGetLastError());
        SHIP(hipStreamSynchronize(st));
        if (loss_out) SHIP(hipMemcpy(loss_out, loss, sizeof(float), hipMemcpyDeviceToHost));
        if (y_out) SHIP(hipMemcpy(y_out, y, sizeof(float) * B, hipMemcpyDeviceToHost));
        if (grad_out) SHIP(hipMemcpy(grad_out, G, sizeof(float) * NP, hipMemcpyDeviceToHost));
        for (void* q : {(void*)P, (void*)G, (void*)dX, (void*)dT, (void*)act, (void*)xin, (void*)y, (void*)dy,
                        (void*)dz, (void*)dz2, (void*)loss, (void*)idx})
            SHIP(hipFree(q));
        return BFM_OK;
    } catch (const HipErr& e) {
        return fail(c, BFM_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
    }
}

// Trains the network on the GPU: n_samples valid samples of the exact solve
// (seeded), Adam with a cosine-decayed learning rate from lr. Installs the
// trained weights in the context and returns the FP32 master parameters
// (params_out[kNParams]) and the per-epoch mean loss (loss_out[epochs]).
int bfm_surrogate_train(bfm_ctx* c, int n_samples, int epochs, int batch, double lr, uint64_t seed,
                        float* params_out, float* loss_out)
{
    if (!c || n_samples < 1 || epochs < 1 || batch < 16) return BFM_ERR_ARG;
    try {
        SHIP(hipSetDevice(c->device));
        auto* s = ensure_sur(c);
        const double Req = c->p.R / 2, E = c->p.E0;
        s->env = default_env(c->p.R);
        hipStream_t st = c->stream;
        // samples: draw in rounds, keep the valid ones
        std::vector<float> hX, hT;
        hX.reserve((size_t)n_samples * 4);
        hT.reserve(n_samples);
        {
            const int chunk = n_samples + n_samples / 4 + 1024;
            float *dX, *dT;
            int* dok;
            dmal(&dX, (size_t)chunk * 4); dmal(&dT, chunk); dmal(&dok, chunk);
            std::vector<float> bx((size_t)chunk * 4), bt(chunk);
            std::vector<int> bok(chunk);
            for (uint64_t round = 0; (int)hT.size() < n_samples; ++round) {
                hipLaunchKernelGGL(k_sur_gen, dim3(1024), dim3(256), 0, st, chunk, seed * 1000003ull + round,
                                   s->env, Req, E, dX, dT, dok);
                SHIP(hipGetLastError());
                SHIP(hipStreamSynchronize(st));
                SHIP(hipMemcpy(bx.data(), dX, sizeof(float) * chunk * 4, hipMemcpyDeviceToHost));
                SHIP(hipMemcpy(bt.data(), dT, sizeof(float) * chunk, hipMemcpyDeviceToHost));
                SHIP(hipMemcpy(bok.data(), dok, sizeof(int) * chunk, hipMemcpyDeviceToHost));
                for (int i = 0; i < chunk && (int)hT.size() < n_samples; ++i)
                    if (bok[i]) {
                        hX.insert(hX.end(), bx.begin() + 4 * (size_t)i, bx.begin() + 4 * (size_t)i + 4);
                        hT.push_back(bt[i]);
                    }
            }
            for (void* q : {(void*)dX, (void*)dT, (void*)dok}) SHIP(hipFree(q));
        }
        const int N = n_samples, NP = kNParams;
        float *X, *T, *P, *G, *M1, *M2, *act, *xin, *y, *dy, *dz, *dz2, *loss;
        int* idx;
        dmal(&X, (size_t)N * 4); dmal(&T, N); dmal(&P, NP); dmal(&G, NP); dmal(&M1, NP); dmal(&M2, NP);
        dmal(&act, (size_t)batch * 6 * 64); dmal(&xin, (size_t)batch * 4); dmal(&y, batch); dmal(&dy, batch);
        dmal(&dz, (size_t)batch * 64); dmal(&dz2, (size_t)batch * 64); dmal(&loss, 1); dmal(&idx, N);
        SHIP(hipMemcpy(X, hX.data(), sizeof(float) * N * 4, hipMemcpyHostToDevice));
        SHIP(hipMemcpy(T, hT.data(), sizeof(float) * N, hipMemcpyHostToDevice));
        // He-style init (seeded)
        std::mt19937_64 rng(seed);
        std::normal_distribution<double> nd(0.0, 1.0);
        std::vector<float> hp(NP, 0.f);
        for (int i = 0; i < 64 * 4; ++i) hp[kW1 + i] = (float)(nd(rng) * std::sqrt(2.0 / 4));
        for (int i = 0; i < 64 * 64; ++i) hp[kW2 + i] = (float)(nd(rng) * std::sqrt(2.0 / 64));
        for (int i = 0; i < 64 * 64; ++i) hp[kW3 + i] = (float)(nd(rng) * std::sqrt(2.0 / 64));
        for (int i = 0; i < 64; ++i) hp[kW4 + i] = (float)(nd(rng) * std::sqrt(1.0 / 64));
        // output bias at the mean target
        double mt = 0;
        for (float t : hT) mt += t;
        hp[kB4] = (float)(mt / N);
        SHIP(hipMemcpy(P, hp.data(), sizeof(float) * NP, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_zero_f, dim3(64), dim3(256), 0, st, M1, NP);
        hipLaunchKernelGGL(k_zero_f, dim3(64), dim3(256), 0, st, M2, NP);
        hipLaunchKernelGGL(k_adam, dim3(64), dim3(256), 0, st, P, G, M1, M2, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, s->w1,
                           s->w2, s->w3, s->w4, s->b, 0);
        std::vector<int> perm(N);
        for (int i = 0; i < N; ++i) perm[i] = i;
        const int steps_per_epoch = N / batch;
        const int total = steps_per_epoch * epochs;
        int step = 0;
        const float b1 = 0.9f, b2 = 0.999f;
        for (int ep = 0; ep < epochs; ++ep) {
            std::shuffle(perm.begin(), perm.end(), rng);
            SHIP(hipMemcpy(idx, perm.data(), sizeof(int) * N, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_zero_f, dim3(1), dim3(64), 0, st, loss, 1);
            for (int k = 0; k < steps_per_epoch; ++k, ++step) {
                const int* bidx = idx + (size_t)k * batch;
                const int B = batch;
                const int gw = (B + 63) / 64;
                hipLaunchKernelGGL(k_zero_f, dim3(64), dim3(256), 0, st, G, NP);
                hipLaunchKernelGGL(k_train_fwd, dim3(std::max(1, (B + 63) / 64)), dim3(256), 0, st, B, bidx, X,
                                   s->w1, s->w2, s->w3, s->w4, s->b, act, xin, y);
                hipLaunchKernelGGL(k_train_out, dim3((B + 255) / 256), dim3(256), 0, st, B, bidx, T, y, s->w4, act,
                                   dy, dz, loss);
                hipLaunchKernelGGL(k_train_wgrad, dim3(gw), dim3(256), 0, st, B, 64, dy, 1, act, (int64_t)6 * 64,
                                   5 * 64, G + kW4, G + kB4, 1);
                hipLaunchKernelGGL(k_train_wgrad, dim3(gw), dim3(256), 0, st, B, 64, dz, 64, act, (int64_t)6 * 64,
                                   3 * 64, G + kW3, G + kB3, 64);
                hipLaunchKernelGGL(k_train_back, dim3((B + 255) / 256), dim3(256), 0, st, B, dz, s->w3, act, 2, dz2);
                hipLaunchKernelGGL(k_train_wgrad, dim3(gw), dim3(256), 0, st, B, 64, dz2, 64, act, (int64_t)6 * 64,
                                   1 * 64, G + kW2, G + kB2, 64);
                hipLaunchKernelGGL(k_train_back, dim3((B + 255) / 256), dim3(256), 0, st, B, dz2, s->w2, act, 0, dz);
                hipLaunchKernelGGL(k_train_wgrad, dim3(gw), dim3(256), 0, st, B, 4, dz, 64, xin, (int64_t)4, 0,
                                   G + kW1, G + kB1, 64);
                const double frac = (double)step / std::max(1, total);
                const float lr_t = (float)(lr * 0.5 * (1 + std::cos(kPi * frac)) + lr * 1e-3);
                const float c1 = 1 - std::pow(b1, (float)(step + 1)), c2 = 1 - std::pow(b2, (float)(step + 1));
                hipLaunchKernelGGL(k_adam, dim3(64), dim3(256), 0, st, P, G, M1, M2, lr_t, b1, b2, 1e-8f, c1, c2,
                                   s->w1, s->w2, s->w3, s->w4, s->b, 1);
            }
            SHIP(hipGetLastError());
            float l = 0;
            SHIP(hipMemcpy(&l, loss, sizeof(float), hipMemcpyDeviceToHost));
            if (loss_out) loss_out[ep] = l / std::max(1, steps_per_epoch);
            if (std::getenv("BFM_TRAIN_LOG")) {
                std::printf("surrogate epoch %d loss %.3e\n", ep, l / std::max(1, steps_per_epoch));
                std::fflush(stdout);
            }
        }
        SHIP(hipStreamSynchronize(st));
        if (params_out) SHIP(hipMemcpy(params_out, P, sizeof(float) * NP, hipMemcpyDeviceToHost));
        for (void* q : {(void*)X, (void*)T, (void*)P, (void*)G, (void*)M1, (void*)M2, (void*)act, (void*)xin,
                        (void*)y, (void*)dy, (void*)dz, (void*)dz2, (void*)loss, (void*)idx})
            SHIP(hipFree(q));
        return BFM_OK;
    } catch (const HipErr& e) {
        return fail(c, BFM_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
    }
}

int bfm_surrogate_generate(bfm_ctx* c, int n, uint64_t seed, float* X, float* T, int* ok)
{
    if (!c || n < 1) return BFM_ERR_ARG;
    try {
        SHIP(hipSetDevice(c->device));
        auto* s = ensure_sur(c);
        s->env = default_env(c->p.R);
        float *dX, *dT;
        int* dok;
        dmal(&dX, (size_t)n * 4); dmal(&dT, n); dmal(&dok, n);
        hipLaunchKernelGGL(k_sur_gen, dim3(1024), dim3(256), 0, c->stream, n, seed, s->env, c->p.R / 2, c->p.E0, dX,
                           dT, dok);
        SHIP(hipGetLastError());
        SHIP(hipStreamSynchronize(c->stream));
        SHIP(hipMemcpy(X, dX, sizeof(float) * n * 4, hipMemcpyDeviceToHost));
        SHIP(hipMemcpy(T, dT, sizeof(float) * n, hipMemcpyDeviceToHost));
        SHIP(hipMemcpy(ok, dok, sizeof(int) * n, hipMemcpyDeviceToHost));
        for (void* q : {(void*)dX, (void*)dT, (void*)dok}) SHIP(hipFree(q));
        return BFM_OK;
    } catch (const HipErr& e) {
        return fail(c, BFM_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
    }
}

}  // extern "C"

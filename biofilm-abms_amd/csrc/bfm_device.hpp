// bfm_device.hpp -- device-side math of the mechanical step (CDNA4, FP64).
//
// Every function restates a reference routine of ABM-3d/ABM-3d-cpp/include
// (cited per function). Expressions keep the reference's evaluation order and
// the library is compiled with -ffp-contract=off, so the geometry (distances,
// contact coordinates, the neighbor predicate) rounds exactly like the CPU
// oracle. Non-integer powers use sqrt products (x^1.5 = x * sqrt(x), ...),
// which differ from libm pow by at most an ulp or two.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bfm {

constexpr double kPi = 3.14159265358979323846;

// Column indices of the cells array, include/indices.hpp:17-47.
enum {
    C_ID = 0, C_RX = 1, C_RY = 2, C_RZ = 3, C_NX = 4, C_NY = 5, C_NZ = 6,
    C_DRX = 7, C_DRY = 8, C_DRZ = 9, C_DNX = 10, C_DNY = 11, C_DNZ = 12,
    C_L = 13, C_HALF_L = 14, C_T0 = 15, C_GROWTH = 16, C_ETA0 = 17,
    C_ETA1 = 18, C_SIGMA0 = 19, C_GROUP = 20
};

// std::max / std::min semantics (first argument wins ties).
__host__ __device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }

__host__ __device__ __forceinline__ double norm3(double x, double y, double z)
{
    return sqrt((x * x + y * y) + z * z);
}

// x^(E2 / 2) for x >= 0 and a small non-negative integer E2.
template <int E2>
__device__ __forceinline__ double pow_half(double x)
{
    if constexpr (E2 == 0) return 1.0;
    else if constexpr (E2 == 1) return sqrt(x);
    else if constexpr (E2 == 2) return x;
    else if constexpr (E2 == 3) return x * sqrt(x);
    else if constexpr (E2 == 4) return x * x;
    else if constexpr (E2 == 5) return (x * x) * sqrt(x);
    else if constexpr (E2 == 6) return (x * x) * x;
    else if constexpr (E2 == 7) return ((x * x) * x) * sqrt(x);
    else static_assert(E2 <= 7, "exponent out of range");
    return 0.0;
}

struct Seg {
    double r[3];
    double n[3];
    double hl;
};

// nearestCellBodyCoordToPoint, distances.hpp:92-105.
__host__ __device__ __forceinline__ double nearest_coord(const Seg& c, double qx, double qy, double qz)
{
    double s = ((qx - c.r[0]) * c.n[0] + (qy - c.r[1]) * c.n[1]) + (qz - c.r[2]) * c.n[2];
    if (fabs(s) <= c.hl) return s;
    else if (s > c.hl) return c.hl;
    else return -c.hl;
}

// distBetweenCells, distances.hpp:206-357. d runs from cell 1 to cell 2.
// The skew branch is the clamped minimiser of |r12 + t n2 - s n1|^2 (the
// reference calls CGAL's segment-segment distance there).
__host__ __device__ __forceinline__ void dist_between_cells(const Seg& c1, const Seg& c2, double d[3],
                                                   double& s_out, double& t_out)
{
    double s = 0, t = 0;
    double cos_theta = (c1.n[0] * c2.n[0] + c1.n[1] * c2.n[1]) + c1.n[2] * c2.n[2];
    if (cos_theta >= 0.9999 || cos_theta <= -0.9999) {
        double p1[3], q1[3], p2[3], q2[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            p1[k] = c1.r[k] - c1.hl * c1.n[k];
            q1[k] = c1.r[k] + c1.hl * c1.n[k];
            p2[k] = c2.r[k] - c2.hl * c2.n[k];
            q2[k] = c2.r[k] + c2.hl * c2.n[k];
        }
        double s_p1 = nearest_coord(c2, p1[0], p1[1], p1[2]);
        double s_q1 = nearest_coord(c2, q1[0], q1[1], q1[2]);
        double dp1[3], dq1[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            dp1[k] = (c2.r[k] + s_p1 * c2.n[k]) - p1[k];
            dq1[k] = (c2.r[k] + s_q1 * c2.n[k]) - q1[k];
        }
        double dist_p1 = norm3(dp1[0], dp1[1], dp1[2]);
        double dist_q1 = norm3(dq1[0], dq1[1], dq1[2]);
        if (s_p1 == s_q1) {
            if (dist_p1 < dist_q1) {
                s = -c1.hl; t = s_p1;
                d[0] = dp1[0]; d[1] = dp1[1]; d[2] = dp1[2];
            } else {
                s = c1.hl; t = s_q1;
                d[0] = dq1[0]; d[1] = dq1[1]; d[2] = dq1[2];
            }
        } else {
            double s_p2 = nearest_coord(c1, p2[0], p2[1], p2[2]);
            double s_q2 = nearest_coord(c1, q2[0], q2[1], q2[2]);
            double dp2[3], dq2[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                dp2[k] = (c1.r[k] + s_p2 * c1.n[k]) - p2[k];
                dq2[k] = (c1.r[k] + s_q2 * c1.n[k]) - q2[k];
            }
            double dists[4] = {dist_p1, dist_q1, norm3(dp2[0], dp2[1], dp2[2]),
                               norm3(dq2[0], dq2[1], dq2[2])};
            // Stable ranking of the four distances (ties keep index order);
            // the two smallest select the branch (distances.hpp:274-335).
            int rank[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                int r = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    r += (dists[b] < dists[a]) || (dists[b] == dists[a] && b < a);
                rank[a] = r;
            }
            bool has0 = rank[0] < 2, has1 = rank[1] < 2, has2 = rank[2] < 2;
            if (has0) {
                if (has1) {
                    s = 0.0;
                    t = nearest_coord(c2, c1.r[0], c1.r[1], c1.r[2]);
#pragma unroll
                    for (int k = 0; k < 3; ++k) d[k] = (c2.r[k] + t * c2.n[k]) - c1.r[k];
                } else {
                    s = has2 ? (-c1.hl + s_p2) / 2 : (-c1.hl + s_q2) / 2;
                    t = nearest_coord(c2, c1.r[0] + s * c1.n[0], c1.r[1] + s * c1.n[1],
                                      c1.r[2] + s * c1.n[2]);
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        d[k] = ((c2.r[k] + t * c2.n[k]) - c1.r[k]) - s * c1.n[k];
                }
            } else if (has1) {
                s = has2 ? (c1.hl + s_p2) / 2 : (c1.hl + s_q2) / 2;
                t = nearest_coord(c2, c1.r[0] + s * c1.n[0], c1.r[1] + s * c1.n[1],
                                  c1.r[2] + s * c1.n[2]);
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    d[k] = ((c2.r[k] + t * c2.n[k]) - c1.r[k]) - s * c1.n[k];
            } else {
                t = 0.0;
                s = nearest_coord(c1, c2.r[0], c2.r[1], c2.r[2]);
#pragma unroll
                for (int k = 0; k < 3; ++k) d[k] = (c2.r[k] - c1.r[k]) - s * c1.n[k];
            }
        }
    } else {
        double r12[3] = {c2.r[0] - c1.r[0], c2.r[1] - c1.r[1], c2.r[2] - c1.r[2]};
        double a = (c1.n[0] * c1.n[0] + c1.n[1] * c1.n[1]) + c1.n[2] * c1.n[2];
        double b = (c1.n[0] * c2.n[0] + c1.n[1] * c2.n[1]) + c1.n[2] * c2.n[2];
        double c = (c2.n[0] * c2.n[0] + c2.n[1] * c2.n[1]) + c2.n[2] * c2.n[2];
        double e = (c1.n[0] * r12[0] + c1.n[1] * r12[1]) + c1.n[2] * r12[2];
        double f = (c2.n[0] * r12[0] + c2.n[1] * r12[1]) + c2.n[2] * r12[2];
        double det = a * c - b * b;
        s = (e * c - b * f) / det;
        s = dmin(dmax(s, -c1.hl), c1.hl);
        t = (b * s - f) / c;
        if (t < -c2.hl || t > c2.hl) {
            t = dmin(dmax(t, -c2.hl), c2.hl);
            s = (e + b * t) / a;
            s = dmin(dmax(s, -c1.hl), c1.hl);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) d[k] = ((c2.r[k] + t * c2.n[k]) - c1.r[k]) - s * c1.n[k];
    }
    s_out = s;
    t_out = t;
}

// Larger root of the JKR quartic (jkr.hpp:67-144), as sqrt(a) = u, the largest
// positive root of u^4 - p u - q (p = Req sqrt(4 pi gamma / E), q = Req delta).
// Newton from an upper bound decreases monotonically onto the root.
__host__ __device__ __forceinline__ double jkr_radius_newton(double delta, double Req, double E,
                                                    double gamma, double u0, int max_iter,
                                                    int* iters)
{
    double p = Req * sqrt(4 * kPi * gamma / E);
    double q = Req * delta;
    double u = u0 > 0 ? u0 : dmax(cbrt(2 * p), sqrt(sqrt(dmax(2 * q, 0.0))));
    int it = 0;
    if (u > 0) {
        for (; it < max_iter; ++it) {
            double u2 = u * u;
            double F = (u2 * u2 - p * u) - q;
            double dF = (4.0 * u2) * u - p;
            if (!(dF > 0)) break;
            double un = u - F / dF;
            if (!(un < u)) break;
            u = un;
        }
    }
    if (iters) *iters = it;
    return u * u;
}

// Force magnitude term1 - term2 of forcesIsotropicJKRLagrange
// (adhesion.hpp:206-212).
__device__ __forceinline__ double jkr_magnitude(double radius, double Req, double E0, double gamma)
{
    double a3 = radius * radius * radius;
    double term1 = 4 * E0 * a3 / (3 * Req);
    double term2 = 4 * sqrt(kPi * a3 * gamma * E0);
    return term1 - term2;
}

// nearestValue, utils.hpp:1169-1215.
__device__ __forceinline__ int nearest_value(const double* v, int n, double x)
{
    if (x <= v[0]) return 0;
    else if (x >= v[n - 1]) return n - 1;
    int low = 0, high = n - 1, nearest = 0;
    while (low <= high) {
        int mid = (low + high) / 2;
        if (v[mid] <= x && x < v[mid + 1]) { nearest = mid; break; }
        else if (x >= v[mid + 1]) low = mid + 1;
        else high = mid - 1;
    }
    if (x - v[nearest] < v[nearest + 1] - x) return nearest;
    return nearest + 1;
}

// 1-D table radius with linear interpolation, adhesion.hpp:312-328.
__device__ __forceinline__ double table_radius_1d(const double* ov, const double* rad, int n,
                                                  double delta)
{
    int i1 = nearest_value(ov, n, delta);
    int i2 = (ov[i1] < delta ? i1 + 1 : i1 - 1);
    if (i1 > i2) { int tmp = i1; i1 = i2; i2 = tmp; }
    double d1 = ov[i1], d2 = ov[i2], r1 = rad[i1], r2 = rad[i2];
    return (r1 * (d2 - delta) + r2 * (delta - d1)) / (d2 - d1);
}

// 2-D table radius with bilinear interpolation, adhesion.hpp:452-490.
__device__ __forceinline__ double table_radius_2d(const double* ov, int nd, const double* gm,
                                                  int ng, const double* rad, double delta,
                                                  double gamma)
{
    int d1i = nearest_value(ov, nd, delta);
    int g1i = nearest_value(gm, ng, gamma);
    int d2i = (ov[d1i] < delta ? d1i + 1 : d1i - 1);
    if (d1i > d2i) { int tmp = d1i; d1i = d2i; d2i = tmp; }
    int g2i = (gm[g1i] < gamma ? g1i + 1 : g1i - 1);
    if (g1i > g2i) { int tmp = g1i; g1i = g2i; g2i = tmp; }
    double d1 = ov[d1i], d2 = ov[d2i], g1 = gm[g1i], g2 = gm[g2i];
    double denom = (d2 - d1) * (g2 - g1);
    double w11 = (d2 - delta) * (g2 - gamma) / denom;
    double w12 = (d2 - delta) * (gamma - g1) / denom;
    double w21 = (delta - d1) * (g2 - gamma) / denom;
    double w22 = (delta - d1) * (gamma - g1) / denom;
    double radius = 0;
    radius += w11 * rad[d1i * ng + g1i];
    radius += w12 * rad[d1i * ng + g2i];
    radius += w21 * rad[d2i * ng + g1i];
    radius += w22 * rad[d2i * ng + g2i];
    return radius;
}

// Anisotropic JKR tables on the device (AnisotropicJKRDataFixedGamma /
// AnisotropicJKRDataVariableGamma, mechanics.hpp:199-327). The seven grids
// lie back to back in `grid` (offsets off[], sizes n[]: theta, half_l,
// centerline coordinate, Rx, phi, overlap, gamma; n[6] == 0 for the fixed-
// gamma table); curv holds Rx at the contact point per (theta, half_l,
// coordinate), fr the (force, radius) pair per (Rx1, Rx2, phi, overlap
// [, gamma]), row-major in that order.
constexpr int kAnisoGrids = 7;
constexpr int kAnisoMaxGrid = 8192;   // grid points in total (staged in LDS)
struct AnisoJKR {
    const double* grid;
    const double* curv;
    const double2* fr;
    int n[kAnisoGrids];
    int off[kAnisoGrids];
    int total;
    double inv_h[kAnisoGrids];   // 1 / spacing of a (nearly) evenly spaced grid, else 0
};

// nearestValue (utils.hpp:1169-1215) on a non-decreasing grid. The
// reference's bisection stops at the interval v[i] <= x < v[i+1], i.e. the
// last i with v[i] <= x; this bisection keeps v[lo] <= x < v[hi] and ends at
// the same i, then takes the nearer end (ties to the upper one). NaN gives 0
// (the reference throws; the non-finite forces that follow are reported).
__device__ __forceinline__ int nearest_sorted(const double* v, int n, double x)
{
    if (!(x > v[0])) return 0;
    if (x >= v[n - 1]) return n - 1;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] <= x) lo = mid;
        else hi = mid;
    }
    return (x - v[lo] < v[hi] - x) ? lo : hi;
}

// nearest_sorted with a first guess for evenly spaced grids (inv_h > 0): the
// guess is moved to the same interval v[lo] <= x < v[lo + 1] by comparisons
// with the grid values, so the result is nearest_sorted's for any sorted
// grid; the spacing only decides how many steps that takes.
__device__ __forceinline__ int nearest_grid(const double* v, int n, double x, double inv_h)
{
    if (inv_h <= 0) return nearest_sorted(v, n, x);
    if (!(x > v[0])) return 0;
    if (x >= v[n - 1]) return n - 1;
    int lo = (int)((x - v[0]) * inv_h);
    lo = min(max(lo, 0), n - 2);
    while (lo > 0 && v[lo] > x) --lo;
    while (lo < n - 2 && v[lo + 1] <= x) ++lo;
    return (x - v[lo] < v[lo + 1] - x) ? lo : lo + 1;
}

// acosSafe, utils.hpp:629-638.
__device__ __forceinline__ double acos_safe(double x)
{
    if (x >= 1) return 0.0;
    if (x <= -1) return kPi;
    return acos(x);
}

// Eigen's 3-vector dot, (a0 b0 + a1 b1) + a2 b2, without contraction into
// FMAs: the angles index tables, so they are formed as the CPU forms them.
__device__ __forceinline__ double dot3_rn(const double* a, const double* b)
{
    return __dadd_rn(__dadd_rn(__dmul_rn(a[0], b[0]), __dmul_rn(a[1], b[1])), __dmul_rn(a[2], b[2]));
}

// Table lookup of forcesAnisotropicJKRLagrange (adhesion.hpp:734-840 fixed
// gamma, :842-1046 variable gamma): the angles between the overlap vector and
// each orientation folded into [0, pi/2], the principal radius Rx at each
// contact point, the pair ordered Rx1 <= Rx2, the angle between the cells
// folded into [0, pi/2], then (force, radius) at the nearest grid points.
// g: the grids (an LDS copy of AnisoJKR::grid); delta: the capped overlap.
__device__ __forceinline__ double2 aniso_jkr_lookup(const AnisoJKR& an, const double* g, const double d[3],
                                                    double dist, const double n1[3], double hl1,
                                                    const double n2[3], double hl2, double s, double t,
                                                    double delta, double gamma)
{
    constexpr double kHalfPi = 1.57079632679489661923;
    const double dn[3] = {d[0] / dist, d[1] / dist, d[2] / dist};
    // The reference folds theta1 = acos(c1) into [0, pi/2] by taking acos(-c1)
    // when acos(c1) > pi/2, i.e. when c1 < 0, and theta2 likewise from
    // (-d12n).n2 = -c2 (sign flips are exact): both are acos(|c|). One
    // arccosine per angle instead of two; the two-step form differs only when
    // acos of a |c| below an ulp rounds to pi/2, a difference under one ulp.
    const double c1 = dot3_rn(dn, n1), c2 = dot3_rn(dn, n2);
    const double th1 = acos_safe(fabs(c1));
    const double th2 = acos_safe(fabs(c2));
    const double* g_th = g + an.off[0];
    const double* g_hl = g + an.off[1];
    const double* g_co = g + an.off[2];
    const int it1 = nearest_grid(g_th, an.n[0], th1, an.inv_h[0]), it2 = nearest_grid(g_th, an.n[0], th2, an.inv_h[0]);
    const int ih1 = nearest_grid(g_hl, an.n[1], hl1, an.inv_h[1]), ih2 = nearest_grid(g_hl, an.n[1], hl2, an.inv_h[1]);
    const int ic1 = nearest_grid(g_co, an.n[2], fabs(s) / hl1, an.inv_h[2]);
    const int ic2 = nearest_grid(g_co, an.n[2], fabs(t) / hl2, an.inv_h[2]);
    double rx1 = an.curv[((int64_t)it1 * an.n[1] + ih1) * an.n[2] + ic1];
    double rx2 = an.curv[((int64_t)it2 * an.n[1] + ih2) * an.n[2] + ic2];
    if (rx1 > rx2) { const double tmp = rx1; rx1 = rx2; rx2 = tmp; }
    double phi = acos_safe(dot3_rn(n1, n2));
    if (phi > kHalfPi) phi = kPi - phi;
    const int ir1 = nearest_grid(g + an.off[3], an.n[3], rx1, an.inv_h[3]);
    const int ir2 = nearest_grid(g + an.off[3], an.n[3], rx2, an.inv_h[3]);
    const int ip = nearest_grid(g + an.off[4], an.n[4], phi, an.inv_h[4]);
    const int id = nearest_grid(g + an.off[5], an.n[5], delta, an.inv_h[5]);
    const int ng = an.n[6] > 0 ? an.n[6] : 1;
    const int ig = an.n[6] > 0 ? nearest_grid(g + an.off[6], an.n[6], gamma, an.inv_h[6]) : 0;
    const int64_t q = ((((int64_t)ir1 * an.n[3] + ir2) * an.n[4] + ip) * an.n[5] + id) * ng + ig;
    return an.fr[q];
}

// ---- surface integrals, integrals.hpp -------------------------------------
__device__ __forceinline__ double phi_(double rz, double nz, double R, double s)
{
    return R - rz - s * nz;
}

// integral1 (integrals.hpp:116-144) with exponent g = E2/2 - 1.
template <int E2>
__device__ __forceinline__ double integral1(double rz, double nz, double R, double hl, double ss)
{
    constexpr double gp1 = E2 / 2.0;
    if (ss > hl) {
        double o1 = pow_half<E2>(phi_(rz, nz, R, -hl));
        double o2 = pow_half<E2>(phi_(rz, nz, R, hl));
        return (o1 - o2) / (nz * gp1);
    } else if (ss > -hl) {
        double o1 = pow_half<E2>(phi_(rz, nz, R, -hl));
        return o1 / (nz * gp1);
    }
    return 0;
}

// integral2 (integrals.hpp:158-184) with exponent g = E2/2 - 1.
template <int E2>
__device__ __forceinline__ double integral2(double rz, double nz, double R, double hl, double ss)
{
    constexpr double gp1 = E2 / 2.0;
    double term1 = integral1<E2 + 2>(rz, nz, R, hl, ss);
    double term2 = 0;
    if (ss > hl) {
        double o1 = pow_half<E2>(phi_(rz, nz, R, -hl));
        double o2 = pow_half<E2>(phi_(rz, nz, R, hl));
        term2 = -hl * (o1 + o2);
    } else if (ss > -hl) {
        double o1 = pow_half<E2>(phi_(rz, nz, R, -hl));
        term2 = -hl * o1;
    }
    return (term1 + term2) / (nz * gp1);
}

__device__ __forceinline__ double integral4(double hl, double ss)
{
    if (ss > hl) return 2 * hl;
    else if (ss > -hl) return ss + hl;
    return 0;
}
__device__ __forceinline__ double integral5(double hl, double ss)
{
    if (-hl < ss && ss <= hl) return (ss * ss - hl * hl) / 2.0;
    return 0.0;
}
__device__ __forceinline__ double integral6(double hl, double ss)
{
    if (ss > hl) return 2 * hl * hl * hl / 3.0;
    else if (ss > -hl) return (ss * ss * ss + hl * hl * hl) / 3.0;
    return 0.0;
}

// areaIntegrals, integrals.hpp:400-463.
__device__ __forceinline__ void area_integrals(double rz, double nz, double R, double hl, double ss,
                                               double sqR, double out[3])
{
    double int4 = integral4(hl, ss);
    double int5 = integral5(hl, ss);
    double int6 = integral6(hl, ss);
    double int1a = integral1<3>(rz, nz, R, hl, ss);
    double int1b = integral1<5>(rz, nz, R, hl, ss);
    double int1c = integral1<7>(rz, nz, R, hl, ss);
    double term1 = sqR * (1 - nz * nz) * int1a;
    double term2 = kPi * R * nz * nz * int4;
    double area1 = term1 + term2;
    double overlap1 = 0, overlap2 = 0;
    if (ss > hl) {
        overlap1 = phi_(rz, nz, R, -hl);
        overlap2 = phi_(rz, nz, R, hl);
    } else if (ss > -hl) {
        overlap1 = phi_(rz, nz, R, -hl);
    }
    double o1p1 = pow_half<3>(overlap1), o2p1 = pow_half<3>(overlap2);
    term2 = -hl * (o1p1 + o2p1);
    double int2a = (int1b + term2) / (1.5 * nz);
    term1 = sqR * (1 - nz * nz) * int2a;
    term2 = kPi * R * nz * nz * int5;
    double area2 = term1 + term2;
    double o1p2 = o1p1 * overlap1, o2p2 = o2p1 * overlap2;
    term2 = -hl * (o1p2 + o2p2);
    double int2b = (int1c + term2) / (2.5 * nz);
    term1 = 2 * int2b;
    term2 = hl * hl * (o1p1 - o2p1);
    double int3 = (term1 + term2) / (1.5 * nz);
    term1 = sqR * (1 - nz * nz) * int3;
    term2 = kPi * R * nz * nz * int6;
    double area3 = term1 + term2;
    out[0] = area1; out[1] = area2; out[2] = area3;
}

struct SurfaceConsts {
    double R, E0, sqR;
    double rep0, rep1, rep2;   // 2 E0, (8/3) E0 sqrt(R), 2 E0 sqrt(R)   (mechanics.hpp:687-689)
    double adh0, adh1, adh2;   // sqrt(R)/2, 2 pi R, 2 sqrt(R)           (mechanics.hpp:812-814)
};

// cellSurfaceRepulsionForces (mechanics.hpp:673-780) and
// cellSurfaceAdhesionForces (mechanics.hpp:800-910) for one cell,
// include_constraint = false: the dE/drz and dE/dnz entries.
__device__ __forceinline__ void surface_forces(const SurfaceConsts& k, bool assume_2d, double rz,
                                               double nz, double hl, double l, double sigma0,
                                               double ss, double& rep2, double& rep5,
                                               double& adh2, double& adh5)
{
    const double R = k.R;
    rep2 = 0; rep5 = 0; adh2 = 0; adh5 = 0;
    if (assume_2d) {
        double ph = R - rz;
        if (ph > 0) {
            rep2 = -k.rep0 * ph * l;
            adh2 = sigma0 * k.adh0 * l / sqrt(ph);
        }
        return;
    }
    double nz2 = nz * nz;
    // repulsion
    double int1 = integral1<4>(rz, nz, R, hl, ss);
    double int2 = integral1<3>(rz, nz, R, hl, ss);
    rep2 = -k.rep0 * ((1.0 - nz2) * int1 + k.sqR * nz2 * int2);
    double int3 = integral1<6>(rz, nz, R, hl, ss);
    double int4 = integral2<4>(rz, nz, R, hl, ss);
    double int5 = integral1<5>(rz, nz, R, hl, ss);
    double int6 = integral2<3>(rz, nz, R, hl, ss);
    double f5 = 0;
    f5 -= k.rep0 * nz * int3;
    f5 -= k.rep0 * (1 - nz2) * int4;
    f5 += k.rep1 * nz * int5;
    f5 -= k.rep2 * nz2 * int6;
    rep5 = f5;
    // adhesion
    double a1 = integral1<1>(rz, nz, R, hl, ss);
    double term2 = 0.0;
    if (fabs(ss) < hl) term2 = (k.adh1 / 2) * nz;
    double g2 = k.adh0 * (1 - nz2) * a1 + term2;
    g2 *= sigma0;
    adh2 = g2;
    double a2 = int2;                                   // integral1, exponent 0.5
    double a3 = integral2<1>(rz, nz, R, hl, ss);
    double a4 = integral4(hl, ss);
    double term4 = 0.0;
    if (fabs(ss) < hl) term4 = (k.adh1 / 2) * (R - rz);
    double g5 = 0;
    g5 += k.adh2 * nz * a2;
    g5 += k.adh0 * (1 - nz2) * a3;
    g5 -= k.adh1 * nz * a4;
    g5 += term4;
    g5 *= sigma0;
    adh5 = g5;
}

// cellSurfaceRepulsionForces / cellSurfaceAdhesionForces with
// include_constraint = true (mechanics.hpp:741-754, 868-884): dE/drz and the
// constrained torques about (n_y, -n_x, 0), entries 2, 3, 4 of rep[] / adh[].
__device__ __forceinline__ void surface_forces_ic(const SurfaceConsts& k, bool assume_2d, double rz, double nx,
                                                  double ny, double nz, double hl, double l, double sigma0,
                                                  double ss, double rep[6], double adh[6])
{
    const double R = k.R;
    if (assume_2d) {
        double ph = R - rz;
        if (ph > 0) {
            rep[2] = -k.rep0 * ph * l;
            adh[2] = sigma0 * k.adh0 * l / sqrt(ph);
        }
        return;
    }
    const double nz2 = nz * nz;
    const double cx = ny, cy = -nx;
    const double int1 = integral1<4>(rz, nz, R, hl, ss);
    const double int2 = integral1<3>(rz, nz, R, hl, ss);
    rep[2] = -k.rep0 * ((1.0 - nz2) * int1 + k.sqR * nz2 * int2);
    const double i3 = integral2<4>(rz, nz, R, hl, ss);
    const double i4 = integral2<3>(rz, nz, R, hl, ss);
    rep[3] = -k.rep0 * cx * ((1.0 - nz2) * i3 + k.sqR * nz2 * i4);
    rep[4] = -k.rep0 * cy * ((1.0 - nz2) * i3 + k.sqR * nz2 * i4);
    const double a1 = integral1<1>(rz, nz, R, hl, ss);
    double term2 = 0.0;
    if (fabs(ss) < hl) term2 = (k.adh1 / 2) * nz;
    double g2 = k.adh0 * (1 - nz2) * a1 + term2;
    g2 *= sigma0;
    adh[2] = g2;
    const double a2 = integral2<1>(rz, nz, R, hl, ss);
    double w1 = 0, w2 = 0;
    if (fabs(ss) < hl) {
        w1 = (k.adh1 / 2) * ss * nz * cx;
        w2 = (k.adh1 / 2) * ss * nz * cy;
    }
    double g3 = k.adh0 * (1 - nz2) * a2 * cx - w1;
    g3 *= sigma0;
    double g4 = k.adh0 * (1 - nz2) * a2 * cy - w2;
    g4 *= sigma0;
    adh[3] = g3;
    adh[4] = g4;
}

struct CellVel {
    double rz, nx, ny, nz, l, hl, eta0, eta1;
    double drx, dry, dnx, dny;   // stored velocities (Coulomb truncation)
};

// Per-cell velocity solve of getVelocities, mechanics.hpp:2069-2344.
__device__ __forceinline__ void velocity_solve(const CellVel& c, bool assume_2d, double ss,
                                               const SurfaceConsts& k, double coeff,
                                               bool no_surface, const double f[6], double v[6])
{
    const double R = k.R, E0 = k.E0;
    const double rz = c.rz, nx = c.nx, ny = c.ny, nz = c.nz, l = c.l, hl = c.hl;
    const double eta0 = c.eta0;
    bool contacting;
    if (!assume_2d) {
        double z1 = rz - hl * nz, z2 = rz + hl * nz;
        contacting = (z1 < R || z2 < R);
    } else {
        contacting = (rz < R);
    }
    if (no_surface || !contacting) {
        double K = eta0 * l;
        double L = K * l * l / 12;
        double mult = -(nx * f[3] + ny * f[4] + nz * f[5]);
        v[0] = f[0] / K; v[1] = f[1] / K; v[2] = f[2] / K;
        v[3] = (f[3] + mult * nx) / L;
        v[4] = (f[4] + mult * ny) / L;
        v[5] = (f[5] + mult * nz) / L;
    } else if (assume_2d) {
        double K, L;
        double eta1 = c.eta1;
        if (coeff == 0) {
            double a = sqrt(R * (R - rz));
            double eta_combined = eta0 + eta1 * a / R;
            K = eta_combined * l;
            L = K * l * l / 12;
        } else {
            double K_vis = eta0 * l;
            double L_vis = K_vis * l * l / 12;
            double a = sqrt(R * (R - rz));
            double K_fric = eta1 * l * a / R;
            double drxy = sqrt(c.drx * c.drx + c.dry * c.dry);
            double friction_norm = K_fric * drxy;
            double max_norm = coeff * (2 * E0 * (R - rz) * l);
            if (friction_norm > max_norm) {
                double eta1_truncated = max_norm * R / (drxy * l * a);
                K_fric = eta1_truncated * l * a / R;
            }
            double L_fric = K_fric * l * l / 12;
            K = K_vis + K_fric;
            L = L_vis + L_fric;
        }
        double mult = -(nx * f[3] + ny * f[4]);
        v[0] = f[0] / K;
        v[1] = f[1] / K;
        v[3] = (f[3] + mult * nx) / L;
        v[4] = (f[4] + mult * ny) / L;
        v[2] = f[2] / (eta0 * l);
        v[5] = 12 * f[5] / (eta0 * l * l * l);
    } else {
        double eta1 = c.eta1;
        double ai[3];
        area_integrals(rz, nz, R, hl, ss, k.sqR, ai);
        if (coeff > 0) {
            double drxy = sqrt(c.drx * c.drx + c.dry * c.dry);
            double dnxy = sqrt(c.dnx * c.dnx + c.dny * c.dny);
            double friction_norm = ai[0] * drxy;
            friction_norm += ai[1] * dnxy;
            friction_norm *= (eta1 / R);
            double max_norm = (1 - nz * nz) * integral1<4>(rz, nz, R, hl, ss);
            max_norm += k.sqR * nz * nz * integral1<3>(rz, nz, R, hl, ss);
            max_norm *= (2 * E0 * coeff);
            if (friction_norm > max_norm)
                eta1 = max_norm * R / (ai[0] * drxy + ai[1] * dnxy);
        }
        const double c1 = eta0 * l + (eta1 / R) * ai[0];
        const double c2 = eta0 * l;
        const double c3 = c2 * l * l / 12 + (eta1 / R) * ai[2];
        const double c4 = c2 * l * l / 12;
        const double c5 = (eta1 / R) * ai[1];
        const double term1 = c3 - (c5 * c5 / c1);
        const double w4 = f[3] - f[0] * (c5 / c1);
        const double w5 = f[4] - f[1] * (c5 / c1);
        const double nx2 = nx / term1, ny2 = ny / term1, nz2 = nz / c4;
        const double alpha = 2 * (nx * nx2 + ny * ny2 + nz * nz2);
        const double beta = -(w4 * nx2 + w5 * ny2 + f[5] * nz2);
        const double lambda = beta / alpha;
        v[5] = (f[5] + 2 * nz * lambda) / c4;
        v[4] = (w5 + 2 * ny * lambda) / term1;
        v[3] = (w4 + 2 * nx * lambda) / term1;
        v[2] = f[2] / c2;
        v[1] = (f[1] - c5 * v[4]) / c1;
        v[0] = (f[0] - c5 * v[3]) / c1;
    }
}

// The packed 64-byte segment {r, n, half_l, 0} of cell c (k_pack_segs layout).
__device__ __forceinline__ Seg load_packed_seg(const double* __restrict__ geom, int c)
{
    const double2* g = reinterpret_cast<const double2*>(geom + (int64_t)c * 8);
    const double2 a = g[0], b = g[1], e = g[2], f = g[3];
    Seg s;
    s.r[0] = a.x; s.r[1] = a.y; s.r[2] = b.x;
    s.n[0] = b.y; s.n[1] = e.x; s.n[2] = e.y;
    s.hl = f.x;
    return s;
}

// updateNeighborDistances (mechanics.hpp:604-650) fused into the first
// kernel that reads the list (the pair forces, or the learned radius's
// staging): each row's distance vector and contact coordinates recomputed
// from the cells' packed segments and written back to the list.
// geom == nullptr: the list's stored values are used.
struct DistRefresh {
    const double* geom;
    int n;                  // rows naming a cell >= n keep their stored values
    double *dx, *dy, *dz, *s, *t;
    const int* gate;        // non-null: the kernel runs only if *gate != 0
};

// Row k's distance vector d and contact coordinates (s, t): refreshed from
// the packed segments (and written back) or read from the list.
__device__ __forceinline__ void row_geometry(const DistRefresh& rf, int k, int i, int j,
                                             const double* __restrict__ pdx, const double* __restrict__ pdy,
                                             const double* __restrict__ pdz, const double* __restrict__ ps,
                                             const double* __restrict__ pt, double d[3], double& sv, double& tv)
{
    if (rf.geom && (unsigned)i < (unsigned)rf.n && (unsigned)j < (unsigned)rf.n) {
        const Seg ci = load_packed_seg(rf.geom, i);
        const Seg cj = load_packed_seg(rf.geom, j);
        dist_between_cells(ci, cj, d, sv, tv);
        rf.dx[k] = d[0]; rf.dy[k] = d[1]; rf.dz[k] = d[2]; rf.s[k] = sv; rf.t[k] = tv;
    } else {
        d[0] = pdx[k]; d[1] = pdy[k]; d[2] = pdz[k]; sv = ps ? ps[k] : 0.0; tv = pt ? pt[k] : 0.0;
    }
}

}  // namespace bfm

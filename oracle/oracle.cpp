/*
 * oracle.cpp -- CPU oracle of the mechanical step (TEST INFRASTRUCTURE ONLY).
 *
 * Restates the reference algorithms of ABM-3d/ABM-3d-cpp/include (paths below
 * are relative to that directory) without Eigen, Boost or CGAL. It is the
 * checker of the HIP path and the CPU baseline of bench.py; the product never
 * links it. Build: oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off).
 *
 * Floating-point conventions: every expression is written with the
 * evaluation order of the reference expression it restates (C++ evaluates
 * a * b * c as (a * b) * c), and the build disables FMA contraction, so the
 * rounding of each step is that of the plain IEEE operation.
 *
 * Deliberate differences from the reference:
 *  - distBetweenCells, non-parallel branch: CGAL's segment-segment routine is
 *    replaced by the clamped closed-form minimiser below (the minimiser of a
 *    strictly convex quadratic is unique, so results agree to rounding).
 *  - jkrContactRadius: the 30/100-digit Aberth solve is replaced by a
 *    monotone FP64 Newton iteration on u = sqrt(a) (pinned against mpmath).
 *  - Eigen's norm summation order is taken as ((x*x + y*y) + z*z).
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>
#include <omp.h>

namespace {

const double kPi = 3.14159265358979323846;

// Column indices, include/indices.hpp:17-47.
enum {
    C_ID = 0, C_RX = 1, C_RY = 2, C_RZ = 3, C_NX = 4, C_NY = 5, C_NZ = 6,
    C_DRX = 7, C_DRY = 8, C_DRZ = 9, C_DNX = 10, C_DNY = 11, C_DNZ = 12,
    C_L = 13, C_HALF_L = 14, C_T0 = 15, C_GROWTH = 16, C_ETA0 = 17,
    C_ETA1 = 18, C_SIGMA0 = 19, C_GROUP = 20
};

inline double cell(const double* cells, int ld, int i, int c) { return cells[(int64_t)c * ld + i]; }

inline double norm3(double x, double y, double z) { return std::sqrt((x * x + y * y) + z * z); }

// JKR table registered by oracle_set_jkr_table (the JKRData of the reference).
std::vector<double> g_tab_overlaps, g_tab_gammas, g_tab_radii;

// Anisotropic JKR tables registered by oracle_set_aniso_tables
// (AnisotropicJKRDataFixedGamma / VariableGamma, mechanics.hpp:199-327).
struct AnisoTab {
    std::vector<double> theta, half_l, coords, curv, rx, phi, delta, gamma, fr;
} g_an;

}  // namespace

extern "C" {

/* nearestCellBodyCoordToPoint, distances.hpp:92-105. */
double oracle_nearest_cell_body_coord(const double* r, const double* n, double half_l,
                                      const double* q)
{
    double s = ((q[0] - r[0]) * n[0] + (q[1] - r[1]) * n[1]) + (q[2] - r[2]) * n[2];
    if (std::fabs(s) <= half_l) return s;
    else if (s > half_l) return half_l;
    else return -half_l;
}

/* distBetweenCells, distances.hpp:206-357. d runs from cell 1 to cell 2. */
void oracle_dist_between_cells(const double* r1, const double* n1, double half_l1,
                               const double* r2, const double* n2, double half_l2,
                               double* d, double* s_out, double* t_out)
{
    double s = 0, t = 0;
    double cos_theta = (n1[0] * n2[0] + n1[1] * n2[1]) + n1[2] * n2[2];
    if (cos_theta >= 0.9999 || cos_theta <= -0.9999) {
        // Near-parallel branch, distances.hpp:228-337.
        double p1[3], q1[3], p2[3], q2[3];
        for (int k = 0; k < 3; ++k) {
            p1[k] = r1[k] - half_l1 * n1[k];
            q1[k] = r1[k] + half_l1 * n1[k];
            p2[k] = r2[k] - half_l2 * n2[k];
            q2[k] = r2[k] + half_l2 * n2[k];
        }
        double s_p1 = oracle_nearest_cell_body_coord(r2, n2, half_l2, p1);
        double s_q1 = oracle_nearest_cell_body_coord(r2, n2, half_l2, q1);
        double dp1[3], dq1[3];
        for (int k = 0; k < 3; ++k) {
            dp1[k] = (r2[k] + s_p1 * n2[k]) - p1[k];
            dq1[k] = (r2[k] + s_q1 * n2[k]) - q1[k];
        }
        double dist_p1 = norm3(dp1[0], dp1[1], dp1[2]);
        double dist_q1 = norm3(dq1[0], dq1[1], dq1[2]);
        if (s_p1 == s_q1) {
            if (dist_p1 < dist_q1) {
                s = -half_l1; t = s_p1;
                d[0] = dp1[0]; d[1] = dp1[1]; d[2] = dp1[2];
            } else {
                s = half_l1; t = s_q1;
                d[0] = dq1[0]; d[1] = dq1[1]; d[2] = dq1[2];
            }
        } else {
            double s_p2 = oracle_nearest_cell_body_coord(r1, n1, half_l1, p2);
            double s_q2 = oracle_nearest_cell_body_coord(r1, n1, half_l1, q2);
            double dp2[3], dq2[3];
            for (int k = 0; k < 3; ++k) {
                dp2[k] = (r1[k] + s_p2 * n1[k]) - p2[k];
                dq2[k] = (r1[k] + s_q2 * n1[k]) - q2[k];
            }
            double dists[4] = {dist_p1, dist_q1, norm3(dp2[0], dp2[1], dp2[2]),
                               norm3(dq2[0], dq2[1], dq2[2])};
            // std::sort of 4 (distance, index) pairs keyed on distance
            // (distances.hpp:274-286); libstdc++ sorts n < 16 by insertion
            // sort, which is stable, so ties keep index order.
            int idx[4] = {0, 1, 2, 3};
            for (int a = 1; a < 4; ++a) {
                int v = idx[a];
                int b = a - 1;
                while (b >= 0 && dists[v] < dists[idx[b]]) { idx[b + 1] = idx[b]; --b; }
                idx[b + 1] = v;
            }
            bool has[4] = {false, false, false, false};
            has[idx[0]] = true; has[idx[1]] = true;
            if (has[0]) {
                if (has[1]) {
                    s = 0.0;
                    t = oracle_nearest_cell_body_coord(r2, n2, half_l2, r1);
                    for (int k = 0; k < 3; ++k) d[k] = (r2[k] + t * n2[k]) - r1[k];
                } else {
                    s = has[2] ? (-half_l1 + s_p2) / 2 : (-half_l1 + s_q2) / 2;
                    double c[3] = {r1[0] + s * n1[0], r1[1] + s * n1[1], r1[2] + s * n1[2]};
                    t = oracle_nearest_cell_body_coord(r2, n2, half_l2, c);
                    for (int k = 0; k < 3; ++k) d[k] = ((r2[k] + t * n2[k]) - r1[k]) - s * n1[k];
                }
            } else if (has[1]) {
                s = has[2] ? (half_l1 + s_p2) / 2 : (half_l1 + s_q2) / 2;
                double c[3] = {r1[0] + s * n1[0], r1[1] + s * n1[1], r1[2] + s * n1[2]};
                t = oracle_nearest_cell_body_coord(r2, n2, half_l2, c);
                for (int k = 0; k < 3; ++k) d[k] = ((r2[k] + t * n2[k]) - r1[k]) - s * n1[k];
            } else {
                t = 0.0;
                s = oracle_nearest_cell_body_coord(r1, n1, half_l1, r2);
                for (int k = 0; k < 3; ++k) d[k] = (r2[k] - r1[k]) - s * n1[k];
            }
        }
    } else {
        // Skew branch (distances.hpp:338-345): closest points of the two
        // centerline segments. Minimise |r12 + t n2 - s n1|^2 over the box
        // [-hl1, hl1] x [-hl2, hl2]: solve the 2x2 normal equations, clamp s,
        // take the optimal t for that s, and if t had to be clamped take the
        // optimal s for the clamped t.
        double r12[3] = {r2[0] - r1[0], r2[1] - r1[1], r2[2] - r1[2]};
        double a = (n1[0] * n1[0] + n1[1] * n1[1]) + n1[2] * n1[2];
        double b = (n1[0] * n2[0] + n1[1] * n2[1]) + n1[2] * n2[2];
        double c = (n2[0] * n2[0] + n2[1] * n2[1]) + n2[2] * n2[2];
        double e = (n1[0] * r12[0] + n1[1] * r12[1]) + n1[2] * r12[2];
        double f = (n2[0] * r12[0] + n2[1] * r12[1]) + n2[2] * r12[2];
        double det = a * c - b * b;
        s = (e * c - b * f) / det;
        s = std::min(std::max(s, -half_l1), half_l1);
        t = (b * s - f) / c;
        if (t < -half_l2 || t > half_l2) {
            t = std::min(std::max(t, -half_l2), half_l2);
            s = (e + b * t) / a;
            s = std::min(std::max(s, -half_l1), half_l1);
        }
        for (int k = 0; k < 3; ++k) d[k] = ((r2[k] + t * n2[k]) - r1[k]) - s * n1[k];
    }
    *s_out = s;
    *t_out = t;
}

static inline void cell_geom(const double* cells, int ld, int i, double* r, double* n, double* hl)
{
    r[0] = cell(cells, ld, i, C_RX); r[1] = cell(cells, ld, i, C_RY); r[2] = cell(cells, ld, i, C_RZ);
    n[0] = cell(cells, ld, i, C_NX); n[1] = cell(cells, ld, i, C_NY); n[2] = cell(cells, ld, i, C_NZ);
    *hl = cell(cells, ld, i, C_HALF_L);
}

/* getCellNeighbors, mechanics.hpp:520-593: serial all-pairs loop over j < i
 * with the center-distance prefilter (:559) and |d| < threshold (:580). */
int64_t oracle_get_cell_neighbors(const double* cells, int n, int ld, double thr,
                                  double R, double Ldiv, double* out, int64_t ld_out)
{
    if (n <= 1) return 0;
    const double pre = thr + Ldiv + 2 * R;
    int64_t idx = 0;
    for (int i = 0; i < n; ++i) {
        double ri[3], ni[3], hli;
        cell_geom(cells, ld, i, ri, ni, &hli);
        for (int j = 0; j < i; ++j) {
            double rj[3], nj[3], hlj;
            cell_geom(cells, ld, j, rj, nj, &hlj);
            double dist_rij = norm3(ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]);
            if (dist_rij < pre) {
                double d[3], si, sj;
                oracle_dist_between_cells(ri, ni, hli, rj, nj, hlj, d, &si, &sj);
                if (norm3(d[0], d[1], d[2]) < thr) {
                    if (idx < ld_out) {
                        out[0 * ld_out + idx] = i; out[1 * ld_out + idx] = j;
                        out[2 * ld_out + idx] = d[0]; out[3 * ld_out + idx] = d[1];
                        out[4 * ld_out + idx] = d[2]; out[5 * ld_out + idx] = si;
                        out[6 * ld_out + idx] = sj;
                    }
                    idx++;
                }
            }
        }
    }
    return idx;
}

/* Rows [row_lo, row_hi) of oracle_get_cell_neighbors (same loop body),
 * counting kept pairs only: a bounded sample of the O(N^2) search for CPU
 * timing. Row i costs O(i), so a block centred on N/2 has the mean cost. */
int64_t oracle_count_neighbor_rows(const double* cells, int n, int ld, double thr, double R,
                                   double Ldiv, int row_lo, int row_hi)
{
    const double pre = thr + Ldiv + 2 * R;
    int64_t idx = 0;
    for (int i = std::max(row_lo, 0); i < std::min(row_hi, n); ++i) {
        double ri[3], ni[3], hli;
        cell_geom(cells, ld, i, ri, ni, &hli);
        for (int j = 0; j < i; ++j) {
            double rj[3], nj[3], hlj;
            cell_geom(cells, ld, j, rj, nj, &hlj);
            double dist_rij = norm3(ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]);
            if (dist_rij < pre) {
                double d[3], si, sj;
                oracle_dist_between_cells(ri, ni, hli, rj, nj, hlj, d, &si, &sj);
                if (norm3(d[0], d[1], d[2]) < thr) idx++;
            }
        }
    }
    return idx;
}

/* Same list as oracle_get_cell_neighbors, found with a uniform grid and
 * OpenMP (context timing only; not the reference's algorithm). A candidate is
 * dropped without the segment test only when |r_i - r_j| exceeds
 * thr + hl_i + hl_j by a margin, which bounds |d| >= thr for any pair of
 * centerline points, so the list is identical. */
int64_t oracle_get_cell_neighbors_binned(const double* cells, int n, int ld, double thr,
                                         double R, double Ldiv, double* out,
                                         int64_t ld_out, int nthreads)
{
    if (n <= 1) return 0;
    const double pre = thr + Ldiv + 2 * R;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, hlmax = 0;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            double v = cell(cells, ld, i, C_RX + k);
            lo[k] = std::min(lo[k], v); hi[k] = std::max(hi[k], v);
        }
        hlmax = std::max(hlmax, cell(cells, ld, i, C_HALF_L));
    }
    double cut = std::min(pre, thr + 2 * hlmax) * (1 + 1e-9) + 1e-9;
    int dims[3];
    double edge = cut;
    for (;;) {
        int64_t tot = 1;
        for (int k = 0; k < 3; ++k) { dims[k] = (int)((hi[k] - lo[k]) / edge) + 1; tot *= dims[k]; }
        if (tot <= 4 * (int64_t)n + 64) break;
        edge *= 1.25;
    }
    int64_t nb = (int64_t)dims[0] * dims[1] * dims[2];
    std::vector<int> key(n), start(nb + 1, 0), order(n);
    for (int i = 0; i < n; ++i) {
        int b[3];
        for (int k = 0; k < 3; ++k)
            b[k] = std::min(dims[k] - 1, (int)((cell(cells, ld, i, C_RX + k) - lo[k]) / edge));
        key[i] = (b[2] * dims[1] + b[1]) * dims[0] + b[0];
        start[key[i] + 1]++;
    }
    for (int64_t b = 0; b < nb; ++b) start[b + 1] += start[b];
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < n; ++i) order[fill[key[i]]++] = i;
    }
    std::vector<std::vector<double>> rows(n);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        double ri[3], ni[3], hli;
        cell_geom(cells, ld, i, ri, ni, &hli);
        int b0 = key[i] % dims[0], b1 = (key[i] / dims[0]) % dims[1], b2 = key[i] / (dims[0] * dims[1]);
        std::vector<std::pair<int, int>> cand;
        for (int z = std::max(0, b2 - 1); z <= std::min(dims[2] - 1, b2 + 1); ++z)
            for (int y = std::max(0, b1 - 1); y <= std::min(dims[1] - 1, b1 + 1); ++y)
                for (int x = std::max(0, b0 - 1); x <= std::min(dims[0] - 1, b0 + 1); ++x) {
                    int bb = (z * dims[1] + y) * dims[0] + x;
                    for (int q = start[bb]; q < start[bb + 1]; ++q) {
                        int j = order[q];
                        if (j < i) cand.push_back({j, 0});
                    }
                }
        std::sort(cand.begin(), cand.end());
        std::vector<double>& mine = rows[i];
        for (auto& cj : cand) {
            int j = cj.first;
            double rj[3], nj[3], hlj;
            cell_geom(cells, ld, j, rj, nj, &hlj);
            double dx = ri[0] - rj[0], dy = ri[1] - rj[1], dz = ri[2] - rj[2];
            double r2 = (dx * dx + dy * dy) + dz * dz;
            double bound = (thr + hli + hlj) * (1 + 1e-9) + 1e-9;
            if (r2 > bound * bound) continue;
            if (!(std::sqrt(r2) < pre)) continue;
            double d[3], si, sj;
            oracle_dist_between_cells(ri, ni, hli, rj, nj, hlj, d, &si, &sj);
            if (norm3(d[0], d[1], d[2]) < thr) {
                double row[6] = {(double)j, d[0], d[1], d[2], si, sj};
                mine.insert(mine.end(), row, row + 6);
            }
        }
    }
    int64_t idx = 0;
    for (int i = 0; i < n; ++i) {
        for (size_t q = 0; q < rows[i].size(); q += 6) {
            if (idx < ld_out) {
                out[idx] = i;
                for (int c = 0; c < 6; ++c) out[(c + 1) * ld_out + idx] = rows[i][q + c];
            }
            idx++;
        }
    }
    return idx;
}

/* updateNeighborDistances, mechanics.hpp:604-648. */
void oracle_update_neighbor_distances(const double* cells, int n, int ld,
                                      double* nb, int64_t m, int64_t ld_nb)
{
    (void)n;
    for (int64_t k = 0; k < m; ++k) {
        int i = (int)nb[k], j = (int)nb[ld_nb + k];
        double ri[3], ni[3], hli, rj[3], nj[3], hlj, d[3], si, sj;
        cell_geom(cells, ld, i, ri, ni, &hli);
        cell_geom(cells, ld, j, rj, nj, &hlj);
        oracle_dist_between_cells(ri, ni, hli, rj, nj, hlj, d, &si, &sj);
        nb[2 * ld_nb + k] = d[0]; nb[3 * ld_nb + k] = d[1]; nb[4 * ld_nb + k] = d[2];
        nb[5 * ld_nb + k] = si; nb[6 * ld_nb + k] = sj;
    }
}

/* jkrContactRadius (jkr.hpp:67-144) returns the larger real root of
 * a^4 - 2 R delta a^2 - (4 pi R^2 gamma / E) a + R^2 delta^2 = 0. Writing the
 * quartic as (a^2 - R delta)^2 = (4 pi R^2 gamma / E) a, the larger root lies
 * on the branch a^2 - R delta = +R sqrt(4 pi gamma a / E), i.e. u = sqrt(a) is
 * the largest positive root of F(u) = u^4 - p u - q with p = R sqrt(4 pi
 * gamma / E), q = R delta. F is convex for u > 0, so Newton started above the
 * root decreases monotonically onto it; the loop stops when an iterate no
 * longer decreases. Returns NaN when no root exists (delta below pull-off). */
double oracle_jkr_contact_radius(double delta, double Req, double E, double gamma, int* iters)
{
    double p = Req * std::sqrt(4 * kPi * gamma / E);
    double q = Req * delta;
    double u = std::max(std::cbrt(2 * p), std::sqrt(std::sqrt(std::max(2 * q, 0.0))));
    int it = 0;
    if (u > 0) {
        for (; it < 200; ++it) {
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
    double u2 = u * u;
    if ((u2 * u2 - p * u) - q > 1e-9 * std::max(1.0, std::fabs(q))) return NAN;
    return u2;
}

/* Force magnitude of forcesIsotropicJKRLagrange (adhesion.hpp:206-212):
 * returns term1 - term2 = 4 E0 a^3 / (3 Req) - 4 sqrt(pi a^3 gamma E0). */
double oracle_jkr_force_magnitude(double radius, double Req, double E0, double gamma)
{
    double a3 = radius * radius * radius;
    double term1 = 4 * E0 * a3 / (3 * Req);
    double term2 = 4 * std::sqrt(kPi * a3 * gamma * E0);
    return term1 - term2;
}

/* newtonRaphson (rootFinding.hpp:42-75): Newton with the derivative from
 * central differences of step dx while |f| > tol and iter < max_iter. */
extern "C++" template <class F>
double newton_raphson(F&& f, double x0, double dx, double tol, int max_iter, double* f_out)
{
    double x = x0;
    double fx = f(x);
    int iter = 0;
    while (std::fabs(fx) > tol && iter < max_iter) {
        double deriv = (f(x + dx) - f(x - dx)) / (2 * dx);
        double update = -fx / deriv;
        x += update;
        fx = f(x);
        iter++;
    }
    if (f_out) *f_out = fx;
    return x;
}

/* jkrEquilibriumDistance (jkr.hpp:1310-1340): Newton on the overlap x of
 * f(x) = 4/3 E0/Req a^3 - 4 sqrt(pi gamma E0) a^1.5, a = JKR radius at x with
 * Req = R/2, from 2R - dinit; overlap capped at 2R - dmin (std::min order).
 * NaN when an iterate has no JKR root (the reference throws there). */
double oracle_jkr_equilibrium_distance(double R, double E0, double gamma, double dinit,
                                       double dmin, double increment_overlap, double newton_tol,
                                       int max_iter, double* residual)
{
    double overlap = 2 * R - dinit;
    double max_overlap = 2 * R - dmin;
    auto func = [&](double x) {
        double Req = R / 2;
        double radius = oracle_jkr_contact_radius(x, Req, E0, gamma, nullptr);
        double prefactor1 = 4.0 / 3.0 * E0 / Req;
        double prefactor2 = 4 * std::sqrt(kPi * gamma * E0);
        double f_hertz = prefactor1 * radius * radius * radius;
        double f_jkr = prefactor2 * std::pow(radius, 1.5);
        return f_hertz - f_jkr;
    };
    double root = newton_raphson(func, overlap, increment_overlap, newton_tol, max_iter, residual);
    double delta = std::min(root, max_overlap);
    return 2 * R - delta;
}

/* jkrOptimalSurfaceEnergyDensity (jkr.hpp:1369-1408): Newton on log10(gamma)
 * of deq(gamma) - deq_target, inner solves from 0.75 * 2R capped at 2 Rcell. */
double oracle_jkr_optimal_surface_energy_density(double R, double Rcell, double E0,
                                                 double deq_target, double init_gamma,
                                                 double log_increment_gamma,
                                                 double increment_overlap, double outer_tol,
                                                 double inner_tol, int outer_max_iter,
                                                 int inner_max_iter, double* residual)
{
    auto func = [&](double x) {
        double dinit = 0.75 * 2 * R;
        double dmin = 2 * Rcell;
        double gamma = std::pow(10.0, x);
        double deq = oracle_jkr_equilibrium_distance(R, E0, gamma, dinit, dmin, increment_overlap,
                                                     inner_tol, inner_max_iter, nullptr);
        return deq - deq_target;
    };
    double x = newton_raphson(func, std::log10(init_gamma), log_increment_gamma, outer_tol,
                              outer_max_iter, residual);
    return std::pow(10.0, x);
}

/* Eigen's LinSpaced(n, lo, hi) for floating point: low half from lo upward,
 * high half from hi downward, so both ends are exact. */
void oracle_linspace(double lo, double hi, int n, double* out)
{
    if (n == 1) { out[0] = hi; return; }
    double step = (hi - lo) / (n - 1);
    int flip = (n - 1) / 2;
    for (int i = 0; i < n; ++i)
        out[i] = (i < flip) ? lo + i * step : hi - (n - 1 - i) * step;
}

/* nearestValue, utils.hpp:1169-1215. */
int oracle_nearest_value(const double* values, int n, double x)
{
    if (x <= values[0]) return 0;
    else if (x >= values[n - 1]) return n - 1;
    int low = 0, high = n - 1, nearest = 0;
    while (low <= high) {
        int mid = (low + high) / 2;
        if (values[mid] <= x && x < values[mid + 1]) { nearest = mid; break; }
        else if (x >= values[mid + 1]) low = mid + 1;
        else high = mid - 1;
    }
    if (x - values[nearest] < values[nearest + 1] - x) return nearest;
    return nearest + 1;
}

/* Interpolated table radius, adhesion.hpp:312-328. */
double oracle_jkr_table_radius_1d(const double* overlaps, const double* radii, int n, double delta)
{
    int i1 = oracle_nearest_value(overlaps, n, delta);
    int i2 = (overlaps[i1] < delta ? i1 + 1 : i1 - 1);
    if (i1 > i2) std::swap(i1, i2);
    double d1 = overlaps[i1], d2 = overlaps[i2], r1 = radii[i1], r2 = radii[i2];
    return (r1 * (d2 - delta) + r2 * (delta - d1)) / (d2 - d1);
}

/* Bilinear table radius, adhesion.hpp:452-490. radii row-major [nd][ng]. */
double oracle_jkr_table_radius_2d(const double* overlaps, int nd, const double* gammas, int ng,
                                  const double* radii, double delta, double gamma)
{
    int d1i = oracle_nearest_value(overlaps, nd, delta);
    int g1i = oracle_nearest_value(gammas, ng, gamma);
    int d2i = (overlaps[d1i] < delta ? d1i + 1 : d1i - 1);
    if (d1i > d2i) std::swap(d1i, d2i);
    int g2i = (gammas[g1i] < gamma ? g1i + 1 : g1i - 1);
    if (g1i > g2i) std::swap(g1i, g2i);
    double d1 = overlaps[d1i], d2 = overlaps[d2i], g1 = gammas[g1i], g2 = gammas[g2i];
    double denom = (d2 - d1) * (g2 - g1);
    double w11 = (d2 - delta) * (g2 - gamma) / denom;
    double w12 = (d2 - delta) * (gamma - g1) / denom;
    double w21 = (delta - d1) * (g2 - gamma) / denom;
    double w22 = (delta - d1) * (gamma - g1) / denom;
    double radius = 0;
    radius += w11 * radii[d1i * ng + g1i];
    radius += w12 * radii[d1i * ng + g2i];
    radius += w21 * radii[d2i * ng + g1i];
    radius += w22 * radii[d2i * ng + g2i];
    return radius;
}

void oracle_set_jkr_table(const double* overlaps, int n_overlap, const double* gammas,
                          int n_gamma, const double* radii)
{
    g_tab_overlaps.assign(overlaps, overlaps + n_overlap);
    g_tab_gammas.assign(gammas ? gammas : overlaps, gammas ? gammas + n_gamma : overlaps);
    g_tab_radii.assign(radii, radii + (int64_t)n_overlap * std::max(n_gamma, 1));
}

void oracle_set_aniso_tables(const double* theta, int nt, const double* half_l, int nh,
                             const double* coords, int nc, const double* curv, const double* rx, int nr,
                             const double* phi, int np, const double* delta, int nd,
                             const double* gamma, int ng, const double* fr)
{
    g_an.theta.assign(theta, theta + nt);
    g_an.half_l.assign(half_l, half_l + nh);
    g_an.coords.assign(coords, coords + nc);
    g_an.curv.assign(curv, curv + (int64_t)nt * nh * nc);
    g_an.rx.assign(rx, rx + nr);
    g_an.phi.assign(phi, phi + np);
    g_an.delta.assign(delta, delta + nd);
    g_an.gamma.assign(gamma ? gamma : rx, gamma ? gamma + ng : rx);
    g_an.fr.assign(fr, fr + (int64_t)nr * nr * np * nd * std::max(ng, 1) * 2);
}

/* acosSafe, utils.hpp:629-638. */
static double acos_safe(double x)
{
    if (x >= 1) return 0;
    else if (x <= -1) return kPi;
    return std::acos(x);
}

static inline double dot3(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

static int nearest_or_zero(const std::vector<double>& v, double x)
{
    // nearestValue throws where its bisection fails (NaN); 0 keeps the oracle total
    if (std::isnan(x)) return 0;
    if (v.size() == 1) return 0;
    return oracle_nearest_value(v.data(), (int)v.size(), x);
}

/* forcesAnisotropicJKRLagrange, table overloads (adhesion.hpp:734-840 with a
 * fixed gamma, :842-1046 with a gamma grid): force magnitude (repulsion
 * positive) and contact radius of a pair at distance vector d12. The hash-map
 * lookups of the reference become reads of the dense tables, in which the
 * combinations absent from the files hold 0 (what operator[] returns). */
void oracle_aniso_jkr_force(const double* n1, double half_l1, const double* n2, double half_l2,
                            const double* d12, double R, double gamma, double s, double t,
                            double max_overlap, double* force_out, double* radius_out)
{
    *force_out = 0;
    *radius_out = 0;
    const double dist = norm3(d12[0], d12[1], d12[2]);
    if (!(dist < 2 * R)) return;
    const double d12n[3] = {d12[0] / dist, d12[1] / dist, d12[2] / dist};
    double delta = 2 * R - dist;
    if (max_overlap >= 0 && delta > max_overlap) delta = max_overlap;
    const double halfpi = kPi / 2;
    const double mn1[3] = {-n1[0], -n1[1], -n1[2]}, mn2[3] = {-n2[0], -n2[1], -n2[2]};
    const double md[3] = {-d12n[0], -d12n[1], -d12n[2]};
    double theta1 = acos_safe(dot3(d12n, n1));
    if (theta1 > halfpi) theta1 = acos_safe(dot3(d12n, mn1));
    double theta2 = acos_safe(dot3(md, n2));
    if (theta2 > halfpi) theta2 = acos_safe(dot3(md, mn2));
    const int it1 = nearest_or_zero(g_an.theta, theta1), it2 = nearest_or_zero(g_an.theta, theta2);
    const int ih1 = nearest_or_zero(g_an.half_l, half_l1), ih2 = nearest_or_zero(g_an.half_l, half_l2);
    const int is = nearest_or_zero(g_an.coords, std::abs(s) / half_l1);
    const int itt = nearest_or_zero(g_an.coords, std::abs(t) / half_l2);
    const int64_t nh = (int64_t)g_an.half_l.size(), nc = (int64_t)g_an.coords.size();
    double Rx1 = g_an.curv[(it1 * nh + ih1) * nc + is];
    double Rx2 = g_an.curv[(it2 * nh + ih2) * nc + itt];
    if (Rx1 > Rx2) std::swap(Rx1, Rx2);
    double phi = acos_safe(dot3(n1, n2));
    if (phi > halfpi) phi = kPi - phi;
    const int ir1 = nearest_or_zero(g_an.rx, Rx1), ir2 = nearest_or_zero(g_an.rx, Rx2);
    const int ip = nearest_or_zero(g_an.phi, phi), id = nearest_or_zero(g_an.delta, delta);
    const int64_t nr = (int64_t)g_an.rx.size(), np = (int64_t)g_an.phi.size(), nd = (int64_t)g_an.delta.size();
    const int64_t ng = std::max<int64_t>((int64_t)g_an.gamma.size(), 1);
    const int ig = g_an.gamma.empty() ? 0 : nearest_or_zero(g_an.gamma, gamma);
    const int64_t q = ((((ir1 * nr + ir2) * np + ip) * nd + id) * ng + ig) * 2;
    *force_out = g_an.fr[q];
    *radius_out = g_an.fr[q + 1];
}

/* integrals.hpp:56-64 (phi). */
static inline double phi_(double rz, double nz, double R, double s) { return R - rz - s * nz; }

/* integral1, integrals.hpp:116-144. */
double oracle_integral1(double rz, double nz, double R, double half_l, double g, double ss)
{
    if (ss > half_l) {
        double o1 = std::pow(phi_(rz, nz, R, -half_l), g + 1);
        double o2 = std::pow(phi_(rz, nz, R, half_l), g + 1);
        return (o1 - o2) / (nz * (g + 1));
    } else if (ss > -half_l) {
        double o1 = std::pow(phi_(rz, nz, R, -half_l), g + 1);
        return o1 / (nz * (g + 1));
    }
    return 0;
}

/* integral2, integrals.hpp:158-184. */
double oracle_integral2(double rz, double nz, double R, double half_l, double g, double ss)
{
    double term1 = oracle_integral1(rz, nz, R, half_l, g + 1, ss);
    double term2 = 0;
    if (ss > half_l) {
        double o1 = std::pow(phi_(rz, nz, R, -half_l), g + 1);
        double o2 = std::pow(phi_(rz, nz, R, half_l), g + 1);
        term2 = -half_l * (o1 + o2);
    } else if (ss > -half_l) {
        double o1 = std::pow(phi_(rz, nz, R, -half_l), g + 1);
        term2 = -half_l * o1;
    }
    return (term1 + term2) / (nz * (g + 1));
}

/* integral3, integrals.hpp:198-224. */
double oracle_integral3(double rz, double nz, double R, double half_l, double g, double ss)
{
    double term1 = 2 * oracle_integral2(rz, nz, R, half_l, g + 1, ss);
    double term2 = 0;
    if (ss > half_l) {
        double o1 = std::pow(phi_(rz, nz, R, -half_l), g + 1);
        double o2 = std::pow(phi_(rz, nz, R, half_l), g + 1);
        term2 = half_l * half_l * (o1 - o2);
    } else if (ss > -half_l) {
        double o1 = std::pow(phi_(rz, nz, R, -half_l), g + 1);
        term2 = half_l * half_l * o1;
    }
    return (term1 + term2) / (nz * (g + 1));
}

/* integral4..6, integrals.hpp:240-318. */
double oracle_integral4(double nz, double half_l, double ss)
{
    (void)nz;
    if (ss > half_l) return 2 * half_l;
    else if (ss > -half_l) return ss + half_l;
    return 0;
}
double oracle_integral5(double nz, double half_l, double ss)
{
    (void)nz;
    if (-half_l < ss && ss <= half_l) return (ss * ss - half_l * half_l) / 2.0;
    return 0.0;
}
double oracle_integral6(double nz, double half_l, double ss)
{
    (void)nz;
    if (ss > half_l) return 2 * half_l * half_l * half_l / 3.0;
    else if (ss > -half_l) return (ss * ss * ss + half_l * half_l * half_l) / 3.0;
    return 0.0;
}

/* areaIntegrals, integrals.hpp:400-463. */
void oracle_area_integrals(double rz, double nz, double R, double half_l, double ss, double* out)
{
    double int4 = oracle_integral4(nz, half_l, ss);
    double int5 = oracle_integral5(nz, half_l, ss);
    double int6 = oracle_integral6(nz, half_l, ss);
    double int1a = oracle_integral1(rz, nz, R, half_l, 0.5, ss);
    double int1b = oracle_integral1(rz, nz, R, half_l, 1.5, ss);
    double int1c = oracle_integral1(rz, nz, R, half_l, 2.5, ss);
    double sqR = std::sqrt(R);
    double term1 = sqR * (1 - nz * nz) * int1a;
    double term2 = kPi * R * nz * nz * int4;
    double area1 = term1 + term2;
    double overlap1 = 0, overlap2 = 0;
    if (ss > half_l) {
        overlap1 = phi_(rz, nz, R, -half_l);
        overlap2 = phi_(rz, nz, R, half_l);
    } else if (ss > -half_l) {
        overlap1 = phi_(rz, nz, R, -half_l);
    }
    double o1p1 = std::pow(overlap1, 1.5), o2p1 = std::pow(overlap2, 1.5);
    term2 = -half_l * (o1p1 + o2p1);
    double int2a = (int1b + term2) / (1.5 * nz);
    term1 = sqR * (1 - nz * nz) * int2a;
    term2 = kPi * R * nz * nz * int5;
    double area2 = term1 + term2;
    double o1p2 = o1p1 * overlap1, o2p2 = o2p1 * overlap2;
    term2 = -half_l * (o1p2 + o2p2);
    double int2b = (int1c + term2) / (2.5 * nz);
    term1 = 2 * int2b;
    term2 = half_l * half_l * (o1p1 - o2p1);
    double int3 = (term1 + term2) / (1.5 * nz);
    term1 = sqR * (1 - nz * nz) * int3;
    term2 = kPi * R * nz * nz * int6;
    double area3 = term1 + term2;
    out[0] = area1; out[1] = area2; out[2] = area3;
}

/* cellCellFrictionForces, mechanics.hpp:1031-1175 (include_constraint =
 * false, dz_threshold 1e-3 as passed at :1343-1346 and :1652-1655). dEdq is
 * n x 6 column-major, ld n, overwritten; rows accumulated in row order. */
static void friction_forces_impl(const double* cells, int n, int ld, const double* nb, int64_t m,
                                 int64_t ld_nb, const bfm_params* p, bool ic, double* dEdq);

void oracle_friction_forces(const double* cells, int n, int ld, const double* nb, int64_t m,
                            int64_t ld_nb, const bfm_params* p, double* dEdq)
{
    friction_forces_impl(cells, n, ld, nb, m, ld_nb, p, false, dEdq);
}

/* (s n) x v, the torque form of include_constraint = true (mechanics.hpp:1300-1305). */
static inline void cross_sn(double s, const double* nrm, const double* v, double* out)
{
    const double a[3] = {s * nrm[0], s * nrm[1], s * nrm[2]};
    out[0] = a[1] * v[2] - a[2] * v[1];
    out[1] = a[2] * v[0] - a[0] * v[2];
    out[2] = a[0] * v[1] - a[1] * v[0];
}

static inline void cell_n(const double* cells, int ld, int i, double* nrm)
{
    nrm[0] = cell(cells, ld, i, C_NX); nrm[1] = cell(cells, ld, i, C_NY); nrm[2] = cell(cells, ld, i, C_NZ);
}

static void friction_forces_impl(const double* cells, int n, int ld, const double* nb, int64_t m,
                                 int64_t ld_nb, const bfm_params* p, bool ic, double* dEdq)
{
    std::memset(dEdq, 0, sizeof(double) * 6 * (size_t)n);
    const double R = p->R, E0 = p->E0, Req = R / 2;
    const double coulomb = p->cell_cell_coulomb_coeff;
    const int ce = p->colidx_eta_cell_cell;
    for (int64_t k = 0; k < m; ++k) {
        int i = (int)nb[k], j = (int)nb[ld_nb + k];
        double dij[3] = {nb[2 * ld_nb + k], nb[3 * ld_nb + k], nb[4 * ld_nb + k]};
        double dist = norm3(dij[0], dij[1], dij[2]);
        double overlap = 2 * R - dist;
        if (!(overlap > 0)) continue;
        double si = nb[5 * ld_nb + k], sj = nb[6 * ld_nb + k];
        double dijn[3] = {dij[0] / dist, dij[1] / dist, dij[2] / dist};
        double eta = std::min(cell(cells, ld, i, ce), cell(cells, ld, j, ce));
        bool assume_2d = false;
        if (!p->no_surface && std::fabs(dijn[2]) < 1e-3) {
            double piz = cell(cells, ld, i, C_RZ) + si * cell(cells, ld, i, C_NZ);
            double pjz = cell(cells, ld, j, C_RZ) + sj * cell(cells, ld, j, C_NZ);
            if (piz < 1.01 * R && pjz < 1.01 * R) assume_2d = true;
        }
        double rvij[3];
        for (int c = 0; c < 3; ++c) {
            double vi = cell(cells, ld, i, C_DRX + c) + si * cell(cells, ld, i, C_DNX + c);
            double vj = cell(cells, ld, j, C_DRX + c) + sj * cell(cells, ld, j, C_DNX + c);
            rvij[c] = vi - vj;
        }
        if (assume_2d) {
            dij[2] = 0;
            dist = norm3(dij[0], dij[1], dij[2]);
            for (int c = 0; c < 3; ++c) dijn[c] = dij[c] / dist;
            overlap = 2 * R - dist;
            rvij[2] = 0;
        }
        double dot = (rvij[0] * dijn[0] + rvij[1] * dijn[1]) + rvij[2] * dijn[2];
        double force[3];
        double pre = -eta * std::sqrt(Req * overlap);
        for (int c = 0; c < 3; ++c) force[c] = pre * (rvij[c] - dot * dijn[c]);
        if (coulomb > 0) {
            double fn = norm3(force[0], force[1], force[2]);
            double max_norm = coulomb * (4. / 3.) * E0 * std::sqrt(Req) * std::pow(overlap, 1.5);
            if (fn > max_norm)
                for (int c = 0; c < 3; ++c) force[c] *= (max_norm / fn);
        }
        double ti[3], tj[3];
        if (!ic) {
            for (int c = 0; c < 3; ++c) { ti[c] = si * force[c]; tj[c] = (-sj) * force[c]; }
        } else {
            double ni[3], nj[3], mf[3] = {-force[0], -force[1], -force[2]};
            cell_n(cells, ld, i, ni);
            cell_n(cells, ld, j, nj);
            cross_sn(si, ni, force, ti);
            cross_sn(sj, nj, mf, tj);
        }
        for (int c = 0; c < 3; ++c) {
            dEdq[c * n + i] -= force[c];
            dEdq[(c + 3) * n + i] -= ti[c];
            dEdq[c * n + j] += force[c];
            dEdq[(c + 3) * n + j] -= tj[c];
        }
    }
}

/* cellCellInteractionForces: without adhesion mechanics.hpp:1209-1350, with
 * adhesion :1390-1659 (include_constraint = false), plus the friction array
 * when friction_mode is KINETIC (:1341-1347, :1650-1656). dEdq is n x 6
 * column-major, ld n, and is overwritten. Rows are accumulated in the order
 * of the neighbor rows, as the reference does. */
static void cell_cell_forces_impl(const double* cells, int n, int ld, const double* nb, int64_t m,
                                  int64_t ld_nb, const bfm_params* p, bool ic, double* dEdq);

void oracle_cell_cell_forces(const double* cells, int n, int ld, const double* nb,
                             int64_t m, int64_t ld_nb, const bfm_params* p, double* dEdq)
{
    cell_cell_forces_impl(cells, n, ld, nb, m, ld_nb, p, false, dEdq);
}

/* include_constraint = true: the torques of the Hertz, EPS-cap and hard-core
 * terms are (s n) x F (mechanics.hpp:1298-1306, 1554-1562, 1608-1616); those of
 * the JKR term are s (-(n . F) n ... ) as adhesion.hpp:225-231 writes them. */
void oracle_cell_cell_forces_ic(const double* cells, int n, int ld, const double* nb,
                                int64_t m, int64_t ld_nb, const bfm_params* p, double* dEdq)
{
    cell_cell_forces_impl(cells, n, ld, nb, m, ld_nb, p, true, dEdq);
}

static void cell_cell_forces_impl(const double* cells, int n, int ld, const double* nb, int64_t m,
                                  int64_t ld_nb, const bfm_params* p, bool ic, double* dEdq)
{
    std::memset(dEdq, 0, sizeof(double) * 6 * (size_t)n);
    const double R = p->R, Rcell = p->Rcell, E0 = p->E0;
    const double Req = R / 2;
    const double max_overlap = 2 * (R - Rcell);
    const bool adhesion = p->adhesion_mode != BFM_ADHESION_NONE;
    for (int64_t k = 0; k < m; ++k) {
        int i = (int)nb[k], j = (int)nb[ld_nb + k];
        double dij[3] = {nb[2 * ld_nb + k], nb[3 * ld_nb + k], nb[4 * ld_nb + k]};
        double dist = norm3(dij[0], dij[1], dij[2]);
        double overlap = 2 * R - dist;
        if (!(overlap > 0)) continue;
        double si = nb[5 * ld_nb + k], sj = nb[6 * ld_nb + k];
        double dijn[3] = {dij[0] / dist, dij[1] / dist, dij[2] / dist};
        double w[3];            // dE/dr_i of the main term; dE/dr_j = -w
        bool hard = false, jkr = false;
        double wh[3] = {0, 0, 0};
        if (!adhesion) {
            double mag;
            if (overlap <= max_overlap) {
                mag = p->repulsion_prefactors[0] * std::pow(overlap, 1.5);
            } else {
                mag = p->repulsion_prefactors[2];
                mag += p->repulsion_prefactors[1] * std::pow(overlap - max_overlap, 1.5);
            }
            for (int c = 0; c < 3; ++c) w[c] = mag * dijn[c];
        } else {
            double gamma = std::min(cell(cells, ld, i, p->colidx_gamma),
                                    cell(cells, ld, j, p->colidx_gamma));
            if (gamma > 0 && p->adhesion_mode == BFM_ADHESION_JKR_ANISOTROPIC) {
                // jkr_data->getJKRForces with anisotropic tables (mechanics.hpp:1468-1479)
                double ni_[3], nj_[3], force, radius;
                cell_n(cells, ld, i, ni_);
                cell_n(cells, ld, j, nj_);
                oracle_aniso_jkr_force(ni_, cell(cells, ld, i, C_HALF_L), nj_, cell(cells, ld, j, C_HALF_L),
                                       dij, R, gamma, si, sj, max_overlap, &force, &radius);
                // v = -force d12n; dEdq(i) gets -v = force d12n (adhesion.hpp:813-822)
                for (int c = 0; c < 3; ++c) w[c] = force * dijn[c];
                jkr = true;
            } else if (gamma > 0) {
                double delta = overlap;
                if (max_overlap >= 0 && delta > max_overlap) delta = max_overlap;
                double radius;
                if (p->jkr_mode == BFM_JKR_TABLE_1D)
                    radius = oracle_jkr_table_radius_1d(g_tab_overlaps.data(), g_tab_radii.data(),
                                                        (int)g_tab_overlaps.size(), delta);
                else if (p->jkr_mode == BFM_JKR_TABLE_2D)
                    radius = oracle_jkr_table_radius_2d(
                        g_tab_overlaps.data(), (int)g_tab_overlaps.size(), g_tab_gammas.data(),
                        (int)g_tab_gammas.size(), g_tab_radii.data(), delta, gamma);
                else
                    radius = oracle_jkr_contact_radius(delta, Req, E0, gamma, nullptr);
                double a3 = radius * radius * radius;
                double term1 = 4 * E0 * a3 / (3 * Req);
                double term2 = 4 * std::sqrt(kPi * a3 * gamma * E0);
                // v = (term2 - term1) * d12n; dEdq(i) gets -v and -s v, dEdq(j)
                // gets v and t v (adhesion.hpp:212-229). -v == (term1 - term2) * d12n.
                for (int c = 0; c < 3; ++c) w[c] = (term1 - term2) * dijn[c];
                jkr = true;
            } else {
                double mag;
                if (overlap <= 2 * (R - Rcell)) mag = p->repulsion_prefactors[0] * std::pow(overlap, 1.5);
                else mag = p->repulsion_prefactors[2];
                for (int c = 0; c < 3; ++c) w[c] = mag * dijn[c];
            }
            if (overlap > 2 * (R - Rcell)) {
                double mag = p->repulsion_prefactors[1] * std::pow(overlap - 2 * (R - Rcell), 1.5);
                for (int c = 0; c < 3; ++c) wh[c] = mag * dijn[c];
                hard = true;
            }
        }
        double ti[3], tj[3], ni[3], nj[3];
        if (ic) { cell_n(cells, ld, i, ni); cell_n(cells, ld, j, nj); }
        if (!ic) {
            for (int c = 0; c < 3; ++c) { ti[c] = si * w[c]; tj[c] = -sj * w[c]; }
        } else if (jkr) {
            // v = -w: dEdq(0, 3:5) = s (-w1 n1 - v), dEdq(1, 3:5) = t (w2 n2 + v),
            // w1 = n1 . (-v), w2 = n2 . (-v)   (adhesion.hpp:225-231)
            double w1 = (ni[0] * w[0] + ni[1] * w[1]) + ni[2] * w[2];
            double w2 = (nj[0] * w[0] + nj[1] * w[1]) + nj[2] * w[2];
            for (int c = 0; c < 3; ++c) {
                ti[c] = si * ((-w1) * ni[c] + w[c]);
                tj[c] = sj * (w2 * nj[c] - w[c]);
            }
        } else {
            double mw[3] = {-w[0], -w[1], -w[2]};
            cross_sn(si, ni, w, ti);
            cross_sn(sj, nj, mw, tj);
        }
        for (int c = 0; c < 3; ++c) {
            dEdq[c * n + i] += w[c];
            dEdq[(c + 3) * n + i] += ti[c];
        }
        for (int c = 0; c < 3; ++c) {
            dEdq[c * n + j] += -w[c];
            dEdq[(c + 3) * n + j] += tj[c];
        }
        if (hard) {
            double thi[3], thj[3];
            if (!ic) {
                for (int c = 0; c < 3; ++c) { thi[c] = si * wh[c]; thj[c] = -sj * wh[c]; }
            } else {
                double mwh[3] = {-wh[0], -wh[1], -wh[2]};
                cross_sn(si, ni, wh, thi);
                cross_sn(sj, nj, mwh, thj);
            }
            for (int c = 0; c < 3; ++c) {
                dEdq[c * n + i] += wh[c];
                dEdq[(c + 3) * n + i] += thi[c];
            }
            for (int c = 0; c < 3; ++c) {
                dEdq[c * n + j] += -wh[c];
                dEdq[(c + 3) * n + j] += thj[c];
            }
        }
    }
    if (p->friction_mode == 1) {
        std::vector<double> fr(6 * (size_t)n);
        friction_forces_impl(cells, n, ld, nb, m, ld_nb, p, ic, fr.data());
        for (size_t q = 0; q < 6 * (size_t)n; ++q) dEdq[q] += fr[q];
    }
}

/* cellSurfaceRepulsionForces, mechanics.hpp:673-780 (include_constraint =
 * false). Writes columns 2 and 5 of dEdq (n x 6, zeroed here). */
static void surface_repulsion_impl(const double* cells, int n, int ld, const double* ss, double R, double E0,
                                   const int* assume_2d, bool ic, double* dEdq);

void oracle_surface_repulsion_forces(const double* cells, int n, int ld, const double* ss,
                                     double R, double E0, const int* assume_2d, double* dEdq)
{
    surface_repulsion_impl(cells, n, ld, ss, R, E0, assume_2d, false, dEdq);
}

static void surface_repulsion_impl(const double* cells, int n, int ld, const double* ss, double R, double E0,
                                   const int* assume_2d, bool ic, double* dEdq)
{
    std::memset(dEdq, 0, sizeof(double) * 6 * (size_t)n);
    const double prefactor0 = 2 * E0;
    const double prefactor1 = (8. / 3.) * E0 * std::pow(R, 0.5);
    const double prefactor2 = 2 * E0 * std::pow(R, 0.5);
    #pragma omp parallel for if(n >= 50)
    for (int i = 0; i < n; ++i) {
        double rz = cell(cells, ld, i, C_RZ), nz = cell(cells, ld, i, C_NZ);
        double hl = cell(cells, ld, i, C_HALF_L);
        if (assume_2d[i]) {
            double ph = R - rz;
            if (ph > 0) dEdq[2 * n + i] = -prefactor0 * ph * cell(cells, ld, i, C_L);
        } else {
            double nz2 = nz * nz;
            double int1 = oracle_integral1(rz, nz, R, hl, 1.0, ss[i]);
            double int2 = oracle_integral1(rz, nz, R, hl, 0.5, ss[i]);
            dEdq[2 * n + i] = -prefactor0 * ((1.0 - nz2) * int1 + std::sqrt(R) * nz2 * int2);
            if (ic) {   // mechanics.hpp:741-754
                double i3 = oracle_integral2(rz, nz, R, hl, 1.0, ss[i]);
                double i4 = oracle_integral2(rz, nz, R, hl, 0.5, ss[i]);
                double cx = cell(cells, ld, i, C_NY), cy = -cell(cells, ld, i, C_NX);
                dEdq[3 * n + i] = -prefactor0 * cx * ((1.0 - nz2) * i3 + std::sqrt(R) * nz2 * i4);
                dEdq[4 * n + i] = -prefactor0 * cy * ((1.0 - nz2) * i3 + std::sqrt(R) * nz2 * i4);
                continue;
            }
            double int3 = oracle_integral1(rz, nz, R, hl, 2.0, ss[i]);
            double int4 = oracle_integral2(rz, nz, R, hl, 1.0, ss[i]);
            double int5 = oracle_integral1(rz, nz, R, hl, 1.5, ss[i]);
            double int6 = oracle_integral2(rz, nz, R, hl, 0.5, ss[i]);
            double f5 = 0;
            f5 -= prefactor0 * nz * int3;
            f5 -= prefactor0 * (1 - nz2) * int4;
            f5 += prefactor1 * nz * int5;
            f5 -= prefactor2 * nz2 * int6;
            dEdq[5 * n + i] = f5;
        }
    }
}

/* cellSurfaceAdhesionForces, mechanics.hpp:800-910 (include_constraint = false). */
static void surface_adhesion_impl(const double* cells, int n, int ld, const double* ss, double R,
                                  const int* assume_2d, bool ic, double* dEdq);

void oracle_surface_adhesion_forces(const double* cells, int n, int ld, const double* ss,
                                    double R, const int* assume_2d, double* dEdq)
{
    surface_adhesion_impl(cells, n, ld, ss, R, assume_2d, false, dEdq);
}

static void surface_adhesion_impl(const double* cells, int n, int ld, const double* ss, double R,
                                  const int* assume_2d, bool ic, double* dEdq)
{
    std::memset(dEdq, 0, sizeof(double) * 6 * (size_t)n);
    const double prefactor0 = std::pow(R, 0.5) / 2;
    const double prefactor1 = 2 * kPi * R;
    const double prefactor2 = 2 * std::pow(R, 0.5);
    #pragma omp parallel for if(n >= 50)
    for (int i = 0; i < n; ++i) {
        double rz = cell(cells, ld, i, C_RZ), nz = cell(cells, ld, i, C_NZ);
        double hl = cell(cells, ld, i, C_HALF_L), sigma0 = cell(cells, ld, i, C_SIGMA0);
        if (assume_2d[i]) {
            double ph = R - rz;
            if (ph > 0) dEdq[2 * n + i] = sigma0 * prefactor0 * cell(cells, ld, i, C_L) / std::pow(ph, 0.5);
        } else {
            double nz2 = nz * nz;
            double int1 = oracle_integral1(rz, nz, R, hl, -0.5, ss[i]);
            double term2 = 0.0;
            if (std::fabs(ss[i]) < hl) term2 = (prefactor1 / 2) * nz;
            double f2 = prefactor0 * (1 - nz2) * int1 + term2;
            f2 *= sigma0;
            dEdq[2 * n + i] = f2;
            if (ic) {   // mechanics.hpp:868-884
                double i2 = oracle_integral2(rz, nz, R, hl, -0.5, ss[i]);
                double cx = cell(cells, ld, i, C_NY), cy = -cell(cells, ld, i, C_NX);
                double w1 = 0, w2 = 0;
                if (std::fabs(ss[i]) < hl) {
                    w1 = (prefactor1 / 2) * ss[i] * nz * cx;
                    w2 = (prefactor1 / 2) * ss[i] * nz * cy;
                }
                double f3 = prefactor0 * (1 - nz2) * i2 * cx - w1;
                f3 *= sigma0;
                double f4 = prefactor0 * (1 - nz2) * i2 * cy - w2;
                f4 *= sigma0;
                dEdq[3 * n + i] = f3;
                dEdq[4 * n + i] = f4;
                continue;
            }
            double int2 = oracle_integral1(rz, nz, R, hl, 0.5, ss[i]);
            double int3 = oracle_integral2(rz, nz, R, hl, -0.5, ss[i]);
            double int4 = oracle_integral4(nz, hl, ss[i]);
            double term4 = 0.0;
            if (std::fabs(ss[i]) < hl) term4 = (prefactor1 / 2) * (R - rz);
            double f5 = 0;
            f5 += prefactor2 * nz * int2;
            f5 += prefactor0 * (1 - nz2) * int3;
            f5 -= prefactor1 * nz * int4;
            f5 += term4;
            f5 *= sigma0;
            dEdq[5 * n + i] = f5;
        }
    }
}

static void sstar_all(const double* cells, int n, int ld, double R, const int* assume_2d, double* ss)
{
    for (int i = 0; i < n; ++i)
        ss[i] = assume_2d[i] ? NAN : (R - cell(cells, ld, i, C_RZ)) / cell(cells, ld, i, C_NZ);
}

/* getConservativeForces, mechanics.hpp:1702-1769 (include_constraint = false):
 * -F_cc - F_rep - F_adh, left to right. */
void oracle_conservative_forces(const double* cells, int n, int ld, const double* nb,
                                int64_t m, int64_t ld_nb, const int* assume_2d,
                                const bfm_params* p, double* forces)
{
    std::vector<double> ss(n), cc(6 * (size_t)n), rep(6 * (size_t)n, 0.0), adh(6 * (size_t)n, 0.0);
    sstar_all(cells, n, ld, p->R, assume_2d, ss.data());
    oracle_cell_cell_forces(cells, n, ld, nb, m, ld_nb, p, cc.data());
    if (!p->no_surface) {
        oracle_surface_repulsion_forces(cells, n, ld, ss.data(), p->R, p->E0, assume_2d, rep.data());
        oracle_surface_adhesion_forces(cells, n, ld, ss.data(), p->R, assume_2d, adh.data());
    }
    for (size_t q = 0; q < 6 * (size_t)n; ++q) forces[q] = (-cc[q] - rep[q]) - adh[q];
}

/* getVelocities, mechanics.hpp:2011-2347. */
void oracle_get_velocities(const double* cells, int n, int ld, const double* nb, int64_t m,
                           int64_t ld_nb, const int* assume_2d, const double* noise,
                           const bfm_params* p, double* vel)
{
    const double R = p->R, E0 = p->E0;
    const double coeff = p->cell_surface_coulomb_coeff;
    std::vector<double> ss(n), F(6 * (size_t)n);
    sstar_all(cells, n, ld, R, assume_2d, ss.data());
    oracle_conservative_forces(cells, n, ld, nb, m, ld_nb, assume_2d, p, F.data());
    if (noise)
        for (size_t q = 0; q < 6 * (size_t)n; ++q) F[q] += noise[q];
    #pragma omp parallel for if(n >= 50)
    for (int i = 0; i < n; ++i) {
        double f[6];
        for (int c = 0; c < 6; ++c) f[c] = F[(size_t)c * n + i];
        double v[6];
        const double rz = cell(cells, ld, i, C_RZ), nx = cell(cells, ld, i, C_NX);
        const double ny = cell(cells, ld, i, C_NY), nz = cell(cells, ld, i, C_NZ);
        const double l = cell(cells, ld, i, C_L), hl = cell(cells, ld, i, C_HALF_L);
        const double eta0 = cell(cells, ld, i, C_ETA0);
        bool contacting;
        if (!assume_2d[i]) {
            double z1 = rz - hl * nz, z2 = rz + hl * nz;
            contacting = (z1 < R || z2 < R);
        } else {
            contacting = (rz < R);
        }
        if (p->no_surface || !contacting) {
            double K = eta0 * l;
            double L = K * l * l / 12;
            double mult = -(nx * f[3] + ny * f[4] + nz * f[5]);
            v[0] = f[0] / K; v[1] = f[1] / K; v[2] = f[2] / K;
            v[3] = (f[3] + mult * nx) / L;
            v[4] = (f[4] + mult * ny) / L;
            v[5] = (f[5] + mult * nz) / L;
        } else if (assume_2d[i]) {
            double K, L;
            double eta1 = cell(cells, ld, i, C_ETA1);
            if (coeff == 0) {
                double a = std::sqrt(R * (R - rz));
                double eta_combined = eta0 + eta1 * a / R;
                K = eta_combined * l;
                L = K * l * l / 12;
            } else {
                double K_vis = eta0 * l;
                double L_vis = K_vis * l * l / 12;
                double a = std::sqrt(R * (R - rz));
                double K_fric = eta1 * l * a / R;
                double drx = cell(cells, ld, i, C_DRX), dry = cell(cells, ld, i, C_DRY);
                double drxy = std::sqrt(drx * drx + dry * dry);
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
            double eta1 = cell(cells, ld, i, C_ETA1);
            double ai[3];
            oracle_area_integrals(rz, nz, R, hl, ss[i], ai);
            if (coeff > 0) {
                double drx = cell(cells, ld, i, C_DRX), dry = cell(cells, ld, i, C_DRY);
                double dnx = cell(cells, ld, i, C_DNX), dny = cell(cells, ld, i, C_DNY);
                double drxy = std::sqrt(drx * drx + dry * dry);
                double dnxy = std::sqrt(dnx * dnx + dny * dny);
                double friction_norm = ai[0] * drxy;
                friction_norm += ai[1] * dnxy;
                friction_norm *= (eta1 / R);
                double max_norm = (1 - nz * nz) * oracle_integral1(rz, nz, R, hl, 1, ss[i]);
                max_norm += std::sqrt(R) * nz * nz * oracle_integral1(rz, nz, R, hl, 0.5, ss[i]);
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
        for (int c = 0; c < 6; ++c) vel[(size_t)c * n + i] = v[c];
    }
}

/* normalizeOrientations, mechanics.hpp:1918-1959. */
void oracle_normalize_orientations(double* cells, int n, int ld)
{
    for (int i = 0; i < n; ++i) {
        double* nx = &cells[(int64_t)C_NX * ld + i];
        double* ny = &cells[(int64_t)C_NY * ld + i];
        double* nz = &cells[(int64_t)C_NZ * ld + i];
        double nrm = norm3(*nx, *ny, *nz);
        *nx /= nrm; *ny /= nrm; *nz /= nrm;
        if (*nz < 0) { *nx *= -1; *ny *= -1; *nz *= -1; }
    }
}

/* stepRungeKuttaAdaptive, mechanics.hpp:2406-2612. A is s x s row-major.
 * cells_new has leading dimension n and the column count of params.ncols. */
void oracle_step_rk(const double* A, const double* b, const double* bs, int s, double dt,
                    const double* cells, int n, int ld, const double* nb, int64_t m,
                    int64_t ld_nb, const double* noise, const bfm_params* p,
                    double* cells_new, double* errors)
{
    const int ncols = p->ncols;
    std::vector<int> assume_2d(n);
    for (int i = 0; i < n; ++i) assume_2d[i] = (cell(cells, ld, i, C_NZ) < p->nz_threshold);
    std::vector<std::vector<double>> k(s, std::vector<double>(6 * (size_t)n));
    std::vector<double> ci((size_t)ncols * n);
    for (int c = 0; c < ncols; ++c)
        for (int i = 0; i < n; ++i) ci[(size_t)c * n + i] = cell(cells, ld, i, c);
    oracle_get_velocities(ci.data(), n, n, nb, m, ld_nb, assume_2d.data(), noise, p, k[0].data());
    for (int st = 1; st < s; ++st) {
        std::vector<double> mult(6 * (size_t)n, 0.0);
        for (int j = 0; j < st; ++j)
            for (size_t q = 0; q < 6 * (size_t)n; ++q) mult[q] += k[j][q] * A[st * s + j];
        for (int c = 0; c < ncols; ++c)
            for (int i = 0; i < n; ++i) ci[(size_t)c * n + i] = cell(cells, ld, i, c);
        for (int c = 0; c < 6; ++c)
            for (int i = 0; i < n; ++i) ci[(size_t)(c + 1) * n + i] += mult[(size_t)c * n + i] * dt;
        oracle_normalize_orientations(ci.data(), n, n);
        oracle_get_velocities(ci.data(), n, n, nb, m, ld_nb, assume_2d.data(), noise, p, k[st].data());
    }
    std::vector<double> v1(6 * (size_t)n, 0.0), v2(6 * (size_t)n, 0.0);
    for (int st = 0; st < s; ++st)
        for (size_t q = 0; q < 6 * (size_t)n; ++q) {
            v1[q] += k[st][q] * b[st];
            v2[q] += k[st][q] * bs[st];
        }
    for (int c = 0; c < ncols; ++c)
        for (int i = 0; i < n; ++i) cells_new[(size_t)c * n + i] = cell(cells, ld, i, c);
    for (size_t q = 0; q < 6 * (size_t)n; ++q) {
        double d1 = v1[q] * dt, d2 = v2[q] * dt;
        cells_new[(size_t)n + q] += d1;
        errors[q] = d1 - d2;
        cells_new[(size_t)7 * n + q] = v1[q];
    }
    oracle_normalize_orientations(cells_new, n, n);
}

/* ---- velocity Verlet ------------------------------------------------------ */

/* getConservativeForces with include_constraint = true (mechanics.hpp:1702-1769). */
void oracle_conservative_forces_ic(const double* cells, int n, int ld, const double* nb, int64_t m,
                                   int64_t ld_nb, const int* assume_2d, const bfm_params* p, double* forces)
{
    std::vector<double> ss(n), cc(6 * (size_t)n), rep(6 * (size_t)n, 0.0), adh(6 * (size_t)n, 0.0);
    sstar_all(cells, n, ld, p->R, assume_2d, ss.data());
    cell_cell_forces_impl(cells, n, ld, nb, m, ld_nb, p, true, cc.data());
    if (!p->no_surface) {
        surface_repulsion_impl(cells, n, ld, ss.data(), p->R, p->E0, assume_2d, true, rep.data());
        surface_adhesion_impl(cells, n, ld, ss.data(), p->R, assume_2d, true, adh.data());
    }
    for (size_t q = 0; q < 6 * (size_t)n; ++q) forces[q] = (-cc[q] - rep[q]) - adh[q];
}

/* ambientViscosityForces, mechanics.hpp:1785-1806: eta0 l^3 dr and
 * eta0 l^9 (n x dn) / 12, as written there. */
void oracle_ambient_viscosity_forces(const double* cells, int n, int ld, double* dEdq)
{
    for (int i = 0; i < n; ++i) {
        double nv[3], dn[3], w[3];
        cell_n(cells, ld, i, nv);
        dn[0] = cell(cells, ld, i, C_DNX); dn[1] = cell(cells, ld, i, C_DNY); dn[2] = cell(cells, ld, i, C_DNZ);
        w[0] = nv[1] * dn[2] - nv[2] * dn[1];
        w[1] = nv[2] * dn[0] - nv[0] * dn[2];
        w[2] = nv[0] * dn[1] - nv[1] * dn[0];
        const double eta = cell(cells, ld, i, C_ETA0);
        const double li = std::pow(cell(cells, ld, i, C_L), 3);
        for (int c = 0; c < 3; ++c) {
            dEdq[c * n + i] = eta * li * cell(cells, ld, i, C_DRX + c);
            dEdq[(c + 3) * n + i] = eta * li * li * li * w[c] / 12;
        }
    }
}

/* cellSurfaceFrictionForces, mechanics.hpp:1822-1900. */
void oracle_surface_friction_forces(const double* cells, int n, int ld, const double* ss, double R,
                                    const int* assume_2d, double* dEdq)
{
    std::memset(dEdq, 0, sizeof(double) * 6 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        double nv[3], dn[3], w[3];
        cell_n(cells, ld, i, nv);
        dn[0] = cell(cells, ld, i, C_DNX); dn[1] = cell(cells, ld, i, C_DNY); dn[2] = cell(cells, ld, i, C_DNZ);
        w[0] = nv[1] * dn[2] - nv[2] * dn[1];
        w[1] = nv[2] * dn[0] - nv[0] * dn[2];
        w[2] = nv[0] * dn[1] - nv[1] * dn[0];
        const double q[3] = {cell(cells, ld, i, C_DRX), cell(cells, ld, i, C_DRY), 0.0};
        const double mv[3] = {w[1] * nv[2] - w[2] * nv[1], w[2] * nv[0] - w[0] * nv[2], 0.0};
        const double eta = cell(cells, ld, i, C_ETA1);
        const double l = cell(cells, ld, i, C_L);
        if (assume_2d[i]) {
            const double phi = R - cell(cells, ld, i, C_RZ);
            if (phi > 0) {
                const double area = std::sqrt(R) * std::sqrt(phi) * l;
                dEdq[0 * n + i] = eta * q[0] * area / R;
                dEdq[1 * n + i] = eta * q[1] * area / R;
                const double cross_z = nv[0] * mv[1] - nv[1] * mv[0];
                const double l2 = l * l;
                dEdq[5 * n + i] = eta * cross_z * area * l2 / (12 * R);
            }
        } else {
            double ai[3];
            oracle_area_integrals(cell(cells, ld, i, C_RZ), nv[2], R, cell(cells, ld, i, C_HALF_L), ss[i], ai);
            const double int1 = ai[0], int2 = ai[1], int3 = ai[2];
            dEdq[0 * n + i] = eta * (q[0] * int1 + mv[0] * int2) / R;
            dEdq[1 * n + i] = eta * (q[1] * int1 + mv[1] * int2) / R;
            const double c1[3] = {nv[1] * q[2] - nv[2] * q[1], nv[2] * q[0] - nv[0] * q[2], nv[0] * q[1] - nv[1] * q[0]};
            const double c2[3] = {nv[1] * mv[2] - nv[2] * mv[1], nv[2] * mv[0] - nv[0] * mv[2],
                                  nv[0] * mv[1] - nv[1] * mv[0]};
            for (int c = 0; c < 3; ++c) dEdq[(3 + c) * n + i] = eta * (c1[c] * int2 + c2[c] * int3) / R;
        }
    }
}

/* The generalized forces of stepVelocityVerlet (mechanics.hpp:2845-2872):
 * getConservativeForces (include_constraint = true) + noise - ambient
 * viscosity - cell-surface friction truncated to cs_coulomb |repulsion|. */
static void verlet_forces(const double* cells, int n, int ld, const double* nb, int64_t m, int64_t ld_nb,
                          const int* a2d, const double* noise, const bfm_params* p, double* F)
{
    oracle_conservative_forces_ic(cells, n, ld, nb, m, ld_nb, a2d, p, F);
    if (noise)
        for (size_t q = 0; q < 6 * (size_t)n; ++q) F[q] += noise[q];
    std::vector<double> amb(6 * (size_t)n);
    oracle_ambient_viscosity_forces(cells, n, ld, amb.data());
    for (size_t q = 0; q < 6 * (size_t)n; ++q) F[q] -= amb[q];
    if (!p->no_surface) {
        std::vector<double> ss(n), sf(6 * (size_t)n), rep(6 * (size_t)n);
        sstar_all(cells, n, ld, p->R, a2d, ss.data());
        oracle_surface_friction_forces(cells, n, ld, ss.data(), p->R, a2d, sf.data());
        if (p->cell_surface_coulomb_coeff > 0) {
            surface_repulsion_impl(cells, n, ld, ss.data(), p->R, p->E0, a2d, true, rep.data());
            for (int i = 0; i < n; ++i) {
                double fn2 = 0, rn2 = 0;
                for (int c = 0; c < 6; ++c) {
                    fn2 += sf[(size_t)c * n + i] * sf[(size_t)c * n + i];
                    rn2 += rep[(size_t)c * n + i] * rep[(size_t)c * n + i];
                }
                const double fn = std::sqrt(fn2), mx = p->cell_surface_coulomb_coeff * std::sqrt(rn2);
                if (fn > mx)
                    for (int c = 0; c < 6; ++c) sf[(size_t)c * n + i] *= (mx / fn);
            }
        }
        for (size_t q = 0; q < 6 * (size_t)n; ++q) F[q] -= sf[q];
    }
}

/* Accelerations of stepVelocityVerlet (mechanics.hpp:2874-2891, 2975-2986):
 * delta1 = F / mass, delta2 = torque_par / I0 + torque_perp / I1. */
static void verlet_accel(const double* F, const double* cells, int n, int ld, int i, const double* mass,
                         const double* I0, const double* I1, double d1[3], double d2[3])
{
    double nv[3], t[3], par[3], perp[3];
    cell_n(cells, ld, i, nv);
    for (int c = 0; c < 3; ++c) t[c] = F[(size_t)(3 + c) * n + i];
    const double dot = (t[0] * nv[0] + t[1] * nv[1]) + t[2] * nv[2];
    for (int c = 0; c < 3; ++c) {
        par[c] = dot * nv[c];
        perp[c] = t[c] - par[c];
    }
    for (int c = 0; c < 3; ++c) {
        d1[c] = F[(size_t)c * n + i] / mass[i];
        d2[c] = par[c] / I0[i];
        d2[c] += perp[c] / I1[i];
    }
}

/* stepVelocityVerlet, mechanics.hpp:2668-2996, with caller-drawn noise (n x 6,
 * or NULL for none). cells_new has leading dimension n; neighbors are used
 * as given for the first force evaluation and with distances updated to the
 * half step for the second (a private copy, as the reference's neighbors_new). */
void oracle_step_verlet(double dt, double rho, const double* cells, int n, int ld, const double* nb, int64_t m,
                        int64_t ld_nb, const double* noise, const bfm_params* p, double* cells_new)
{
    const int ncols = p->ncols;
    const double R = p->R;
    std::vector<int> a2d(n);
    for (int i = 0; i < n; ++i) a2d[i] = cell(cells, ld, i, C_NZ) < p->nz_threshold;
    std::vector<double> mass(n), I0(n), I1(n);
    for (int i = 0; i < n; ++i) {
        const double length = cell(cells, ld, i, C_L);
        const double volume = kPi * ((4. / 3.) * std::pow(R, 3) + R * R * length);
        mass[i] = rho * volume;
        const double diam = 2 * R, d2 = diam * diam, d3 = d2 * diam, d4 = d3 * diam;
        const double l2 = length * length, l3 = l2 * length;
        I0[i] = d4 * (length / 32 + diam / 60);
        I1[i] = d2 * (4 * d3 + 15 * length * d2 + 20 * l2 * diam + 10 * l3) / 480;
        I0[i] *= (rho * kPi);
        I1[i] *= (rho * kPi);
    }
    std::vector<double> F(6 * (size_t)n);
    verlet_forces(cells, n, ld, nb, m, ld_nb, a2d.data(), noise, p, F.data());
    for (int c = 0; c < ncols; ++c)
        for (int i = 0; i < n; ++i) cells_new[(size_t)c * n + i] = cell(cells, ld, i, c);
    std::vector<double> wh(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        double d1[3], d2[3], nv[3], dn[3], w[3], r[3], v[3], wi[3];
        verlet_accel(F.data(), cells, n, ld, i, mass.data(), I0.data(), I1.data(), d1, d2);
        cell_n(cells, ld, i, nv);
        dn[0] = cell(cells, ld, i, C_DNX); dn[1] = cell(cells, ld, i, C_DNY); dn[2] = cell(cells, ld, i, C_DNZ);
        w[0] = nv[1] * dn[2] - nv[2] * dn[1];
        w[1] = nv[2] * dn[0] - nv[0] * dn[2];
        w[2] = nv[0] * dn[1] - nv[1] * dn[0];
        const double h = 0.5 * dt;
        for (int c = 0; c < 3; ++c) {
            r[c] = cell(cells, ld, i, C_RX + c);
            v[c] = cell(cells, ld, i, C_DRX + c) + h * d1[c];
            wi[c] = w[c] + h * d2[c];
            wh[(size_t)c * n + i] = wi[c];
        }
        const double hl = cell(cells, ld, i, C_HALF_L), l = cell(cells, ld, i, C_L);
        double pp[3], qq[3], pq[3], cr[3], pn[3], qn[3];
        for (int c = 0; c < 3; ++c) {
            pp[c] = r[c] - hl * nv[c];
            qq[c] = r[c] + hl * nv[c];
            pq[c] = pp[c] - qq[c];
        }
        cr[0] = (wi[1] * pq[2] - wi[2] * pq[1]) / 2;
        cr[1] = (wi[2] * pq[0] - wi[0] * pq[2]) / 2;
        cr[2] = (wi[0] * pq[1] - wi[1] * pq[0]) / 2;
        for (int c = 0; c < 3; ++c) {
            pn[c] = pp[c] + dt * (v[c] + cr[c]);
            qn[c] = qq[c] + dt * (v[c] - cr[c]);
            cells_new[(size_t)(C_RX + c) * n + i] = (pn[c] + qn[c]) / 2;
            cells_new[(size_t)(C_NX + c) * n + i] = (qn[c] - pn[c]) / l;
            cells_new[(size_t)(C_DRX + c) * n + i] = v[c];
        }
        cells_new[(size_t)C_DNX * n + i] = wi[1] * nv[2] - wi[2] * nv[1];
        cells_new[(size_t)C_DNY * n + i] = wi[2] * nv[0] - wi[0] * nv[2];
        cells_new[(size_t)C_DNZ * n + i] = wi[0] * nv[1] - wi[1] * nv[0];
    }
    oracle_normalize_orientations(cells_new, n, n);
    std::vector<double> nb2((size_t)7 * (m > 0 ? m : 1));
    for (int c = 0; c < 7; ++c)
        for (int64_t k = 0; k < m; ++k) nb2[(size_t)c * m + k] = nb[(size_t)c * ld_nb + k];
    oracle_update_neighbor_distances(cells_new, n, n, nb2.data(), m, m > 0 ? m : 1);
    for (int i = 0; i < n; ++i) a2d[i] = cells_new[(size_t)C_NZ * n + i] < p->nz_threshold;
    verlet_forces(cells_new, n, n, nb2.data(), m, m > 0 ? m : 1, a2d.data(), noise, p, F.data());
    for (int i = 0; i < n; ++i) {
        double d1[3], d2[3], nv[3], wf[3];
        verlet_accel(F.data(), cells_new, n, n, i, mass.data(), I0.data(), I1.data(), d1, d2);
        cell_n(cells_new, n, i, nv);
        const double h = 0.5 * dt;
        for (int c = 0; c < 3; ++c) {
            cells_new[(size_t)(C_DRX + c) * n + i] += h * d1[c];
            wf[c] = wh[(size_t)c * n + i] + h * d2[c];
        }
        cells_new[(size_t)C_DNX * n + i] = wf[1] * nv[2] - wf[2] * nv[1];
        cells_new[(size_t)C_DNY * n + i] = wf[2] * nv[0] - wf[0] * nv[2];
        cells_new[(size_t)C_DNZ * n + i] = wf[0] * nv[1] - wf[1] * nv[0];
    }
    oracle_normalize_orientations(cells_new, n, n);
}

/* The controller's error measure, simulation.hpp:1688-1695:
 * max(max_k errors_k / (tol * (1 + |y_k|)), min_error), a signed maximum. */
double oracle_max_scaled_error(const double* cells, int n, int ld, const double* errors,
                               double tol, double min_error)
{
    double mx = -INFINITY;
    for (int c = 0; c < 6; ++c)
        for (int i = 0; i < n; ++i) {
            double z = 1.0 + std::fabs(cell(cells, ld, i, c + 1));
            double v = errors[(size_t)c * n + i] / (tol * z);
            if (v > mx) mx = v;
        }
    return std::max(mx, min_error);
}

/* Step-size controller, simulation.hpp:1688-1769, driven by a script of
 * errors instead of real steps. */
int oracle_controller_trace(double dt, double error_order, int max_tries,
                            double min_stepsize, double max_stepsize,
                            const double* err_seq, int n_err, double* dts)
{
    int k = 0, nd = 0;
    dts[nd++] = dt;
    double max_error = err_seq[std::min(k++, n_err - 1)];
    bool exceeded = max_error > 1.0;
    double factor = 0.9 * std::pow(1.0 / max_error, 1.0 / (error_order + 1));
    if (factor >= 10) factor = 10;
    else if (factor < 0.2) factor = 0.2;
    int j = 0;
    while (exceeded && j < max_tries) {
        dts[nd++] = dt * factor;
        max_error = err_seq[std::min(k++, n_err - 1)];
        exceeded = max_error > 1.0;
        factor *= 0.9 * std::pow(1.0 / max_error, 1.0 / (error_order + 1));
        if (factor >= 10) { factor = 10; break; }
        else if (factor < 0.2) { factor = 0.2; break; }
        j++;
    }
    if (dt * factor < min_stepsize) factor = min_stepsize / dt;
    else if (dt * factor > max_stepsize) factor = max_stepsize / dt;
    dt *= factor;
    dts[nd++] = dt;
    return nd;
}


}  // extern "C"

/* ---- growth, division and the simulation loop ------------------------- */

namespace {

/* boost::random::uniform_01<> on mt19937: x * 2^-32 (growth.hpp callers,
 * simulation.hpp:1294). std::mt19937 is the same generator as boost's. */
double orc_uniform(std::mt19937& g) { return (double)g() * (1.0 / 4294967296.0); }

/* standardNormal, utils.hpp:648-655. */
double orc_normal(std::mt19937& g)
{
    double u = orc_uniform(g);
    double v = orc_uniform(g);
    double c = std::sqrt(-2 * std::log(u));
    return c * std::cos(2 * kPi * v);
}

/* rotateXY, utils.hpp:791-801: the 3x3 rotation about z applied to n. */
void orc_rotate_xy(const double* n, double theta, double* out)
{
    double rot[3][3] = {{std::cos(theta), -std::sin(theta), 0},
                        {std::sin(theta), std::cos(theta), 0},
                        {0, 0, 1}};
    for (int r = 0; r < 3; ++r) out[r] = (rot[r][0] * n[0] + rot[r][1] * n[1]) + rot[r][2] * n[2];
}

/* rotateOutOfXY, utils.hpp:814-829: Rodrigues rotation about n x e_z. */
void orc_rotate_out_of_xy(const double* n, double theta, double* out)
{
    const double z[3] = {0, 0, 1};
    double v[3] = {n[1] * z[2] - n[2] * z[1], n[2] * z[0] - n[0] * z[2], n[0] * z[1] - n[1] * z[0]};
    double nrm = norm3(v[0], v[1], v[2]);
    for (double& x : v) x /= nrm;
    double w[3] = {v[1] * n[2] - v[2] * n[1], v[2] * n[0] - v[0] * n[2], v[0] * n[1] - v[1] * n[0]};
    for (int k = 0; k < 3; ++k) out[k] = n[k] * std::cos(theta) + w[k] * std::sin(theta);
}

/* minDistToCell, growth.hpp:128-162: brute force over all other cells. */
double orc_min_dist_to_cell(const double* cells, int n, int ld, int i)
{
    double ri[3], ni[3], hli;
    cell_geom(cells, ld, i, ri, ni, &hli);
    double mindist = INFINITY;
    for (int j = 0; j < n; ++j) {
        if (i == j) continue;
        double rj[3], nj[3], hlj, d[3], si, sj;
        cell_geom(cells, ld, j, rj, nj, &hlj);
        oracle_dist_between_cells(ri, ni, hli, rj, nj, hlj, d, &si, &sj);
        double dist = norm3(d[0], d[1], d[2]);
        if (dist < mindist) mindist = dist;
    }
    return mindist;
}

/* divideCells, growth.hpp:456-720, on a column-major population with leading
 * dimension cap; daughters are appended. Returns the number of tries. */
int orc_divide(const bfm_sim_config* cfg, std::mt19937& rng, double t, double* cells, int& n, int cap,
               int ncols, std::vector<int>& parents)
{
    std::vector<int> idx;
    for (int i = 0; i < n; ++i)
        if (cells[(int64_t)C_L * cap + i] > cfg->Ldiv) idx.push_back(i);   // divideMaxLength
    const int nd = (int)idx.size();
    if (nd == 0) return 0;
    const double R = cfg->R, mindist_default = 2 * cfg->Rcell;
    auto growth = [&](int group) {
        return cfg->growth_mean[group - 1] + cfg->growth_std[group - 1] * orc_normal(rng);
    };
    auto dlen = [&] {
        double r = 0.5 + cfg->daughter_length_std * orc_normal(rng);
        return r < 0 ? 0.0 : (r > 1 ? 1.0 : r);
    };
    auto axy = [&] {
        double r = orc_uniform(rng);
        return -cfg->daughter_angle_xy_bound + 2 * cfg->daughter_angle_xy_bound * r;
    };
    auto az = [&] {
        double r = orc_uniform(rng);
        return -cfg->daughter_angle_z_bound + 2 * cfg->daughter_angle_z_bound * r;
    };
    std::vector<int> check(nd, 1), satisfies(nd, 0);
    for (int m = 0; m < nd; ++m)
        if (orc_min_dist_to_cell(cells, n, cap, idx[m]) < mindist_default) check[m] = 0;
    for (int m = 0; m < nd; ++m)
        if (check[m] == 0) satisfies[m] = 1;
    // cells_total: n + nd rows, leading dimension n + nd
    const int nt = n + nd;
    std::vector<double> total((size_t)nt * ncols);
    auto T = [&](int i, int c) -> double& { return total[(size_t)c * nt + i]; };
    int ntries = 0, sat = 0;
    for (int v : satisfies) sat += v;
    while (ntries == 0 || (ntries < 10 && sat < nd)) {
        for (int c = 0; c < ncols; ++c)
            for (int i = 0; i < n; ++i) T(i, c) = cells[(int64_t)c * cap + i];
        for (int m = 0; m < nd; ++m)
            for (int c = 0; c < ncols; ++c) T(n + m, c) = cells[(int64_t)c * cap + idx[m]];
        std::vector<double> txy1(nd, 0), tz1(nd, 0), txy2(nd, 0), tz2(nd, 0);
        for (int m = 0; m < nd; ++m)
            if (ntries < 9 && check[m] == 1) {
                txy1[m] = axy(); tz1[m] = az(); txy2[m] = axy(); tz2[m] = az();
            }
        for (int m = 0; m < nd; ++m) {
            double u[3] = {T(idx[m], C_NX), T(idx[m], C_NY), T(idx[m], C_NZ)}, w[3], v[3];
            orc_rotate_xy(u, txy1[m], w);
            orc_rotate_out_of_xy(w, tz1[m], v);
            for (int k = 0; k < 3; ++k) T(idx[m], C_NX + k) = v[k];
            orc_rotate_xy(u, txy2[m], w);
            orc_rotate_out_of_xy(w, tz2[m], v);
            for (int k = 0; k < 3; ++k) T(n + m, C_NX + k) = v[k];
        }
        std::vector<double> M(nd);
        for (int m = 0; m < nd; ++m) M[m] = dlen();
        for (int m = 0; m < nd; ++m) {
            const int i = idx[m], d = n + m;
            double Ld = T(i, C_L) - 2 * R;
            double L1 = M[m] * Ld, L2 = (1 - M[m]) * Ld;
            double div = -T(i, C_HALF_L) + L1 + R;
            double delta1 = L1 / 2 + R, delta2 = L2 / 2 + R;
            T(i, C_L) = L1; T(i, C_HALF_L) = L1 / 2;
            T(d, C_L) = L2; T(d, C_HALF_L) = L2 / 2;
            for (int k = 0; k < 3; ++k) {
                T(i, C_RX + k) = T(i, C_RX + k) + (div - delta1) * T(i, C_NX + k);
                T(d, C_RX + k) = T(d, C_RX + k) + (div + delta2) * T(d, C_NX + k);
            }
            T(i, C_T0) = t;
            T(d, C_T0) = t;
        }
        for (int m = 0; m < nd; ++m) T(idx[m], C_GROWTH) = growth((int)T(idx[m], C_GROUP));
        for (int m = 0; m < nd; ++m) T(n + m, C_GROWTH) = growth((int)T(n + m, C_GROUP));
        for (int m = 0; m < nd; ++m) {
            if (check[m] != 1) continue;
            double d1 = orc_min_dist_to_cell(total.data(), nt, nt, idx[m]);
            double d2 = orc_min_dist_to_cell(total.data(), nt, nt, n + m);
            if (d1 >= mindist_default && d2 >= mindist_default) satisfies[m] = 1;
        }
        sat = 0;
        for (int v : satisfies) sat += v;
        ntries++;
    }
    int max_id = (int)parents.size() - 1;
    for (int m = 0; m < nd; ++m) {
        int parent_id = (int)T(idx[m], C_ID);
        T(idx[m], C_ID) = max_id + (2 * m + 1);
        T(n + m, C_ID) = max_id + (2 * m + 2);
        parents.push_back(parent_id);
        parents.push_back(parent_id);
    }
    for (int c = 0; c < ncols; ++c)
        for (int i = 0; i < nt; ++i) cells[(int64_t)c * cap + i] = T(i, c);
    n = nt;
    return ntries;
}

}  // namespace

extern "C" {

int oracle_divide_cells(const bfm_sim_config* cfg, double* cells, int* n, int cap, int ncols, double t,
                        uint64_t seed, int* parents, int* n_parents, int parents_cap, int* tries_out)
{
    int nd = 0;
    for (int i = 0; i < *n; ++i) nd += cells[(int64_t)C_L * cap + i] > cfg->Ldiv;
    if (*n + nd > cap || *n_parents + 2 * nd > parents_cap) return BFM_ERR_CAPACITY;
    std::mt19937 rng((uint32_t)seed);
    std::vector<int> par(parents, parents + *n_parents);
    int tries = orc_divide(cfg, rng, t, cells, *n, cap, ncols, par);
    std::copy(par.begin(), par.end(), parents);
    *n_parents = (int)par.size();
    if (tries_out) *tries_out = tries;
    return BFM_OK;
}

namespace {
/* Butcher tableaux of runSimulation (simulation.hpp:1207-1268 of ABM-3d;
 * Bogacki-Shampine is also the 2-D tree's fixed method, simulation.hpp:216-228
 * of ABM-2d); -1 is velocity Verlet (no tableau, Heun-Euler placeholders).
 * Returns false for an unknown id. */
bool orc_tableau(int which, const double*& A, const double*& b, const double*& bs, int& s, double& order)
{
    static const double HE_A[4] = {0, 0, 1, 0}, HE_b[2] = {0.5, 0.5}, HE_bs[2] = {1, 0};
    static const double BS_A[16] = {0, 0, 0, 0, 1. / 2., 0, 0, 0, 0, 3. / 4., 0, 0, 2. / 9., 1. / 3., 4. / 9., 0};
    static const double BS_b[4] = {2. / 9., 1. / 3., 4. / 9., 0}, BS_bs[4] = {7. / 24., 1. / 4., 1. / 3., 1. / 8.};
    static const double RKF_A[36] = {0, 0, 0, 0, 0, 0, 1. / 4., 0, 0, 0, 0, 0, 3. / 32., 9. / 32., 0, 0, 0, 0,
                                     1932. / 2197., -7200. / 2197., 7296. / 2197., 0, 0, 0,
                                     439. / 216., -8.0, 3680. / 513., -845. / 4104., 0, 0,
                                     -8. / 27., 2.0, -3544. / 2565., 1859. / 4104., -11. / 40., 0};
    static const double RKF_b[6] = {16. / 135., 0, 6656. / 12825., 28561. / 56430., -9. / 50., 2. / 55.};
    static const double RKF_bs[6] = {25. / 216., 0, 1408. / 2565., 2197. / 4104., -1. / 5., 0};
    static const double DP_A[49] = {0, 0, 0, 0, 0, 0, 0, 1. / 5., 0, 0, 0, 0, 0, 0, 3. / 40., 9. / 40., 0, 0, 0, 0, 0,
                                    44. / 45., -56. / 15., 32. / 9., 0, 0, 0, 0,
                                    19372. / 6561., -25360. / 2187., 64448. / 6561., -212. / 729, 0, 0, 0,
                                    9017. / 3168., -355. / 33., 46732. / 5247., 49. / 176., -5103. / 18656., 0, 0,
                                    35. / 384., 0, 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84., 0};
    static const double DP_b[7] = {35. / 384., 0, 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84., 0};
    static const double DP_bs[7] = {5179. / 57600., 0, 7571. / 16695., 393. / 640., -92097. / 339200.,
                                    187. / 2100., 1. / 40.};
    switch (which) {
    case 0: A = HE_A; b = HE_b; bs = HE_bs; s = 2; order = 1; break;
    case 1: A = BS_A; b = BS_b; bs = BS_bs; s = 4; order = 2; break;
    case 2: A = RKF_A; b = RKF_b; bs = RKF_bs; s = 6; order = 4; break;
    case 3: A = DP_A; b = DP_b; bs = DP_bs; s = 7; order = 4; break;
    case -1: A = HE_A; b = HE_b; bs = HE_bs; s = 2; order = 1; break;   // velocity Verlet (no tableau)
    default: return false;
    }
    return true;
}
}  // namespace

/* runSimulation's loop, simulation.hpp:1616-1864 (noise, switching, basal
 * pruning and file output off): divide, rebuild the neighbor list after a
 * division, adaptive RK step with the controller, grow, refresh the neighbor
 * distances, advance t and iter, periodic rebuild. The neighbor list comes
 * from the binned builder (the same list as getCellNeighbors). */
int oracle_simulate(const bfm_sim_config* cfg, const bfm_params* p, const double* cells_in, int n_in,
                    double* cells_out, int cap, int* n_out, double* dt_trace, int cap_trace,
                    int64_t* iters_out, int64_t* rk_steps_out, int nthreads)
{
    const int ncols = p->ncols;
    const double *A, *b, *bs;
    int s;
    double order;
    if (!orc_tableau(cfg->tableau, A, b, bs, s, order)) return BFM_ERR_ARG;
    const bool verlet = cfg->tableau == -1;
    if (verlet && !(cfg->rho > 0)) return BFM_ERR_ARG;
    if (n_in < 1 || n_in > cap || cfg->iter_update_neighbors < 1 || cfg->iter_update_stepsize < 1)
        return BFM_ERR_ARG;
    std::vector<double> cells((size_t)cap * ncols, 0.0), cells_new((size_t)cap * ncols), errors;
    for (int c = 0; c < ncols; ++c)
        for (int i = 0; i < n_in; ++i) cells[(size_t)c * cap + i] = cells_in[(size_t)c * n_in + i];
    int max_id = 0;
    for (int i = 0; i < n_in; ++i) max_id = std::max(max_id, (int)cells_in[(size_t)C_ID * n_in + i]);
    std::vector<int> parents(max_id + 1, -1);
    std::mt19937 rng((uint32_t)cfg->seed);
    int n = n_in;
    int64_t iter = 0, rk_steps = 0;
    double t = 0, dt = cfg->max_stepsize;
    std::vector<double> nb;
    int64_t m = 0, ld_nb = 0;
    auto build = [&] {
        // compact copy with ld = n for the builder
        int64_t need = oracle_get_cell_neighbors_binned(cells.data(), n, cap, cfg->neighbor_threshold, cfg->R,
                                                        cfg->Ldiv, nullptr, 0, nthreads);
        ld_nb = std::max<int64_t>(need, 1);
        nb.assign((size_t)ld_nb * 7, 0.0);
        m = oracle_get_cell_neighbors_binned(cells.data(), n, cap, cfg->neighbor_threshold, cfg->R, cfg->Ldiv,
                                             nb.data(), ld_nb, nthreads);
    };
    std::vector<double> packed, packed_new;
    auto rk = [&](double h) {
        packed.assign((size_t)n * ncols, 0.0);
        for (int c = 0; c < ncols; ++c)
            std::memcpy(&packed[(size_t)c * n], &cells[(size_t)c * cap], sizeof(double) * n);
        packed_new.assign((size_t)n * ncols, 0.0);
        errors.assign((size_t)n * 6, 0.0);
        oracle_step_rk(A, b, bs, s, h, packed.data(), n, n, nb.data(), m, ld_nb, nullptr, p, packed_new.data(),
                       errors.data());
        ++rk_steps;
        return std::max(oracle_max_scaled_error(packed.data(), n, n, errors.data(), cfg->max_error_allowed, 0.0),
                        cfg->min_error);
    };
    auto terminate = [&] {
        bool hn = cfg->n_cells > 0, hi = cfg->max_iter > 0, ht = cfg->max_time > 0;
        if (!hn && !hi && !ht) return true;
        return (hn && n >= cfg->n_cells) || (hi && iter >= cfg->max_iter) || (ht && t >= cfg->max_time);
    };
    build();
    while (!terminate()) {
        int nd = 0;
        for (int i = 0; i < n; ++i) nd += cells[(size_t)C_L * cap + i] > cfg->Ldiv;
        if (nd > 0) {
            if (n + nd > cap) return BFM_ERR_CAPACITY;
            orc_divide(cfg, rng, t, cells.data(), n, cap, ncols, parents);
            build();
        }
        if (verlet) {   // stepVelocityVerlet at the fixed dt (simulation.hpp:1787-1798)
            packed.assign((size_t)n * ncols, 0.0);
            for (int c = 0; c < ncols; ++c)
                std::memcpy(&packed[(size_t)c * n], &cells[(size_t)c * cap], sizeof(double) * n);
            packed_new.assign((size_t)n * ncols, 0.0);
            oracle_step_verlet(dt, cfg->rho, packed.data(), n, n, nb.data(), m, ld_nb, nullptr, p,
                               packed_new.data());
            ++rk_steps;
        }
        double max_error = verlet ? 0.0 : rk(dt);
        if (!verlet && iter % cfg->iter_update_stepsize == 0) {
            bool exceeded = max_error > 1.0;
            double factor = 0.9 * std::pow(1.0 / max_error, 1.0 / (order + 1));
            if (factor >= 10) factor = 10;
            else if (factor < 0.2) factor = 0.2;
            int j = 0;
            while (exceeded && j < cfg->max_tries_update_stepsize) {
                max_error = rk(dt * factor);
                exceeded = max_error > 1.0;
                factor *= 0.9 * std::pow(1.0 / max_error, 1.0 / (order + 1));
                if (factor >= 10) { factor = 10; break; }
                else if (factor < 0.2) { factor = 0.2; break; }
                j++;
            }
            if (dt * factor < cfg->min_stepsize) factor = cfg->min_stepsize / dt;
            else if (dt * factor > cfg->max_stepsize) factor = cfg->max_stepsize / dt;
            dt *= factor;
            rk(dt);
        }
        for (int c = 0; c < ncols; ++c)
            std::memcpy(&cells[(size_t)c * cap], &packed_new[(size_t)c * n], sizeof(double) * n);
        for (int i = 0; i < n; ++i) {   // growCells, growth.hpp:51-58
            double& l = cells[(size_t)C_L * cap + i];
            l += (cells[(size_t)C_GROWTH * cap + i] * (4 * cfg->R / 3 + l)) * dt;
            cells[(size_t)C_HALF_L * cap + i] = l / 2;
        }
        oracle_update_neighbor_distances(cells.data(), n, cap, nb.data(), m, ld_nb);
        if (dt_trace && iter < cap_trace) dt_trace[iter] = dt;
        t += dt;
        iter++;
        if (iter % cfg->iter_update_neighbors == 0) build();
    }
    for (int c = 0; c < ncols; ++c)
        std::memcpy(&cells_out[(size_t)c * cap], &cells[(size_t)c * cap], sizeof(double) * n);
    *n_out = n;
    if (iters_out) *iters_out = iter;
    if (rk_steps_out) *rk_steps_out = rk_steps;
    return BFM_OK;
}

}  // extern "C"

/* ---- the 2-D model (ABM-2d/ABM-2d-cpp) -------------------------------------
 * Cells in the 2-D tree's layout (id 0, r 1-2, n 3-4, dr 5-6, dn 7-8, l 9,
 * half_l 10, eta0 13, eta1 14), neighbor rows (i, j, dx, dy, s, t). */
namespace {
enum { D2_RX = 1, D2_RY = 2, D2_NX = 3, D2_NY = 4, D2_DRX = 5, D2_L = 9, D2_HALF_L = 10, D2_ETA0 = 13, D2_ETA1 = 14 };

/* cellCellRepulsiveForces (mechanics.hpp:448-570 of ABM-2d) and
 * cellCellAdhesiveForces (:572-694) with forcesJKRLagrange<T, 2> /
 * forcesKiharaLagrange<T, 2> (adhesion.hpp:334-462 of ABM-2d). */
void forces_2d(const double* cells, int n, int ld, const double* nb, int64_t m, int64_t ld_nb,
               const int* to_adhere, const bfm2d_params* p, double* rep, double* adh)
{
    std::memset(rep, 0, sizeof(double) * 4 * (size_t)n);
    std::memset(adh, 0, sizeof(double) * 4 * (size_t)n);
    if (n == 1) return;
    const double R = p->R, Rcell = p->Rcell;
    for (int64_t k = 0; k < m; ++k) {
        const int i = (int)nb[k], j = (int)nb[ld_nb + k];
        const double dx = nb[2 * ld_nb + k], dy = nb[3 * ld_nb + k];
        const double si = nb[4 * ld_nb + k], sj = nb[5 * ld_nb + k];
        const double mag = std::sqrt(dx * dx + dy * dy);
        const double dir[2] = {dx / mag, dy / mag};
        const double overlap = 2 * R - mag;
        const double ni[2] = {cell(cells, ld, i, D2_NX), cell(cells, ld, i, D2_NY)};
        const double nj[2] = {cell(cells, ld, j, D2_NX), cell(cells, ld, j, D2_NY)};
        double prefactor = 0;
        if (overlap > 0 && overlap < 2 * (R - Rcell)) {
            prefactor = p->prefactors[0] * std::pow(overlap, 1.5);
        } else if (overlap >= 2 * (R - Rcell)) {
            double term = p->prefactors[2] * std::pow(overlap - 2 * (R - Rcell), 1.5);
            prefactor = p->prefactors[1] + term;
        }
        if (overlap > 0) {
            const double vij[2] = {prefactor * dir[0], prefactor * dir[1]};
            const double wi = ni[0] * vij[0] + ni[1] * vij[1];
            const double wj = nj[0] * vij[0] + nj[1] * vij[1];
            const double fi[4] = {vij[0], vij[1], si * (-wi * ni[0] + vij[0]), si * (-wi * ni[1] + vij[1])};
            const double fj[4] = {-vij[0], -vij[1], sj * (wj * nj[0] - vij[0]), sj * (wj * nj[1] - vij[1])};
            for (int q = 0; q < 4; ++q) {
                rep[(size_t)q * n + i] += fi[q];
                rep[(size_t)q * n + j] += fj[q];
            }
        }
    }
    if (p->adhesion_mode == BFM2D_ADHESION_NONE) return;
    for (int64_t k = 0; k < m; ++k) {
        const int i = (int)nb[k], j = (int)nb[ld_nb + k];
        const double d12[2] = {nb[2 * ld_nb + k], nb[3 * ld_nb + k]};
        const double s = nb[4 * ld_nb + k], t = nb[5 * ld_nb + k];
        const double overlap = 2 * R - std::sqrt(d12[0] * d12[0] + d12[1] * d12[1]);
        if (!(to_adhere[k] && overlap > 0)) continue;
        const double n1[2] = {cell(cells, ld, i, D2_NX), cell(cells, ld, i, D2_NY)};
        const double n2[2] = {cell(cells, ld, j, D2_NX), cell(cells, ld, j, D2_NY)};
        const double dist = std::sqrt(d12[0] * d12[0] + d12[1] * d12[1]);
        double f0[4] = {0, 0, 0, 0}, f1[4] = {0, 0, 0, 0};
        if (dist <= 2 * R) {
            const double d12n[2] = {d12[0] / dist, d12[1] / dist};
            double v[2];
            if (p->adhesion_mode == BFM2D_ADHESION_JKR) {
                double term = kPi * R;
                if (dist <= p->mindist) term *= (2 * R - p->mindist);
                else term *= (2 * R - dist);
                v[0] = term * d12n[0];
                v[1] = term * d12n[1];
            } else if (p->adhesion_mode == BFM2D_ADHESION_JKR_ISOTROPIC) {
                // forcesIsotropicJKRLagrange<T, 2> (adhesion.hpp:171-233 of
                // ABM-3d) with max_overlap = 2 (R - Rcell) (the 3-D call,
                // mechanics.hpp:1482-1487 of ABM-3d): v = (term2 - term1) d12n
                double delta = 2 * R - dist;
                if (delta > 2 * (R - Rcell)) delta = 2 * (R - Rcell);
                const double Req = R / 2;
                const double radius = oracle_jkr_contact_radius(delta, Req, p->E0, p->gamma, nullptr);
                const double a3 = radius * radius * radius;
                const double term1 = 4 * p->E0 * a3 / (3 * Req);
                const double term2 = 4 * std::sqrt(kPi * a3 * p->gamma * p->E0);
                v[0] = (term2 - term1) * d12n[0];
                v[1] = (term2 - term1) * d12n[1];
            } else {
                const double e = p->distance_exp;
                double term1 = (dist <= p->mindist ? 1.0 / std::pow(p->mindist, e + 1) : 1.0 / std::pow(dist, e + 1));
                double term2 = 1.0 / std::pow(2 * R, e + 1);
                const double a = e * (term1 - term2);
                v[0] = a * d12n[0];
                v[1] = a * d12n[1];
            }
            const double w1 = n1[0] * (-v[0]) + n1[1] * (-v[1]);
            const double w2 = n2[0] * (-v[0]) + n2[1] * (-v[1]);
            f0[0] = -v[0]; f0[1] = -v[1];
            f1[0] = v[0]; f1[1] = v[1];
            f0[2] = s * (-w1 * n1[0] - v[0]); f0[3] = s * (-w1 * n1[1] - v[1]);
            f1[2] = t * (w2 * n2[0] + v[0]); f1[3] = t * (w2 * n2[1] + v[1]);
        }
        for (int q = 0; q < 4; ++q) {
            adh[(size_t)q * n + i] += p->strength * f0[q];
            adh[(size_t)q * n + j] += p->strength * f1[q];
        }
    }
}

void normalize_2d(double* cells, int n, int ld)
{
    for (int i = 0; i < n; ++i) {
        const double nx = cells[(size_t)D2_NX * ld + i], ny = cells[(size_t)D2_NY * ld + i];
        const double nrm = std::sqrt(nx * nx + ny * ny);
        cells[(size_t)D2_NX * ld + i] = nx / nrm;
        cells[(size_t)D2_NY * ld + i] = ny / nrm;
    }
}
}  // namespace

extern "C" {

/* getCellNeighbors of the 2-D tree (mechanics.hpp:293-374 of ABM-2d): the
 * 3-D closest-point computation of oracle_dist_between_cells on the cells in
 * the z = 0 plane (the reference's non-parallel branch calls CGAL's segment
 * distance, CGAL::Distance_3::internal::squared_distance, not available here;
 * its parallel branch is the same as the 3-D tree's). Rows i, j, dx, dy, s, t. */
int64_t oracle2d_get_cell_neighbors(const double* cells, int n, int ld, double thr, double R, double Ldiv,
                                    double* out, int64_t ld_out)
{
    std::vector<double> c3((size_t)15 * n, 0.0);
    for (int i = 0; i < n; ++i) {
        c3[(size_t)0 * n + i] = cell(cells, ld, i, 0);
        c3[(size_t)C_RX * n + i] = cell(cells, ld, i, D2_RX);
        c3[(size_t)C_RY * n + i] = cell(cells, ld, i, D2_RY);
        c3[(size_t)C_NX * n + i] = cell(cells, ld, i, D2_NX);
        c3[(size_t)C_NY * n + i] = cell(cells, ld, i, D2_NY);
        c3[(size_t)C_L * n + i] = cell(cells, ld, i, D2_L);
        c3[(size_t)C_HALF_L * n + i] = cell(cells, ld, i, D2_HALF_L);
    }
    const int64_t cap = std::max<int64_t>(ld_out, 1);
    std::vector<double> nb7((size_t)7 * cap);
    const int64_t m = oracle_get_cell_neighbors(c3.data(), n, n, thr, R, Ldiv, nb7.data(), cap);
    const int src[6] = {0, 1, 2, 3, 5, 6};
    for (int q = 0; q < 6; ++q)
        for (int64_t k = 0; k < std::min(m, ld_out); ++k) out[q * ld_out + k] = nb7[(size_t)src[q] * cap + k];
    return m;
}

/* The repulsive and adhesive dE/dq (n x 4 each) of the 2-D tree. */
void oracle2d_forces(const double* cells, int n, int ld, const double* nb, int64_t m, int64_t ld_nb,
                     const int* to_adhere, const bfm2d_params* p, double* rep, double* adh)
{
    forces_2d(cells, n, ld, nb, m, ld_nb, to_adhere, p, rep, adh);
}

/* getVelocities of the 2-D tree (mechanics.hpp:1503-1708 of ABM-2d) without
 * confinement: (repulsion + adhesion + 0) + noise, divided by the composite
 * drag prefactors (compositeViscosityForcePrefactors, :264-291). */
void oracle2d_get_velocities(const double* cells, int n, int ld, const double* nb, int64_t m, int64_t ld_nb,
                             const int* to_adhere, const double* noise, const bfm2d_params* p, double* vel)
{
    std::vector<double> rep(4 * (size_t)n), adh(4 * (size_t)n);
    forces_2d(cells, n, ld, nb, m, ld_nb, to_adhere, p, rep.data(), adh.data());
    for (int i = 0; i < n; ++i) {
        const double l = cell(cells, ld, i, D2_L);
        const double drag = cell(cells, ld, i, D2_ETA0) + cell(cells, ld, i, D2_ETA1) * p->surface_contact_density / p->R;
        const double K = l * drag;
        const double L = l * l * l * drag / 12;
        for (int q = 0; q < 4; ++q) {
            double e = rep[(size_t)q * n + i] + adh[(size_t)q * n + i] + 0.0;
            e = e + (noise ? noise[(size_t)q * n + i] : 0.0);
            vel[(size_t)q * n + i] = -e / (q < 2 ? K : L);
        }
    }
}

/* stepRungeKuttaAdaptive of the 2-D tree (mechanics.hpp:1710-1817 of ABM-2d). */
void oracle2d_step_rk(const double* A, const double* b, const double* bs, int s, double dt, const double* cells,
                      int n, int ld, int ncols, const double* nb, int64_t m, int64_t ld_nb, const int* to_adhere,
                      const double* noise, const bfm2d_params* p, double* cells_new, double* errors)
{
    std::vector<std::vector<double>> k(s, std::vector<double>(4 * (size_t)n));
    std::vector<double> ci((size_t)ncols * n);
    auto copy_in = [&]() {
        for (int c = 0; c < ncols; ++c)
            for (int i = 0; i < n; ++i) ci[(size_t)c * n + i] = cell(cells, ld, i, c);
    };
    copy_in();
    oracle2d_get_velocities(ci.data(), n, n, nb, m, ld_nb, to_adhere, noise, p, k[0].data());
    for (int st = 1; st < s; ++st) {
        std::vector<double> mult(4 * (size_t)n, 0.0);
        for (int j = 0; j < st; ++j)
            for (size_t q = 0; q < 4 * (size_t)n; ++q) mult[q] += k[j][q] * A[st * s + j];
        copy_in();
        for (int c = 0; c < 4; ++c)
            for (int i = 0; i < n; ++i) ci[(size_t)(c + D2_RX) * n + i] += mult[(size_t)c * n + i] * dt;
        normalize_2d(ci.data(), n, n);
        oracle2d_get_velocities(ci.data(), n, n, nb, m, ld_nb, to_adhere, noise, p, k[st].data());
    }
    std::vector<double> v1(4 * (size_t)n, 0.0), v2(4 * (size_t)n, 0.0);
    for (int st = 0; st < s; ++st)
        for (size_t q = 0; q < 4 * (size_t)n; ++q) {
            v1[q] += k[st][q] * b[st];
            v2[q] += k[st][q] * bs[st];
        }
    for (int c = 0; c < ncols; ++c)
        for (int i = 0; i < n; ++i) cells_new[(size_t)c * n + i] = cell(cells, ld, i, c);
    for (size_t q = 0; q < 4 * (size_t)n; ++q) {
        const double d1 = v1[q] * dt, d2 = v2[q] * dt;
        cells_new[(size_t)D2_RX * n + q] += d1;
        errors[q] = d1 - d2;
        cells_new[(size_t)D2_DRX * n + q] = v1[q];
    }
    normalize_2d(cells_new, n, n);
}

}  // extern "C"

/* ---- the 2-D growth loop (ABM-2d/ABM-2d-cpp) -------------------------------- */
namespace {
enum { D2_ID = 0, D2_T0 = 11, D2_GROWTH = 12, D2_GROUP = 16 };

/* rotate, utils.hpp:373-380 of ABM-2d: the 2x2 rotation applied to n. */
void orc2d_rotate(const double* n, double theta, double* out)
{
    const double rot[2][2] = {{std::cos(theta), -std::sin(theta)}, {std::sin(theta), std::cos(theta)}};
    for (int r = 0; r < 2; ++r) out[r] = rot[r][0] * n[0] + rot[r][1] * n[1];
}

/* minDistToCell, growth.hpp:82-117 of ABM-2d: brute force over all other
 * cells, |d| of distBetweenCells on the cells in the z = 0 plane. */
double orc2d_min_dist_to_cell(const double* cells, int n, int ld, int i)
{
    const double ri[3] = {cell(cells, ld, i, D2_RX), cell(cells, ld, i, D2_RY), 0.0};
    const double ni[3] = {cell(cells, ld, i, D2_NX), cell(cells, ld, i, D2_NY), 0.0};
    const double hli = cell(cells, ld, i, D2_HALF_L);
    double mindist = INFINITY;
    for (int j = 0; j < n; ++j) {
        if (i == j) continue;
        const double rj[3] = {cell(cells, ld, j, D2_RX), cell(cells, ld, j, D2_RY), 0.0};
        const double nj[3] = {cell(cells, ld, j, D2_NX), cell(cells, ld, j, D2_NY), 0.0};
        double d[3], si, sj;
        oracle_dist_between_cells(ri, ni, hli, rj, nj, cell(cells, ld, j, D2_HALF_L), d, &si, &sj);
        const double dist = std::sqrt(d[0] * d[0] + d[1] * d[1]);
        if (dist < mindist) mindist = dist;
    }
    return mindist;
}

/* divideCells of the 2-D tree (growth.hpp:438-697 of ABM-2d; the overload
 * with one growth distribution per group) on a column-major population with
 * leading dimension cap; daughters are appended. Returns the number of tries. */
int orc2d_divide(const bfm_sim_config* cfg, std::mt19937& rng, double t, double* cells, int& n, int cap, int ncols,
                 std::vector<int>& parents)
{
    std::vector<int> idx;
    for (int i = 0; i < n; ++i)
        if (cells[(int64_t)D2_L * cap + i] > cfg->Ldiv) idx.push_back(i);   // divideMaxLength, :67-71
    const int nd = (int)idx.size();
    if (nd == 0) return 0;
    const double R = cfg->R, mindist_default = 2 * cfg->Rcell;
    auto growth = [&](int group) {   // simulation.hpp:267-280 of ABM-2d
        return cfg->growth_mean[group - 1] + cfg->growth_std[group - 1] * orc_normal(rng);
    };
    auto dlen = [&] {                // simulation.hpp:302-315
        double r = 0.5 + cfg->daughter_length_std * orc_normal(rng);
        return r < 0 ? 0.0 : (r > 1 ? 1.0 : r);
    };
    auto angle = [&] {               // simulation.hpp:316-324
        double r = orc_uniform(rng);
        return -cfg->daughter_angle_xy_bound + 2 * cfg->daughter_angle_xy_bound * r;
    };
    std::vector<int> check(nd, 1), satisfies(nd, 0);
    for (int m = 0; m < nd; ++m)
        if (orc2d_min_dist_to_cell(cells, n, cap, idx[m]) < mindist_default) check[m] = 0;
    for (int m = 0; m < nd; ++m)
        if (check[m] == 0) satisfies[m] = 1;
    const int nt = n + nd;
    std::vector<double> total((size_t)nt * ncols);
    auto T = [&](int i, int c) -> double& { return total[(size_t)c * nt + i]; };
    int ntries = 0, sat = 0;
    for (int v : satisfies) sat += v;
    while (ntries == 0 || (ntries < 10 && sat < nd)) {
        for (int c = 0; c < ncols; ++c)
            for (int i = 0; i < n; ++i) T(i, c) = cells[(int64_t)c * cap + i];
        for (int m = 0; m < nd; ++m)
            for (int c = 0; c < ncols; ++c) T(n + m, c) = cells[(int64_t)c * cap + idx[m]];
        std::vector<double> th1(nd, 0), th2(nd, 0);
        for (int m = 0; m < nd; ++m)
            if (ntries < 9 && check[m] == 1) {
                th1[m] = angle();
                th2[m] = angle();
            }
        for (int m = 0; m < nd; ++m) {
            const double u[2] = {T(idx[m], D2_NX), T(idx[m], D2_NY)};
            double v[2];
            orc2d_rotate(u, th1[m], v);
            T(idx[m], D2_NX) = v[0]; T(idx[m], D2_NY) = v[1];
            orc2d_rotate(u, th2[m], v);
            T(n + m, D2_NX) = v[0]; T(n + m, D2_NY) = v[1];
        }
        std::vector<double> M(nd);
        for (int m = 0; m < nd; ++m) M[m] = dlen();
        for (int m = 0; m < nd; ++m) {
            const int i = idx[m], d = n + m;
            const double Ld = T(i, D2_L) - 2 * R;
            const double L1 = M[m] * Ld, L2 = (1 - M[m]) * Ld;
            const double div = -T(i, D2_HALF_L) + L1 + R;
            const double delta1 = L1 / 2 + R, delta2 = L2 / 2 + R;
            T(i, D2_L) = L1; T(i, D2_HALF_L) = L1 / 2;
            T(d, D2_L) = L2; T(d, D2_HALF_L) = L2 / 2;
            T(i, D2_RX) = T(i, D2_RX) + (div - delta1) * T(i, D2_NX);
            T(d, D2_RX) = T(d, D2_RX) + (div + delta2) * T(d, D2_NX);
            T(i, D2_RY) = T(i, D2_RY) + (div - delta1) * T(i, D2_NY);
            T(d, D2_RY) = T(d, D2_RY) + (div + delta2) * T(d, D2_NY);
            T(i, D2_T0) = t;
            T(d, D2_T0) = t;
        }
        for (int m = 0; m < nd; ++m) T(idx[m], D2_GROWTH) = growth((int)T(idx[m], D2_GROUP));
        for (int m = 0; m < nd; ++m) T(n + m, D2_GROWTH) = growth((int)T(n + m, D2_GROUP));
        for (int m = 0; m < nd; ++m) {
            if (check[m] != 1) continue;
            const double d1 = orc2d_min_dist_to_cell(total.data(), nt, nt, idx[m]);
            const double d2 = orc2d_min_dist_to_cell(total.data(), nt, nt, n + m);
            if (d1 >= mindist_default && d2 >= mindist_default) satisfies[m] = 1;
        }
        sat = 0;
        for (int v : satisfies) sat += v;
        ntries++;
    }
    const int max_id = (int)parents.size() - 1;   // ids and lineage, growth.hpp:672-684
    for (int m = 0; m < nd; ++m) {
        const int parent_id = (int)T(idx[m], D2_ID);
        T(idx[m], D2_ID) = max_id + (2 * m + 1);
        T(n + m, D2_ID) = max_id + (2 * m + 2);
        parents.push_back(parent_id);
        parents.push_back(parent_id);
    }
    for (int c = 0; c < ncols; ++c)
        for (int i = 0; i < nt; ++i) cells[(int64_t)c * cap + i] = T(i, c);
    n = nt;
    return ntries;
}
}  // namespace

extern "C" {

int oracle2d_divide_cells(const bfm_sim_config* cfg, double* cells, int* n, int cap, int ncols, double t,
                          uint64_t seed, int* parents, int* n_parents, int parents_cap, int* tries_out)
{
    int nd = 0;
    for (int i = 0; i < *n; ++i) nd += cells[(int64_t)D2_L * cap + i] > cfg->Ldiv;
    if (*n + nd > cap || *n_parents + 2 * nd > parents_cap) return BFM_ERR_CAPACITY;
    std::mt19937 rng((uint32_t)seed);
    std::vector<int> par(parents, parents + *n_parents);
    const int tries = orc2d_divide(cfg, rng, t, cells, *n, cap, ncols, par);
    std::copy(par.begin(), par.end(), parents);
    *n_parents = (int)par.size();
    if (tries_out) *tries_out = tries;
    return BFM_OK;
}

/* runSimulationAdaptiveLagrangian's loop of the 2-D tree (simulation.hpp:
 * 549-765 of ABM-2d; no noise, switching, confinement, growth void or file
 * output): divide, rebuild the list and to_adhere after a division (:551-625),
 * adaptive RK step with the controller (:627-729), grow (growth.hpp:51-57),
 * updateNeighborDistances, advance t and iter, periodic rebuild (:742-764).
 * The list comes from the binned builder on the cells placed in the z = 0
 * plane (the list of getCellNeighbors, oracle2d_get_cell_neighbors). */
int oracle2d_simulate(const bfm_sim_config* cfg, const bfm2d_params* p, const int* amap, int n_groups,
                      const double* cells_in, int n_in, double* cells_out, int cap, int* n_out, int* parents_out,
                      int* n_parents_out, double* dt_trace, int cap_trace, int64_t* iters_out,
                      int64_t* rk_steps_out)
{
    const int ncols = p->ncols;
    const double *A, *b, *bs;
    int s;
    double order;
    if (cfg->tableau < 0 || !orc_tableau(cfg->tableau, A, b, bs, s, order)) return BFM_ERR_ARG;
    if (n_in < 1 || n_in > cap || ncols < 17 || cfg->iter_update_neighbors < 1 || cfg->iter_update_stepsize < 1)
        return BFM_ERR_ARG;
    std::vector<double> cells((size_t)cap * ncols, 0.0);
    for (int c = 0; c < ncols; ++c)
        for (int i = 0; i < n_in; ++i) cells[(size_t)c * cap + i] = cells_in[(size_t)c * n_in + i];
    int max_id = 0;
    for (int i = 0; i < n_in; ++i) max_id = std::max(max_id, (int)cells_in[i]);
    std::vector<int> parents(max_id + 1, -1);
    std::mt19937 rng((uint32_t)cfg->seed);
    int n = n_in;
    int64_t iter = 0, rk_steps = 0;
    double t = 0, dt = cfg->max_stepsize;
    const double R = cfg->R, Rcell = cfg->Rcell;
    std::vector<double> nb6, c3;
    std::vector<int> to_adhere;
    int64_t m = 0;
    auto embed = [&] {
        c3.assign((size_t)15 * n, 0.0);
        for (int i = 0; i < n; ++i) {
            c3[(size_t)0 * n + i] = cells[i];
            c3[(size_t)C_RX * n + i] = cells[(size_t)D2_RX * cap + i];
            c3[(size_t)C_RY * n + i] = cells[(size_t)D2_RY * cap + i];
            c3[(size_t)C_NX * n + i] = cells[(size_t)D2_NX * cap + i];
            c3[(size_t)C_NY * n + i] = cells[(size_t)D2_NY * cap + i];
            c3[(size_t)C_L * n + i] = cells[(size_t)D2_L * cap + i];
            c3[(size_t)C_HALF_L * n + i] = cells[(size_t)D2_HALF_L * cap + i];
        }
    };
    auto build = [&] {   // getCellNeighbors + the driver's to_adhere rule (:242-262)
        embed();
        const int64_t need = oracle_get_cell_neighbors_binned(c3.data(), n, n, cfg->neighbor_threshold, R,
                                                              cfg->Ldiv, nullptr, 0, 1);
        const int64_t ld = std::max<int64_t>(need, 1);
        std::vector<double> nb7((size_t)ld * 7, 0.0);
        m = oracle_get_cell_neighbors_binned(c3.data(), n, n, cfg->neighbor_threshold, R, cfg->Ldiv, nb7.data(), ld, 1);
        nb6.assign((size_t)std::max<int64_t>(m, 1) * 6, 0.0);
        const int src[6] = {0, 1, 2, 3, 5, 6};
        for (int q = 0; q < 6; ++q)
            for (int64_t k = 0; k < m; ++k) nb6[(size_t)q * std::max<int64_t>(m, 1) + k] = nb7[(size_t)src[q] * ld + k];
        to_adhere.assign(std::max<int64_t>(m, 1), 0);
        const int64_t L = std::max<int64_t>(m, 1);
        for (int64_t k = 0; k < m; ++k) {
            const int i = (int)nb6[k], j = (int)nb6[L + k];
            const int gi = (int)cells[(size_t)D2_GROUP * cap + i], gj = (int)cells[(size_t)D2_GROUP * cap + j];
            const int g0 = std::min(gi, gj), g1 = std::max(gi, gj);
            const double dx = nb6[2 * L + k], dy = nb6[3 * L + k];
            const double dist = std::sqrt(dx * dx + dy * dy);
            const bool in_map = amap && g0 >= 0 && g1 < n_groups && amap[g0 * n_groups + g1] != 0;
            to_adhere[k] = in_map && dist > R + Rcell && dist < 2 * R;
        }
    };
    auto update_distances = [&] {   // updateNeighborDistances (mechanics.hpp:376-446 of ABM-2d)
        embed();
        const int64_t L = std::max<int64_t>(m, 1);
        std::vector<double> nb7((size_t)L * 7, 0.0);
        const int dst[6] = {0, 1, 2, 3, 5, 6};
        for (int q = 0; q < 6; ++q)
            for (int64_t k = 0; k < m; ++k) nb7[(size_t)dst[q] * L + k] = nb6[(size_t)q * L + k];
        oracle_update_neighbor_distances(c3.data(), n, n, nb7.data(), m, L);
        for (int q = 0; q < 6; ++q)
            for (int64_t k = 0; k < m; ++k) nb6[(size_t)q * L + k] = nb7[(size_t)dst[q] * L + k];
    };
    std::vector<double> packed, packed_new, errors;
    auto rk = [&](double h) {   // stepRungeKuttaAdaptive + the scaled error (:639-650)
        packed.assign((size_t)n * ncols, 0.0);
        for (int c = 0; c < ncols; ++c)
            std::memcpy(&packed[(size_t)c * n], &cells[(size_t)c * cap], sizeof(double) * n);
        packed_new.assign((size_t)n * ncols, 0.0);
        errors.assign((size_t)n * 4, 0.0);
        const int64_t L = std::max<int64_t>(m, 1);
        oracle2d_step_rk(A, b, bs, s, h, packed.data(), n, n, ncols, nb6.data(), m, L, to_adhere.data(), nullptr, p,
                         packed_new.data(), errors.data());
        ++rk_steps;
        double mx = -INFINITY;
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < n; ++i) {
                const double z = 1.0 + std::fabs(packed[(size_t)(D2_RX + q) * n + i]);
                mx = std::max(mx, errors[(size_t)q * n + i] / (cfg->max_error_allowed * z));
            }
        return std::max(mx, cfg->min_error);
    };
    auto terminate = [&] {
        bool hn = cfg->n_cells > 0, hi = cfg->max_iter > 0, ht = cfg->max_time > 0;
        if (!hn && !hi && !ht) return true;
        return (hn && n >= cfg->n_cells) || (hi && iter >= cfg->max_iter) || (ht && t >= cfg->max_time);
    };
    build();
    while (!terminate()) {
        int nd = 0;
        for (int i = 0; i < n; ++i) nd += cells[(size_t)D2_L * cap + i] > cfg->Ldiv;
        if (nd > 0) {
            if (n + nd > cap) return BFM_ERR_CAPACITY;
            orc2d_divide(cfg, rng, t, cells.data(), n, cap, ncols, parents);
            build();
        }
        double max_error = rk(dt);
        if (iter % cfg->iter_update_stepsize == 0) {
            bool exceeded = max_error > 1.0;
            double factor = 0.9 * std::pow(1.0 / max_error, 1.0 / (order + 1));
            if (factor >= 10) factor = 10;
            else if (factor < 0.2) factor = 0.2;
            int j = 0;
            while (exceeded && j < cfg->max_tries_update_stepsize) {
                max_error = rk(dt * factor);
                exceeded = max_error > 1.0;
                factor *= 0.9 * std::pow(1.0 / max_error, 1.0 / (order + 1));
                if (factor >= 10) { factor = 10; break; }
                else if (factor < 0.2) { factor = 0.2; break; }
                j++;
            }
            if (dt * factor < cfg->min_stepsize) factor = cfg->min_stepsize / dt;
            else if (dt * factor > cfg->max_stepsize) factor = cfg->max_stepsize / dt;
            dt *= factor;
            rk(dt);
        }
        for (int c = 0; c < ncols; ++c)
            std::memcpy(&cells[(size_t)c * cap], &packed_new[(size_t)c * n], sizeof(double) * n);
        for (int i = 0; i < n; ++i) {   // growCells, growth.hpp:51-57 of ABM-2d
            double& l = cells[(size_t)D2_L * cap + i];
            l += (cells[(size_t)D2_GROWTH * cap + i] * (4 * R / 3 + l)) * dt;
            cells[(size_t)D2_HALF_L * cap + i] = l / 2;
        }
        update_distances();
        if (dt_trace && iter < cap_trace) dt_trace[iter] = dt;
        t += dt;
        iter++;
        if (iter % cfg->iter_update_neighbors == 0) build();
    }
    for (int c = 0; c < ncols; ++c)
        std::memcpy(&cells_out[(size_t)c * cap], &cells[(size_t)c * cap], sizeof(double) * n);
    *n_out = n;
    if (parents_out) std::copy(parents.begin(), parents.end(), parents_out);
    if (n_parents_out) *n_parents_out = (int)parents.size();
    if (iters_out) *iters_out = iter;
    if (rk_steps_out) *rk_steps_out = rk_steps;
    return BFM_OK;
}

}  // extern "C"

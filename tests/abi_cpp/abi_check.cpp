// abi_check.cpp -- the C ABI (include/bfm.h) driven from compiled C++ with
// plain column-major buffers, the way the reference's driver would call it
// through a thin adapter (a pointer + leading dimension per array instead of
// an Eigen array): getCellNeighbors, getVelocities and stepRungeKuttaAdaptive
// (mechanics.hpp:520-522, 2011-2028, 2406-2427) on a small dome, compared
// with the CPU oracle (test infrastructure) on the same buffers.
//
//   abi_check --layout   prints sizeof / offsetof of bfm_params, bfm_sim_config
//                        and bfm_sim_stats as JSON (no GPU work)
//   abi_check [n]        runs the three calls on an n-cell dome (default 3000);
//                        exit status 0 when every comparison holds
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bfm.h"
#include "../../oracle/oracle.h"

extern "C" int bfm_make_dome(int n, int ncols, double R, double Rcell, double phi, double hl_lo, double hl_hi,
                             double gamma, double eta0, double eta1, double sigma0, double growth, uint64_t seed,
                             double* cells, double* dome_radius);

namespace {

// Column-major n x cols array with leading dimension n (Eigen's default layout).
struct ColMajor {
    int rows = 0, cols = 0;
    std::vector<double> v;
    ColMajor() = default;
    ColMajor(int r, int c) : rows(r), cols(c), v((size_t)r * c, 0.0) {}
    double& operator()(int i, int j) { return v[(size_t)j * rows + i]; }
    double operator()(int i, int j) const { return v[(size_t)j * rows + i]; }
    double* data() { return v.data(); }
    const double* data() const { return v.data(); }
};

// The adapter a maintainer adds on the reference side: one context, the
// reference's argument order, exceptions on a non-zero status.
struct BfmContext {
    bfm_ctx* ctx = nullptr;
    BfmContext(int max_cells, int64_t max_pairs, const bfm_params& p)
    {
        if (bfm_create(&ctx, 0, max_cells, max_pairs) != BFM_OK)
            throw std::runtime_error(std::string("bfm_create: ") + (ctx ? bfm_last_error(ctx) : ""));
        check(bfm_set_params(ctx, &p));
    }
    ~BfmContext() { bfm_destroy(ctx); }
    void check(int rc) const
    {
        if (rc != BFM_OK) throw std::runtime_error(bfm_last_error(ctx));
    }
};

ColMajor getCellNeighbors(BfmContext& b, const ColMajor& cells, double thr, double R, double Ldiv)
{
    ColMajor nb(16 * cells.rows + 16, 7);
    int64_t m = 0;
    int rc = bfm_get_cell_neighbors(b.ctx, cells.data(), cells.rows, cells.rows, thr, R, Ldiv, nb.data(),
                                    nb.rows, &m);
    if (rc == BFM_ERR_CAPACITY) {
        nb = ColMajor((int)m, 7);
        rc = bfm_get_cell_neighbors(b.ctx, cells.data(), cells.rows, cells.rows, thr, R, Ldiv, nb.data(),
                                    nb.rows, &m);
    }
    b.check(rc);
    ColMajor out((int)m, 7);
    for (int c = 0; c < 7; ++c)
        for (int64_t k = 0; k < m; ++k) out((int)k, c) = nb((int)k, c);
    return out;
}

ColMajor getVelocities(BfmContext& b, const ColMajor& cells, const ColMajor& nb, const std::vector<int>& assume_2d)
{
    ColMajor vel(cells.rows, 6);
    b.check(bfm_get_velocities(b.ctx, cells.data(), cells.rows, cells.rows, nb.data(), nb.rows, nb.rows,
                               assume_2d.data(), nullptr, vel.data()));
    return vel;
}

void stepRungeKuttaAdaptive(BfmContext& b, const std::vector<double>& A_rowmajor, const std::vector<double>& bw,
                            const std::vector<double>& bs, const ColMajor& cells, const ColMajor& nb, double dt,
                            ColMajor& cells_new, ColMajor& errors)
{
    cells_new = ColMajor(cells.rows, cells.cols);
    errors = ColMajor(cells.rows, 6);
    b.check(bfm_step_rk(b.ctx, A_rowmajor.data(), bw.data(), bs.data(), (int)bw.size(), dt, cells.data(),
                        cells.rows, cells.rows, nb.data(), nb.rows, nb.rows, nullptr, cells_new.data(),
                        errors.data()));
}

// max over columns c0..c1 of max_i |a - b| / max_i |b| (normwise per column)
double rel_cols(const ColMajor& a, const ColMajor& b, int c0, int c1)
{
    double worst = 0;
    for (int c = c0; c < c1; ++c) {
        double scale = 1e-300, diff = 0;
        for (int i = 0; i < a.rows; ++i) {
            scale = std::fmax(scale, std::fabs(b(i, c)));
            diff = std::fmax(diff, std::fabs(a(i, c) - b(i, c)));
        }
        worst = std::fmax(worst, diff / scale);
    }
    return worst;
}

void print_layout()
{
#define F(S, m) std::printf("\"%s.%s\": %zu, ", #S, #m, offsetof(S, m))
    std::printf("{");
    std::printf("\"sizeof(bfm_params)\": %zu, ", sizeof(bfm_params));
    F(bfm_params, R); F(bfm_params, Rcell); F(bfm_params, E0); F(bfm_params, repulsion_prefactors);
    F(bfm_params, nz_threshold); F(bfm_params, cell_cell_coulomb_coeff); F(bfm_params, cell_surface_coulomb_coeff);
    F(bfm_params, no_surface); F(bfm_params, adhesion_mode); F(bfm_params, jkr_mode); F(bfm_params, colidx_gamma);
    F(bfm_params, friction_mode); F(bfm_params, ncols); F(bfm_params, colidx_eta_cell_cell);
    std::printf("\"sizeof(bfm_sim_config)\": %zu, ", sizeof(bfm_sim_config));
    F(bfm_sim_config, R); F(bfm_sim_config, Rcell); F(bfm_sim_config, Ldiv); F(bfm_sim_config, max_stepsize);
    F(bfm_sim_config, min_stepsize); F(bfm_sim_config, max_error_allowed); F(bfm_sim_config, min_error);
    F(bfm_sim_config, neighbor_threshold); F(bfm_sim_config, growth_mean); F(bfm_sim_config, growth_std);
    F(bfm_sim_config, daughter_length_std); F(bfm_sim_config, daughter_angle_xy_bound);
    F(bfm_sim_config, daughter_angle_z_bound); F(bfm_sim_config, max_time); F(bfm_sim_config, iter_update_neighbors);
    F(bfm_sim_config, iter_update_stepsize); F(bfm_sim_config, max_tries_update_stepsize); F(bfm_sim_config, n_cells);
    F(bfm_sim_config, max_iter); F(bfm_sim_config, tableau); F(bfm_sim_config, seed);
    F(bfm_sim_config, report_seconds); F(bfm_sim_config, max_wall_seconds); F(bfm_sim_config, rho);
    std::printf("\"sizeof(bfm2d_params)\": %zu, ", sizeof(bfm2d_params));
    F(bfm2d_params, R); F(bfm2d_params, Rcell); F(bfm2d_params, prefactors);
    F(bfm2d_params, surface_contact_density); F(bfm2d_params, adhesion_mode); F(bfm2d_params, ncols);
    F(bfm2d_params, strength); F(bfm2d_params, mindist); F(bfm2d_params, distance_exp); F(bfm2d_params, E0);
    F(bfm2d_params, gamma);
    std::printf("\"sizeof(bfm_sim_stats)\": %zu, ", sizeof(bfm_sim_stats));
    F(bfm_sim_stats, iterations); F(bfm_sim_stats, rk_steps); F(bfm_sim_stats, divisions);
    F(bfm_sim_stats, neighbor_builds); F(bfm_sim_stats, t_end); F(bfm_sim_stats, wall_seconds);
    std::printf("\"bfm_sim_stats.host_division_seconds\": %zu}\n", offsetof(bfm_sim_stats, host_division_seconds));
#undef F
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc > 1 && std::strcmp(argv[1], "--layout") == 0) {
        print_layout();
        return 0;
    }
    const int n = argc > 1 ? std::atoi(argv[1]) : 3000;
    const double R = 0.8, Rcell = 0.5, E0 = 3900.0, Ecell = 3.9e6, Ldiv = 2.4;
    const double thr = 0.5 * (2 * R + Ldiv);
    ColMajor cells(n, 22);
    double rd = 0;
    if (bfm_make_dome(n, 22, R, Rcell, 0.5, 0.4, 1.6, 100.0, 0.7, 720.0, 100.0, 1.5, 11, cells.data(), &rd) != n) {
        std::fprintf(stderr, "dome generator failed\n");
        return 2;
    }
    bfm_params p{};
    p.R = R;
    p.Rcell = Rcell;
    p.E0 = E0;
    p.repulsion_prefactors[0] = (4. / 3.) * E0 * std::sqrt(R / 2);        // simulation.hpp:1271-1274
    p.repulsion_prefactors[1] = (4. / 3.) * Ecell * std::sqrt(Rcell / 2);
    p.repulsion_prefactors[2] = (4. / 3.) * E0 * std::sqrt(R / 2) * std::pow(2 * (R - Rcell), 1.5);
    p.nz_threshold = 1e-8;
    p.cell_cell_coulomb_coeff = 1.0;
    p.cell_surface_coulomb_coeff = 1.0;
    p.no_surface = 0;
    p.adhesion_mode = BFM_ADHESION_JKR_ISOTROPIC;
    p.jkr_mode = BFM_JKR_EXACT;
    p.colidx_gamma = 21;
    p.friction_mode = 0;
    p.ncols = 22;
    p.colidx_eta_cell_cell = -1;
    int failures = 0;
    try {
        BfmContext b(n, 16 * (int64_t)n, p);
        // getCellNeighbors: identical to the oracle's list, element for element
        ColMajor nb = getCellNeighbors(b, cells, thr, R, Ldiv);
        ColMajor ref(16 * n + 16, 7);
        const int64_t m_ref = oracle_get_cell_neighbors(cells.data(), n, n, thr, R, Ldiv, ref.data(), ref.rows);
        bool same = m_ref == nb.rows;
        for (int c = 0; same && c < 7; ++c)
            for (int k = 0; same && k < nb.rows; ++k) same = nb(k, c) == ref(k, c);
        std::printf("getCellNeighbors: m = %d, oracle m = %lld, identical = %d\n", nb.rows, (long long)m_ref,
                    (int)same);
        failures += !same;
        // getVelocities: within 1e-10 of the oracle
        std::vector<int> a2d(n);
        for (int i = 0; i < n; ++i) a2d[i] = cells(i, 6) < p.nz_threshold;
        ColMajor vel = getVelocities(b, cells, nb, a2d);
        ColMajor vref(n, 6);
        oracle_get_velocities(cells.data(), n, n, nb.data(), nb.rows, nb.rows, a2d.data(), nullptr, &p, vref.data());
        const double ev = rel_cols(vel, vref, 0, 6);
        std::printf("getVelocities: rel err %.3e\n", ev);
        failures += !(ev <= 1e-10);
        // stepRungeKuttaAdaptive (Bogacki-Shampine): within 1e-10 of the oracle
        const std::vector<double> A = {0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0.75, 0, 0, 2. / 9., 1. / 3., 4. / 9., 0};
        const std::vector<double> bw = {2. / 9., 1. / 3., 4. / 9., 0};
        const std::vector<double> bs = {7. / 24., 1. / 4., 1. / 3., 1. / 8.};
        ColMajor cnew, err;
        stepRungeKuttaAdaptive(b, A, bw, bs, cells, nb, 1e-5, cnew, err);
        ColMajor cref(n, 22), eref(n, 6);
        oracle_step_rk(A.data(), bw.data(), bs.data(), 4, 1e-5, cells.data(), n, n, nb.data(), nb.rows, nb.rows,
                       nullptr, &p, cref.data(), eref.data());
        const double ec = rel_cols(cnew, cref, 1, 13);
        std::printf("stepRungeKuttaAdaptive: rel err %.3e\n", ec);
        failures += !(ec <= 1e-10);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    std::printf(failures ? "ABI CHECK FAILED\n" : "ABI CHECK OK\n");
    return failures ? 1 : 0;
}

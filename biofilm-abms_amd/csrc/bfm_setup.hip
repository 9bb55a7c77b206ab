// bfm_setup.hip -- host-side JKR setup of runSimulation: the surface energy
// density that gives a target cell-cell equilibrium distance
// (simulation.hpp:1354-1358, jkr.hpp:1310-1408).
//
// The reference solves the JKR quartic with 100-digit Aberth iterations; here
// the larger root is the FP64 monotone Newton of the force kernels
// (jkr_radius_newton), with its residual checked so that an overlap below the
// pull-off point (no admissible root, where the reference throws,
// jkr.hpp:99-105) is reported instead of iterated on.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/bfm.h"
#include "bfm_device.hpp"

using namespace bfm;

namespace {

struct NoRoot {};

// Larger JKR contact radius with the admissibility check of the quartic
// u^4 - p u - q at u^2 = a.
double contact_radius(double delta, double Req, double E, double gamma)
{
    const double a = jkr_radius_newton(delta, Req, E, gamma, 0.0, 400, nullptr);
    const double p = Req * std::sqrt(4 * kPi * gamma / E);
    const double q = Req * delta;
    const double u = std::sqrt(a);
    if (!(std::isfinite(a)) || (a * a - p * u) - q > 1e-9 * std::fmax(1.0, std::fabs(q)))
        throw NoRoot{};
    return a;
}

// newtonRaphson (rootFinding.hpp:42-75): Newton steps with the derivative
// from central differences of step dx, while |f| > tol and fewer than
// max_iter steps were taken. Returns x; *f_out receives f(x).
template <class F>
double newton_central(F&& f, double x0, double dx, double tol, int max_iter, double* f_out)
{
    double x = x0;
    double fx = f(x);
    int iter = 0;
    while (std::fabs(fx) > tol && iter < max_iter) {
        const double deriv = (f(x + dx) - f(x - dx)) / (2 * dx);
        const double update = -fx / deriv;
        x += update;
        fx = f(x);
        ++iter;
    }
    if (f_out) *f_out = fx;
    return x;
}

// jkrEquilibriumDistance (jkr.hpp:1310-1340): the overlap at which the JKR
// force of two cells of radius R (equivalent radius R / 2) vanishes, found by
// Newton from 2R - dinit and capped at 2R - dmin; returns 2R - overlap.
double equilibrium_distance(double R, double E0, double gamma, double d_init, double d_min,
                            double increment_overlap, double tol, int max_iter, double* f_out)
{
    const double overlap = 2 * R - d_init;
    const double max_overlap = 2 * R - d_min;
    const double Req = R / 2;
    const double prefactor1 = 4.0 / 3.0 * E0 / Req;
    const double prefactor2 = 4 * std::sqrt(kPi * gamma * E0);
    auto force = [&](double x) {
        const double radius = contact_radius(x, Req, E0, gamma);
        const double f_hertz = prefactor1 * radius * radius * radius;
        const double f_jkr = prefactor2 * std::pow(radius, 1.5);
        return f_hertz - f_jkr;
    };
    const double root = newton_central(force, overlap, increment_overlap, tol, max_iter, f_out);
    const double delta = dmin(root, max_overlap);   // std::min: a NaN root stays NaN
    return 2 * R - delta;
}

}  // namespace

extern "C" {

int bfm_jkr_equilibrium_distance(double R, double E0, double gamma, double d_init, double d_min,
                                 double increment_overlap, double newton_tol, int max_iter,
                                 double* deq_out, double* residual_out)
{
    if (!deq_out || !(R > 0) || !(E0 > 0) || !(gamma >= 0) || !(increment_overlap > 0) ||
        !(newton_tol >= 0) || max_iter < 0)
        return BFM_ERR_ARG;
    try {
        double f = 0;
        const double deq = equilibrium_distance(R, E0, gamma, d_init, d_min, increment_overlap,
                                                newton_tol, max_iter, &f);
        if (residual_out) *residual_out = f;
        if (!std::isfinite(deq) || !std::isfinite(f)) return BFM_ERR_NAN;
        *deq_out = deq;
        return BFM_OK;
    } catch (const NoRoot&) {
        return BFM_ERR_JKR_ROOT;
    }
}

int bfm_jkr_optimal_surface_energy_density(double R, double Rcell, double E0, double deq_target,
                                           double init_gamma, double log_increment_gamma,
                                           double increment_overlap, double outer_newton_tol,
                                           double inner_newton_tol, int outer_max_iter,
                                           int inner_max_iter, double* gamma_out,
                                           double* residual_out)
{
    if (!gamma_out || !(R > 0) || !(Rcell > 0) || !(E0 > 0) || !(init_gamma > 0) ||
        !(log_increment_gamma > 0) || !(increment_overlap > 0) || !(outer_newton_tol >= 0) ||
        !(inner_newton_tol >= 0) || outer_max_iter < 0 || inner_max_iter < 0)
        return BFM_ERR_ARG;
    try {
        // jkr.hpp:1386-1398: distance minus target as a function of log10(gamma),
        // each inner solve started at 0.75 * 2R and capped at 2 Rcell
        auto miss = [&](double x) {
            const double d_init = 0.75 * 2 * R;
            const double d_min = 2 * Rcell;
            const double gamma = std::pow(10.0, x);
            const double deq = equilibrium_distance(R, E0, gamma, d_init, d_min, increment_overlap,
                                                    inner_newton_tol, inner_max_iter, nullptr);
            return deq - deq_target;
        };
        double f = 0;
        const double x = newton_central(miss, std::log10(init_gamma), log_increment_gamma,
                                        outer_newton_tol, outer_max_iter, &f);
        const double gamma = std::pow(10.0, x);
        if (residual_out) *residual_out = f;
        if (!std::isfinite(gamma) || !std::isfinite(f)) return BFM_ERR_NAN;
        *gamma_out = gamma;
        return BFM_OK;
    } catch (const NoRoot&) {
        return BFM_ERR_JKR_ROOT;
    }
}

}  // extern "C"

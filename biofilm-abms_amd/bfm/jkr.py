"""Host-side JKR setup: contact-radius tables as runSimulation builds them.

The reference tabulates the JKR contact radius at setup
(calculateJKRContactRadii, ABM-3d/ABM-3d-cpp/include/simulation.hpp:86-143, used
at :1376-1405) on Eigen LinSpaced meshes: overlaps in [0, 2 (R - Rcell)] and,
for variable gamma, gammas in [0, max_gamma]. The radius is the larger real
root of a^4 - 2 Req d a^2 - (4 pi Req^2 g / E) a + Req^2 d^2 (jkr.hpp:67-144),
solved here in FP64 as the device does: u = sqrt(a) is the largest positive
root of u^4 - p u - q, found by Newton from an upper bound (monotone).
"""
import numpy as np


def linspaced(lo, hi, n):
    """Eigen's LinSpaced for floating point (both end points exact)."""
    if n == 1:
        return np.array([hi], dtype=np.float64)
    step = (hi - lo) / (n - 1)
    i = np.arange(n)
    flip = (n - 1) // 2
    return np.where(i < flip, lo + i * step, hi - (n - 1 - i) * step)


def contact_radius(delta, Req, E, gamma, max_iter=200):
    delta = np.asarray(delta, dtype=np.float64)
    gamma = np.broadcast_to(np.asarray(gamma, dtype=np.float64), delta.shape)
    p = Req * np.sqrt(4 * np.pi * gamma / E)
    q = Req * delta
    u = np.maximum(np.cbrt(2 * p), np.sqrt(np.sqrt(np.maximum(2 * q, 0.0))))
    active = u > 0
    for _ in range(max_iter):
        if not active.any():
            break
        u2 = u * u
        F = (u2 * u2 - p * u) - q
        dF = (4.0 * u2) * u - p
        with np.errstate(divide="ignore", invalid="ignore"):
            un = u - F / dF
        step = active & (dF > 0) & (un < u)
        u = np.where(step, un, u)
        active = step
    a = u * u
    bad = ((a * a - 2 * Req * delta * a) - p * p * a) + q * q > 1e-9 * np.maximum(1.0, np.abs(q))
    return np.where(bad & (delta < 0), np.nan, a)


def reference_table(R, Rcell, E0, gamma, n_overlap=100):
    """1-D table of IsotropicJKRDataFixedGamma (mechanics.hpp:113-153)."""
    ov = linspaced(0.0, 2 * (R - Rcell), n_overlap)
    return ov, contact_radius(ov, R / 2, E0, gamma)


def reference_table_2d(R, Rcell, E0, max_gamma, n_overlap=100, n_gamma=100):
    """2-D table of IsotropicJKRDataVariableGamma (mechanics.hpp:155-197);
    radii row-major [overlap][gamma]."""
    ov = linspaced(0.0, 2 * (R - Rcell), n_overlap)
    gm = linspaced(0.0, max_gamma, n_gamma)
    D, G = np.meshgrid(ov, gm, indexing="ij")
    return ov, gm, contact_radius(D, R / 2, E0, G)


def force_magnitude(radius, Req, E0, gamma):
    """term1 - term2 of forcesIsotropicJKRLagrange (adhesion.hpp:206-212)."""
    a3 = radius * radius * radius
    return 4 * E0 * a3 / (3 * Req) - 4 * np.sqrt(np.pi * a3 * gamma * E0)

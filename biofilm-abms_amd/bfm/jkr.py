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


def synthetic_aniso_tables(R=0.8, Rcell=0.5, E0=3900.0, gamma=100.0, n_theta=24, n_half_l=16,
                           n_coords=12, n_rx=20, n_phi=10, n_delta=40, gammas=None, seed=0):
    """Anisotropic JKR tables of the shapes AnisotropicJKRDataFixedGamma /
    VariableGamma hold (mechanics.hpp:199-327), filled from closed forms
    instead of the reference's ellipsoid pipeline (precomputeJKRCurvatureRadii
    and precomputeJKRForces, not built here): Rx at a contact point grows
    from R at the cap toward 3R along the body; the force and contact radius
    are the isotropic JKR ones for the effective radius sqrt(Rx1 Rx2) / 2,
    modulated by 1 + 0.05 cos(2 phi). As in the files, only Rx1 <= Rx2 rows
    are filled (the lookup orders the pair). gammas: a gamma grid (variable-
    gamma table) or None (fixed gamma). For tests and benchmarks of the
    lookup, not physics."""
    from .core import AnisoJKRTables
    theta = linspaced(0.0, np.pi / 2, n_theta)
    half_l = linspaced(0.2, 2.0, n_half_l)
    coords = linspaced(0.0, 1.0, n_coords)
    T, H, Cc = np.meshgrid(theta, half_l, coords, indexing="ij")
    curv = R * (1 + 2 * np.sin(T) ** 2 * (1 - Cc ** 4) * H / (H + R))
    rx = linspaced(R, 3 * R, n_rx)
    phi = linspaced(0.0, np.pi / 2, n_phi)
    delta = linspaced(0.0, 2 * (R - Rcell), n_delta)
    gm = None if gammas is None else np.asarray(gammas, dtype=np.float64)
    gv = np.array([gamma]) if gm is None else gm
    R1, R2, P, D, G = np.meshgrid(rx, rx, phi, delta, gv, indexing="ij")
    req = np.sqrt(R1 * R2) / 2
    a = contact_radius(D, req, E0, G)
    f = force_magnitude(a, req, E0, G) * (1 + 0.05 * np.cos(2 * P))
    fr = np.stack([f, a], axis=-1)
    fr[np.arange(n_rx)[:, None] > np.arange(n_rx)[None, :]] = 0.0
    if gm is None:
        fr = fr[..., 0, :]
    return AnisoJKRTables(theta, half_l, coords, curv, rx, phi, delta, fr, gamma=gm,
                          fixed_gamma=None if gm is not None else gamma)


def write_aniso_table_files(tables, curvature_path, forces_path):
    """The two text files runSimulation reads for adhesion_mode 2 (the format
    precomputeJKRCurvatureRadii / precomputeJKRForces write: tab-separated,
    10 significant digits, theta-major and Rx1-major rows, only Rx1 <= Rx2;
    the reduced force file starts with the fixed gamma). Ry is written as R
    of the first Rx grid point."""
    t = tables
    g10 = lambda x: "%.10g" % x   # noqa: E731
    with open(curvature_path, "w") as fh:
        for i, th in enumerate(t.theta):
            for j, hl in enumerate(t.half_l):
                for k, co in enumerate(t.coords):
                    fh.write("\t".join(g10(v) for v in (th, hl, co, t.curvature_rx[i, j, k],
                                                         t.rx[0])) + "\n")
    with open(forces_path, "w") as fh:
        if t.gamma is None:
            fh.write(g10(t.fixed_gamma) + "\n")
        n = len(t.rx)
        for i in range(n):
            for j in range(i, n):
                for p, ph in enumerate(t.phi):
                    for d, de in enumerate(t.delta):
                        if t.gamma is None:
                            f, a = t.force_radius[i, j, p, d]
                            fh.write("\t".join(g10(v) for v in (t.rx[i], t.rx[j], ph, de, f, a)) + "\n")
                        else:
                            for g, ga in enumerate(t.gamma):
                                f, a = t.force_radius[i, j, p, d, g]
                                fh.write("\t".join(g10(v) for v in (t.rx[i], t.rx[j], ph, de, ga, f,
                                                                     a)) + "\n")

"""Surface integrals, surface forces and the per-cell velocity solve of the oracle.

Integrals (integrals.hpp) against mpmath quadrature (the reference uses trapezoid
quadrature, test/src/testIntegrals.cpp); surface forces against finite differences
of the surface energies (testCellSurfaceRepulsionForces.cpp:41-85,
testCellSurfaceAdhesionForces.cpp:67-90), with the energies integrated by mpmath;
the velocity solve against the 7x7 constrained system built from
compositeViscosityForceMatrix (mechanics.hpp:947-1009) with quadrature area
integrals (the residual check of test/src/testVelocities.cpp:89-131).
"""
import math

import mpmath as mp
import numpy as np
import pytest

mp.mp.dps = 30
R = 0.8


def _delta(rz, nz, s):
    return max(R - rz - s * nz, 0)


def _quad(f, hl, ss):
    pts = [-hl, hl] if not (-hl < ss < hl) else [-hl, ss, hl]
    return mp.quad(f, pts)


CASES = [(rz, nz, hl) for nz in (0.1, 0.35, 0.7, 0.95) for hl in (0.5, 1.2)
         for rz in (0.3, 0.75, 1.1)]


@pytest.mark.parametrize("rz,nz,hl", CASES)
def test_integrals_vs_quadrature(orc, rz, nz, hl):
    rz, nz, hl = mp.mpf(rz), mp.mpf(nz), mp.mpf(hl)
    ss = (R - rz) / nz
    for g in (-0.5, 0.5, 1.0, 1.5, 2.0):
        d = lambda s: _delta(rz, nz, s)
        i1 = _quad(lambda s: d(s) ** g if d(s) > 0 else 0, hl, ss)
        i2 = _quad(lambda s: s * d(s) ** g if d(s) > 0 else 0, hl, ss)
        i3 = _quad(lambda s: s * s * d(s) ** g if d(s) > 0 else 0, hl, ss)
        args = (float(rz), float(nz), R, float(hl), g, float(ss))
        for k, ref in ((1, i1), (2, i2), (3, i3)):
            got = orc.integral(k, *args)
            assert abs(got - float(ref)) < 1e-9 * max(1.0, abs(float(ref))), (k, g, got, ref)
    th = lambda s: 1 if _delta(rz, nz, s) > 0 else 0
    for k, f in ((4, lambda s: th(s)), (5, lambda s: s * th(s)), (6, lambda s: s * s * th(s))):
        assert abs(orc.integral(k, float(nz), float(hl), float(ss)) - float(_quad(f, hl, ss))) < 1e-12


def _area(rz, nz, hl):
    ss = (R - rz) / nz
    a = lambda s: (mp.sqrt(R) * (1 - nz * nz) * mp.sqrt(_delta(rz, nz, s))
                   + mp.pi * R * nz * nz * (1 if _delta(rz, nz, s) > 0 else 0))
    return [float(_quad(lambda s: s ** k * a(s), hl, ss)) for k in range(3)]


@pytest.mark.parametrize("rz,nz,hl", CASES)
def test_area_integrals_vs_quadrature(orc, rz, nz, hl):
    got = orc.area_integrals(rz, nz, R, hl, (R - rz) / nz)
    ref = _area(mp.mpf(rz), mp.mpf(nz), mp.mpf(hl))
    assert np.allclose(got, ref, rtol=1e-9, atol=1e-12)


def _cell_row(rz, nz, hl, sigma0=100.0, eta0=0.7, eta1=720.0, nx=None, dr=(0, 0), dn=(0, 0)):
    nx = math.sqrt(1 - nz * nz) if nx is None else nx
    ny = math.sqrt(max(0.0, 1 - nz * nz - nx * nx))
    row = np.zeros(22)
    row[1:7] = [0.0, 0.0, rz, nx, ny, nz]
    row[7:9] = dr
    row[10:12] = dn
    row[13], row[14], row[16], row[17], row[18], row[19], row[20] = 2 * hl, hl, 1.5, eta0, eta1, sigma0, 1
    return row


def _energy_rep(rz, nz, hl, E0):
    ss = (R - rz) / nz
    i2 = _quad(lambda s: _delta(rz, nz, s) ** 2, hl, ss)
    i15 = _quad(lambda s: _delta(rz, nz, s) ** mp.mpf(1.5), hl, ss)
    return E0 * mp.sqrt(R) * (mp.power(R, -0.5) * (1 - nz * nz) * i2 + mp.mpf(4) / 3 * nz * nz * i15)


def _energy_adh(rz, nz, hl, sigma0):
    ss = (R - rz) / nz
    a = lambda s: (mp.sqrt(R) * (1 - nz * nz) * mp.sqrt(_delta(rz, nz, s))
                   + mp.pi * R * nz * nz * (1 if _delta(rz, nz, s) > 0 else 0))
    return -sigma0 * _quad(a, hl, ss)


@pytest.mark.parametrize("nz", [0.15, 0.4, 0.7, 0.95])
@pytest.mark.parametrize("overlap", [0.05, 0.2])
def test_surface_forces_vs_energy_derivatives(orc, nz, overlap):
    E0, sigma0, hl = 3900.0, 100.0, 0.5
    rz = R + hl * nz - overlap * R        # maximum overlap at the lower pole
    cells = _cell_row(rz, nz, hl)[None, :]
    a2d = np.zeros(1, np.int32)
    rep = orc.surface_repulsion_forces(cells, R, E0, a2d)
    adh = orc.surface_adhesion_forces(cells, R, a2d)
    h = mp.mpf("1e-12")
    mrz, mnz = mp.mpf(rz), mp.mpf(nz)
    drep_rz = mp.diff(lambda x: _energy_rep(x, mnz, hl, E0), mrz, h=h)
    drep_nz = mp.diff(lambda x: _energy_rep(mrz, x, hl, E0), mnz, h=h)
    dadh_rz = mp.diff(lambda x: _energy_adh(x, mnz, hl, sigma0), mrz, h=h)
    dadh_nz = mp.diff(lambda x: _energy_adh(mrz, x, hl, sigma0), mnz, h=h)
    for got, ref in ((rep[0, 2], drep_rz), (rep[0, 5], drep_nz), (adh[0, 2], dadh_rz),
                     (adh[0, 5], dadh_nz)):
        assert abs(got - float(ref)) < 1e-8 * max(1.0, abs(float(ref))), (got, ref)


def _params(**kw):
    import bfm
    return bfm.core.make_params(**kw)


def _solve7(M, n, F):
    A = np.zeros((7, 7))
    A[:6, :6] = M
    A[3:6, 6] = -2 * n
    A[6, 3:6] = n
    return np.linalg.solve(A, np.concatenate([F, [0.0]]))[:6]


@pytest.mark.parametrize("nz", [0.2, 0.5, 0.8])
def test_tilted_velocity_solve_residual(orc, nz):
    """Tilted contacting cell, Coulomb off: M v = F + 2 lambda n, n . dn = 0."""
    hl, eta0, eta1 = 0.9, 0.7, 720.0
    rz = R + hl * nz - 0.1
    row = _cell_row(rz, nz, hl, eta0=eta0, eta1=eta1, nx=0.6 * math.sqrt(1 - nz * nz))
    cells = row[None, :]
    p = _params(cell_surface_coulomb_coeff=0.0)
    a2d = np.zeros(1, np.int32)
    noise = np.array([[3.0, -2.0, 1.5, 0.7, -0.4, 0.9]])
    nb = np.zeros((0, 7), order="F")
    F = orc.conservative_forces(cells, nb, a2d, p)[0] + noise[0]
    v = orc.get_velocities(cells, nb, a2d, noise, p)[0]
    A1, A2, A3 = _area(mp.mpf(rz), mp.mpf(nz), mp.mpf(hl))
    l = 2 * hl
    M = np.zeros((6, 6))
    M[0, 0] = M[1, 1] = eta0 * l + eta1 / R * A1
    M[0, 3] = M[1, 4] = M[3, 0] = M[4, 1] = eta1 / R * A2
    M[2, 2] = eta0 * l
    M[3, 3] = M[4, 4] = eta0 * l ** 3 / 12 + eta1 / R * A3
    M[5, 5] = eta0 * l ** 3 / 12
    ref = _solve7(M, row[4:7], F)
    assert np.allclose(v, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def _integral1(rz, nz, hl, g):
    ss = (R - rz) / nz
    return float(_quad(lambda s: _delta(rz, nz, s) ** g if _delta(rz, nz, s) > 0 else 0, hl, ss))


@pytest.mark.parametrize("nz", [0.25, 0.6, 0.9])
@pytest.mark.parametrize("coeff", [0.5, 1.0])
def test_tilted_velocity_solve_coulomb_cap(orc, nz, coeff):
    """Tilted contacting cell with large stored velocities: the surface friction
    coefficient is truncated so that (eta1_eff / R)(A0 |dr_xy| + A1 |dn_xy|)
    equals 2 E0 coeff [(1 - nz^2) I1(1) + sqrt(R) nz^2 I1(1/2)]
    (mechanics.hpp:2246-2262), and the velocities solve the 7x7 constrained
    system built with eta1_eff. Integrals by quadrature."""
    hl, eta0, eta1, E0 = 0.9, 0.7, 720.0, 3900.0
    rz = R + hl * nz - 0.15
    dr, dn = (40.0, -25.0), (12.0, 30.0)
    row = _cell_row(rz, nz, hl, eta0=eta0, eta1=eta1, nx=0.6 * math.sqrt(1 - nz * nz), dr=dr, dn=dn)
    cells = row[None, :]
    p = _params(cell_surface_coulomb_coeff=coeff)
    a2d = np.zeros(1, np.int32)
    noise = np.array([[3.0, -2.0, 1.5, 0.7, -0.4, 0.9]])
    nb = np.zeros((0, 7), order="F")
    F = orc.conservative_forces(cells, nb, a2d, p)[0] + noise[0]
    v = orc.get_velocities(cells, nb, a2d, noise, p)[0]
    mrz, mnz, mhl = mp.mpf(rz), mp.mpf(nz), mp.mpf(hl)
    A1, A2, A3 = _area(mrz, mnz, mhl)
    drxy, dnxy = math.hypot(*dr), math.hypot(*dn)
    max_norm = 2 * E0 * coeff * ((1 - nz * nz) * _integral1(mrz, mnz, mhl, 1)
                                 + math.sqrt(R) * nz * nz * _integral1(mrz, mnz, mhl, mp.mpf(0.5)))
    friction = (eta1 / R) * (A1 * drxy + A2 * dnxy)
    assert friction > 2 * max_norm          # the cap is active in these cases
    eta1_eff = max_norm * R / (A1 * drxy + A2 * dnxy)
    assert abs((eta1_eff / R) * (A1 * drxy + A2 * dnxy) - max_norm) < 1e-12 * max_norm
    l = 2 * hl

    def drag(e1):
        M = np.zeros((6, 6))
        M[0, 0] = M[1, 1] = eta0 * l + e1 / R * A1
        M[0, 3] = M[1, 4] = M[3, 0] = M[4, 1] = e1 / R * A2
        M[2, 2] = eta0 * l
        M[3, 3] = M[4, 4] = eta0 * l ** 3 / 12 + e1 / R * A3
        M[5, 5] = eta0 * l ** 3 / 12
        return M

    ref = _solve7(drag(eta1_eff), row[4:7], F)
    assert np.allclose(v, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    untruncated = _solve7(drag(eta1), row[4:7], F)
    assert not np.allclose(v, untruncated, rtol=1e-3)
    assert abs(np.dot(v[3:], row[4:7])) < 1e-12 * np.abs(v[3:]).max()


def test_free_cell_velocity_solve(orc):
    row = _cell_row(5.0, 0.6, 0.8)
    cells = row[None, :]
    p = _params()
    a2d = np.zeros(1, np.int32)
    noise = np.array([[1.0, 2.0, -3.0, 0.5, -0.25, 0.125]])
    v = orc.get_velocities(cells, np.zeros((0, 7), order="F"), a2d, noise, p)[0]
    K, l = 0.7 * 1.6, 1.6
    L = K * l * l / 12
    ref = _solve7(np.diag([K, K, K, L, L, L]), row[4:7], noise[0])
    assert np.allclose(v, ref, rtol=1e-12)


def test_horizontal_cell_constraint_and_coulomb(orc):
    """assume_2d branch: dn_xy is orthogonal to n_xy, Coulomb truncation caps the
    friction force at coeff * 2 E0 (R - rz) l (mechanics.hpp:2144-2224)."""
    row = _cell_row(0.7, 0.0, 1.0, nx=0.8, dr=(50.0, -20.0))
    cells = row[None, :]
    a2d = np.ones(1, np.int32)
    noise = np.array([[4.0, -1.0, 0.3, 0.6, 0.2, -0.1]])
    nb = np.zeros((0, 7), order="F")
    for coeff in (0.0, 1.0):
        p = _params(cell_surface_coulomb_coeff=coeff)
        v = orc.get_velocities(cells, nb, a2d, noise, p)[0]
        assert abs(v[3] * row[4] + v[4] * row[5]) < 1e-12
    # with a large stored velocity the friction is truncated: K_fric * |dr_xy| = max_norm
    p = _params(cell_surface_coulomb_coeff=1.0)
    F = orc.conservative_forces(cells, nb, a2d, p)[0] + noise[0]
    v = orc.get_velocities(cells, nb, a2d, noise, p)[0]
    l, rz, E0 = 2.0, 0.7, 3900.0
    K = F[0] / v[0]
    K_fric = K - 0.7 * l
    assert abs(K_fric * math.hypot(50.0, -20.0) - 2 * E0 * (R - rz) * l) < 1e-9 * K_fric * 53.8

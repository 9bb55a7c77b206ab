"""Velocity Verlet (stepVelocityVerlet, ABM-3d/ABM-3d-cpp/include/mechanics.hpp:
2668-2996) and the include_constraint = true forces it uses, in the oracle,
pinned by what the mathematics fixes:

* a free cell (no surface, no drag, no neighbors) moves its end points with
  v +- w x (p - q) / 2: r advances by dt v, n by dt w x n before
  renormalisation, and the stored velocities stay v and w x n;
* internal forces (cell-cell repulsion, adhesion, friction) come in equal and
  opposite pairs, so without surface and ambient drag the total momentum
  sum m_i v_i is conserved by both half steps;
* with include_constraint = true every torque is perpendicular to its cell's
  orientation: the (s n) x F form, the JKR form s (-(n . v) n - v)
  (adhesion.hpp:225-231) and the surface forms (n_y, -n_x, 0) (...);
* the scheme is second order: halving dt divides the error after a fixed
  time by about four.
"""
import numpy as np
import pytest

R = 0.8


def _params(**kw):
    from bfm.core import make_params
    return make_params(**kw)


def _cells(n, seed, side=5.0, z0=3.0, vel=1.0, ncols=22):
    rng = np.random.default_rng(seed)
    c = np.zeros((n, ncols), order="F")
    c[:, 0] = np.arange(n)
    c[:, 1:3] = rng.uniform(0, side, (n, 2))
    c[:, 3] = z0 + rng.uniform(0, side / 2, n)
    v = rng.normal(size=(n, 3))
    c[:, 4:7] = v / np.linalg.norm(v, axis=1)[:, None]
    c[:, 6] = np.abs(c[:, 6])
    hl = rng.uniform(0.5, 1.2, n)
    c[:, 13], c[:, 14] = 2 * hl, hl
    c[:, 16], c[:, 17], c[:, 18], c[:, 19], c[:, 20] = 1.5, 0.7, 720, 100, 1
    c[:, 21] = 100.0
    c[:, 7:10] = rng.normal(scale=vel, size=(n, 3))
    w = rng.normal(scale=vel, size=(n, 3))
    c[:, 10:13] = np.cross(w, c[:, 4:7])           # dn = w x n (perpendicular to n)
    if ncols > 22:
        c[:, 22] = rng.uniform(0.2, 2.0, n)
    return c


def _masses(cells, rho):
    l = cells[:, 13]
    return rho * np.pi * ((4 / 3) * R ** 3 + R * R * l)


def test_free_cell_closed_form(orc):
    c = _cells(1, 3)
    c[:, 17] = 0.0                                   # no ambient drag
    nb = np.zeros((0, 7), order="F")
    dt = 1e-3
    new = orc.step_verlet(dt, 2.0, c, nb, _params(no_surface=True))
    n0, v = c[0, 4:7], c[0, 7:10]
    w = np.cross(n0, c[0, 10:13])                    # angular velocity n x dn
    np.testing.assert_allclose(new[0, 1:4], c[0, 1:4] + dt * v, rtol=0, atol=1e-14)
    n1 = n0 + dt * np.cross(w, n0)
    n1 /= np.linalg.norm(n1)
    np.testing.assert_allclose(new[0, 4:7], n1, rtol=0, atol=1e-14)
    np.testing.assert_allclose(new[0, 7:10], v, rtol=0, atol=1e-14)
    np.testing.assert_allclose(new[0, 10:13], np.cross(w, new[0, 4:7]), rtol=0, atol=1e-13)


@pytest.mark.parametrize("adhesion,friction", [(False, False), (True, False), (True, True)])
def test_momentum_conserved_without_external_forces(orc, adhesion, friction):
    ncols = 23 if friction else 22
    c = _cells(60, 5, side=6.0, ncols=ncols)
    c[:, 17] = 0.0
    p = _params(no_surface=True, adhesion=adhesion, friction=friction)
    nb = orc.get_cell_neighbors(c, 2.0, R, 2.4)
    assert nb.shape[0] > 30
    rho = 1e-3
    m = _masses(c, rho)
    new = orc.step_verlet(1e-6, rho, c, nb, p)
    p0 = (m[:, None] * c[:, 7:10]).sum(axis=0)
    p1 = (m[:, None] * new[:, 7:10]).sum(axis=0)
    scale = np.abs(m[:, None] * new[:, 7:10]).sum()
    assert scale > 10 * np.abs(p0).max()            # the forces did change the velocities
    assert np.max(np.abs(p1 - p0)) < 1e-12 * scale


@pytest.mark.parametrize("adhesion,friction", [(False, False), (True, True)])
def test_constraint_torques_perpendicular_to_orientation(orc, adhesion, friction):
    ncols = 23 if friction else 22
    c = _cells(80, 7, side=6.0, z0=0.3, ncols=ncols)
    p = _params(adhesion=adhesion, friction=friction)
    nb = orc.get_cell_neighbors(c, 2.0, R, 2.4)
    f = orc.cell_cell_forces_ic(c, nb, p)
    tq = f[:, 3:6]
    dots = np.abs(np.sum(tq * c[:, 4:7], axis=1))
    assert np.max(np.abs(tq)) > 1.0
    assert np.max(dots) < 1e-12 * np.max(np.abs(tq))
    # the free (include_constraint = false) form is not perpendicular
    f0 = orc.cell_cell_forces(c, nb, p)
    assert np.max(np.abs(np.sum(f0[:, 3:6] * c[:, 4:7], axis=1))) > 1e-3 * np.max(np.abs(f0[:, 3:6]))
    # surface terms: conservative forces with the constraint, cells near the surface
    a2d = (c[:, 6] < p.nz_threshold).astype(np.int32)
    cons = orc.conservative_forces_ic(c, nb, a2d, p)
    tq = cons[:, 3:6]
    assert np.max(np.abs(np.sum(tq * c[:, 4:7], axis=1))) < 1e-12 * np.max(np.abs(tq))


@pytest.mark.parametrize("eta0,lo,hi", [(0.0, 3.8, 4.2), (0.7, 1.8, 2.4)])
def test_convergence_order(orc, eta0, lo, hi):
    """Two overlapping cells (Hertz + JKR, no surface) advanced to t = 0.02 h:
    with position-dependent forces only the error against a fine run shrinks
    4x per halving of dt (second order); with the ambient drag, which the
    second force evaluation takes at the half-step velocities, 2x (first
    order, as the reference's scheme is written)."""
    c = np.zeros((2, 22), order="F")
    c[:, 0] = [0, 1]
    c[0, 1:4] = [0.0, 0.0, 3.0]
    c[1, 1:4] = [0.3, 1.45, 3.1]
    c[0, 4:7] = np.array([1, 0, 0.3]) / np.linalg.norm([1, 0, 0.3])   # n_z well above 0: no flip
    c[1, 4:7] = np.array([1, 0.2, 0.4]) / np.linalg.norm([1, 0.2, 0.4])
    c[:, 13], c[:, 14] = 2.0, 1.0
    c[:, 16], c[:, 17], c[:, 18], c[:, 19], c[:, 20] = 1.5, eta0, 720, 100, 1
    c[:, 21] = 100.0
    p = _params(no_surface=True)

    def run(dt, steps):
        x = c.copy(order="F")
        nb = orc.get_cell_neighbors(x, 2.0, R, 2.4)
        for _ in range(steps):
            x = orc.step_verlet(dt, 1.0, x, nb, p)
            orc.update_neighbor_distances(x, nb)
        return x

    T = 2e-2
    ref = run(T / 1024, 1024)
    errs = [np.max(np.abs(run(T / k, k)[:, 1:7] - ref[:, 1:7])) for k in (8, 16, 32)]
    for a, b in zip(errs, errs[1:]):
        assert lo < a / b < hi

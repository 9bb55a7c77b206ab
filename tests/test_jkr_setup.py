"""JKR setup of runSimulation (jkr.hpp:1310-1408, simulation.hpp:1352-1374):
the equilibrium distance for a surface energy density and the density for a
target distance, in the oracle and in the product's host code, against the
closed forms the JKR model fixes.

With equivalent radius Req = R/2, the JKR force 4/3 E0/Req a^3 - 4 sqrt(pi
gamma E0) a^1.5 vanishes at a^3 = 9 pi gamma Req^2 / E0, and the JKR relation
delta = a^2/Req - sqrt(4 pi gamma a / E0) then gives delta = a^2 / (3 Req); so
deq = 2R - a^2/(3 Req) and, inverted, gamma = (3 Req delta)^1.5 E0 / (9 pi Req^2).
The reference's own test (testJKR.cpp:85-116) checks the first identity with
the cell radius R where jkrEquilibriumDistance uses Req = R/2; the identity
holds with Req."""
import math

import numpy as np
import pytest

R, RCELL, E0 = 0.8, 0.5, 3900.0
REQ = R / 2


def deq_closed(gamma):
    a = (9 * math.pi * gamma * REQ * REQ / E0) ** (1.0 / 3.0)
    return 2 * R - a * a / (3 * REQ)


def gamma_closed(deq):
    delta = 2 * R - deq
    return (3 * REQ * delta) ** 1.5 * E0 / (9 * math.pi * REQ * REQ)


GAMMAS = np.linspace(10.0, 100.0, 41)   # the reference test's grid


@pytest.mark.parametrize("gamma", GAMMAS[::4].tolist() + [300.0, 500.0])
def test_equilibrium_distance_closed_form(orc, gamma):
    from bfm import sim
    d_o = orc.jkr_equilibrium_distance(R, E0, gamma, 0.98 * 2 * R, 2 * RCELL)
    d_p = sim.jkr_equilibrium_distance(R, E0, gamma, 0.98 * 2 * R, 2 * RCELL)
    assert d_p == d_o
    assert abs(d_o - deq_closed(gamma)) < 1e-9
    # the contact radius at the equilibrium overlap satisfies a^3 = 9 pi gamma Req^2 / E0
    a, _ = orc.jkr_contact_radius(2 * R - d_o, REQ, E0, gamma)
    assert abs(a - (9 * math.pi * gamma * REQ * REQ / E0) ** (1 / 3)) < 1e-8
    # and the JKR force there is zero to the Newton tolerance
    assert abs(orc.jkr_force_magnitude(a, REQ, E0, gamma)) < 1e-8


def test_equilibrium_distance_cap(orc):
    """An equilibrium overlap beyond 2R - dmin is capped: deq = dmin."""
    from bfm import sim
    assert deq_closed(1000.0) < 2 * RCELL
    assert orc.jkr_equilibrium_distance(R, E0, 1000.0, 0.98 * 2 * R, 2 * RCELL) == 2 * RCELL
    assert sim.jkr_equilibrium_distance(R, E0, 1000.0, 0.98 * 2 * R, 2 * RCELL) == 2 * RCELL


@pytest.mark.parametrize("target,init", [(1.2, 100.0), (1.4, 100.0), (1.5, 10.0), (1.5, 100.0),
                                         (1.55, 10.0), (1.59, 100.0)])
def test_optimal_surface_energy_density(orc, target, init):
    from bfm import sim
    g_o = orc.jkr_optimal_surface_energy_density(R, RCELL, E0, target, init)
    g_p = sim.jkr_optimal_surface_energy_density(R, RCELL, E0, target, init)
    assert g_p == g_o
    assert abs(g_o / gamma_closed(target) - 1) < 1e-7
    assert abs(orc.jkr_equilibrium_distance(R, E0, g_o, 1.5 * R, 2 * RCELL) - target) < 1e-8


def test_divergent_newton_is_reported(orc):
    """From init_gamma 10 the reference's Newton on log10(gamma) for a 1.2 um
    target steps into the capped region (zero derivative) and diverges; the
    oracle returns a non-finite value and the product an error status."""
    from bfm import sim
    from bfm.core import BFMError
    assert not math.isfinite(orc.jkr_optimal_surface_energy_density(R, RCELL, E0, 1.2, 10.0))
    with pytest.raises(BFMError) as e:
        sim.jkr_optimal_surface_energy_density(R, RCELL, E0, 1.2, 10.0)
    assert e.value.status in (3, 4)


def test_below_pull_off_is_reported():
    """An overlap below the JKR pull-off point has no contact radius: the
    reference throws (jkr.hpp:99-105), the product returns BFM_ERR_JKR_ROOT."""
    from bfm import sim
    from bfm.core import BFMError
    with pytest.raises(BFMError) as e:
        sim.jkr_equilibrium_distance(R, E0, 100.0, 2 * R + 0.5, 2 * RCELL, max_iter=0)
    assert e.value.status == 4


def test_init_surface_energy_by_group():
    """simulation.hpp:1368-1374: group 1 gets the solved density, others 0."""
    from bfm import sim
    from bfm.core import COL
    cells = np.zeros((4, 22), order="F")
    cells[:, COL["group"]] = [1, 2, 1, 2]
    g = sim.init_surface_energy(cells, 1.5)
    assert abs(g / gamma_closed(1.5) - 1) < 1e-7
    assert np.array_equal(cells[:, COL["gamma"]], [g, 0.0, g, 0.0])

"""RK step and step-size controller of the oracle.

stepRungeKuttaAdaptive (mechanics.hpp:2406-2612): tableau identities of the four
Butcher tableaux (simulation.hpp:1207-1268), the embedded error's order (a wrong
stage index or weight breaks the dt scaling), frozen neighbor geometry, the
normalizeOrientations flip. Controller (simulation.hpp:1688-1769): retry and
redo semantics on scripted error sequences.
"""
import math

import numpy as np
import pytest

import bfm


def _two_cells(no_surface=True):
    cells = np.zeros((2, 22))
    cells[0, 1:7] = [0.0, 0.0, 3.0, 1.0, 0.0, 0.0]
    cells[1, 1:7] = [0.3, 1.45, 3.1, 0.8, 0.0, 0.6]     # skew pair, overlap 0.15
    for c in cells:
        c[13], c[14], c[16], c[17], c[18], c[19], c[20], c[21] = 2.0, 1.0, 1.5, 0.7, 720, 100, 1, 100.0
    cells[:, 0] = [0, 1]
    return np.asfortranarray(cells)


@pytest.mark.parametrize("name", list(bfm.TABLEAUX))
def test_tableau_identities(name):
    A, b, bs, order = bfm.tableau(name)
    s = len(b)
    assert A.shape == (s, s) and np.allclose(np.triu(A), 0)
    assert abs(b.sum() - 1) < 1e-14 and abs(bs.sum() - 1) < 1e-14
    c = A.sum(1)
    if s > 2:   # second-order conditions of both weight vectors
        assert abs(b @ c - 0.5) < 1e-14 and abs(bs @ c - 0.5) < 1e-14


@pytest.mark.parametrize("name", list(bfm.TABLEAUX))
def test_embedded_error_order(orc, name):
    A, b, bs, order = bfm.tableau(name)
    p = bfm.core.make_params(no_surface=True)
    cells = _two_cells()
    nb = orc.get_cell_neighbors(cells, 2.0, 0.8, 2.4)
    assert nb.shape[0] == 1
    errs = []
    for dt in (4e-4, 2e-4, 1e-4):
        _, err = orc.step_rk(A, b, bs, dt, cells, nb, p)
        errs.append(np.abs(err).max())
    for e1, e2 in zip(errs, errs[1:]):
        ratio = e1 / e2
        assert abs(math.log2(ratio) - (order + 1)) < 0.05, (name, errs)


def test_step_keeps_unit_orientation_and_upper_hemisphere(orc):
    p = bfm.core.make_params(no_surface=True)
    A, b, bs, _ = bfm.tableau("dormand_prince")
    cells = _two_cells()
    cells[1, 4:7] = [0.6, 0.0, -0.8]     # nz < 0 is flipped at the end of the step
    nb = orc.get_cell_neighbors(cells, 2.0, 0.8, 2.4)
    new, err = orc.step_rk(A, b, bs, 1e-6, cells, nb, p)
    assert np.allclose(np.linalg.norm(new[:, 4:7], axis=1), 1.0, atol=1e-15)
    assert np.all(new[:, 6] >= 0)
    # stored velocities are the b-weighted stage velocities: new - old = dt * v (pre-normalisation)
    assert np.allclose(new[:, 1:4] - cells[:, 1:4], 1e-6 * new[:, 7:10], rtol=1e-9, atol=1e-18)


def test_noise_enters_every_stage_identically(orc):
    p = bfm.core.make_params(no_surface=True)
    A, b, bs, _ = bfm.tableau("heun_euler")
    cells = _two_cells()
    cells[1, 1] = 9.0         # no contact: velocities come from the noise alone
    nb = np.zeros((0, 7), order="F")
    noise = np.asfortranarray(np.tile([1.0, -2.0, 0.5, 0.0, 0.0, 0.0], (2, 1)))
    new, err = orc.step_rk(A, b, bs, 1e-3, cells, nb, p, noise)
    K = 0.7 * 2.0
    assert np.allclose(new[:, 7:10], noise[:, :3] / K, rtol=1e-14)
    assert np.allclose(err, 0, atol=1e-18)


def test_controller_accepts_small_error_with_one_redo(orc):
    dts = orc.controller_trace(1e-3, 1, 3, 1e-9, 1e-2, [0.25])
    assert len(dts) == 2 and dts[0] == 1e-3
    assert dts[1] > 1e-3                 # error below 1 grows dt
    dts = orc.controller_trace(1e-3, 1, 3, 1e-9, 1e-2, [1e-12])
    assert dts[1] == 1e-2                # growth capped at 10x and at max_stepsize


def test_controller_retries_then_redoes(orc):
    # three retries all failing: 1 + 3 + redo attempts, dt shrinking each time
    f = 0.9 * (1 / 1.5) ** 0.5
    dts = orc.controller_trace(1e-3, 1, 3, 1e-12, 1e-2, [1.5, 1.5, 1.5, 1.5])
    assert len(dts) == 5 and np.all(np.diff(dts[:4]) < 0)
    assert np.allclose(dts, 1e-3 * f ** np.array([0, 1, 2, 3, 4]), rtol=1e-14)  # factors accumulate
    # factor falling below 0.2 during the retries clips to 0.2 and stops retrying
    dts = orc.controller_trace(1e-3, 1, 3, 1e-12, 1e-2, [4.0, 4.0, 4.0, 4.0])
    assert len(dts) == 4 and abs(dts[-1] - 0.2e-3) < 1e-18
    # a huge error clips the factor at 0.2 and breaks out of the retry loop
    dts = orc.controller_trace(1e-3, 1, 3, 1e-12, 1e-2, [1e12, 1e12])
    assert len(dts) == 3 and abs(dts[-1] - 0.2 * 1e-3) < 1e-18   # reset to 0.2, not multiplied
    # min_stepsize clamp
    dts = orc.controller_trace(1e-3, 1, 3, 5e-4, 1e-2, [1e12, 1e12])
    assert dts[-1] == 5e-4


def test_max_scaled_error_is_signed_maximum(orc):
    cells = _two_cells()
    err = np.zeros((2, 6), order="F")
    err[0, 0] = -5e-8
    err[1, 2] = 2e-9
    got = orc.max_scaled_error(cells, err, 1e-8)
    assert abs(got - 2e-9 / (1e-8 * (1 + 3.1))) < 1e-15

"""The 2-D tree's growth loop on the GPU (bfm2d_simulate) against the oracle's
restatement of runSimulationAdaptiveLagrangian (ABM-2d simulation.hpp:549-765)
from the 2-D driver's founder cell, and the isotropic-JKR adhesion of the 2-D
composition against the oracle's 2-D velocities.

Tolerances as for the 3-D loop: device and oracle velocities agree to ~1e-13
relative (different summation order), so trajectories keep the same division
events, lineage and iteration counts unless a threshold test lands within
rounding, the state agrees to 4e-12 um per iteration (at least 1e-9), and the
accepted dt to 1e-6 relative (dt follows err^(-1/(p+1)), err being the small
difference of two weighted sums)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R, RCELL, L0 = 0.8, 0.5, 0.4
LDIV = 2 * L0 + 2 * R
AMAP = np.array([[0, 0], [0, 1]], dtype=np.int32)       # group-1 cells adhere (the 2-D drivers)
TOL_STATE_PER_ITER = 4e-12
TOL_DT = 1e-6


def _model(cap, mode=4, gamma=100.0):
    from bfm.model2d import Model2D, make_params
    return Model2D(cap, make_params(adhesion_mode=mode, gamma=gamma), max_pairs=32 * cap + 1024)


def test_jkr_isotropic_velocities_match_oracle(orc):
    """getVelocities of the 2-D step with mode-4 adhesion on a 3,000-cell
    layer (rows with the driver's to_adhere), and one Bogacki-Shampine step."""
    from bfm import tableau
    from bfm.model2d import random_layer, to_adhere_rule, make_params
    n = 3000
    c = random_layer(n, 11)
    print(f"2-D isotropic JKR velocities n={n}", flush=True)
    m2 = _model(n)
    nb = m2.getCellNeighbors(c, 2 * R + LDIV, R, LDIV)
    nb_o = orc.get_cell_neighbors_2d(c, 2 * R + LDIV, R, LDIV)
    assert np.array_equal(nb, nb_o)
    ta = to_adhere_rule(c, nb, R, RCELL, AMAP)
    assert ta.sum() > 100
    p = make_params(adhesion_mode=4, gamma=100.0)
    v = m2.getVelocities(c, nb, ta)
    v_o = orc.get_velocities_2d(c, nb_o, p, ta)
    scale = np.max(np.abs(v_o), axis=0)
    assert np.max(np.abs(v - v_o) / scale) < 1e-10
    A, b, bs, _ = tableau("bogacki_shampine")
    new, err = m2.stepRungeKuttaAdaptive(A, b, bs, c, nb, 1e-4, ta)
    new_o, err_o = orc.step_rk_2d(A, b, bs, 1e-4, c, nb_o, p, ta)
    assert np.max(np.abs(new[:, 1:9] - new_o[:, 1:9]) / np.max(np.abs(new_o[:, 1:9]), axis=0)) < 1e-10


@pytest.mark.parametrize("tableau,n_cells,mode", [("bogacki_shampine", 24, 4),
                                                  ("bogacki_shampine", 16, 0),
                                                  ("heun_euler", 12, 4),
                                                  ("dormand_prince", 12, 4)])
def test_founder_growth_matches_oracle(orc, tableau, n_cells, mode):
    """Founder of the 2-D driver to n_cells: division events and lineage,
    iteration count, accepted dt sequence and final state."""
    from bfm.model2d import founder_cell_2d, make_params, sim_config_2d, simulate_2d
    cfg = sim_config_2d(n_cells=n_cells, tableau=tableau, seed=42)
    cap = 64
    print(f"2-D founder growth {tableau} n_cells={n_cells} mode={mode}", flush=True)
    m2 = _model(cap, mode=mode)
    got, parents, trace, st = simulate_2d(m2, cfg, founder_cell_2d(), AMAP, cap=cap)
    want, par_o, trace_o, iters_o, steps_o = orc.simulate_2d(cfg, make_params(adhesion_mode=mode), AMAP,
                                                             founder_cell_2d(), cap=cap)
    assert st["divisions"] == got.shape[0] - 1 and got.shape[0] >= n_cells
    assert got.shape == want.shape
    assert st["iterations"] == iters_o
    assert abs(st["rk_steps"] - steps_o) <= max(2, steps_o // 1000)
    assert np.array_equal(parents, par_o)
    assert np.array_equal(got[:, 0], want[:, 0]) and np.array_equal(got[:, 16], want[:, 16])
    tol = max(1e-9, TOL_STATE_PER_ITER * iters_o)
    for cols in (slice(1, 5), slice(9, 11), slice(11, 13)):
        assert np.max(np.abs(got[:, cols] - want[:, cols])) < tol
    assert np.max(np.abs(trace - trace_o) / trace_o) < TOL_DT
    assert abs(st["t_end"] - trace.sum()) < 1e-9


def test_stop_rules_and_errors(orc):
    """max_iter and max_time stop the loop at the reference's terminate rule;
    a division past the capacity is BFM_ERR_CAPACITY; the loop needs the 2-D
    parameters."""
    from bfm.core import BFMError, Context
    from bfm.model2d import founder_cell_2d, sim_config_2d, simulate_2d
    print("2-D growth stop rules", flush=True)
    m2 = _model(16)
    _, _, tr, st = simulate_2d(m2, sim_config_2d(max_iter=500, seed=1), founder_cell_2d(), AMAP, cap=16)
    assert st["iterations"] == 500 and len(tr) == 500
    _, _, tr, st = simulate_2d(m2, sim_config_2d(max_time=0.05, seed=1), founder_cell_2d(), AMAP, cap=16)
    assert st["t_end"] >= 0.05 and st["t_end"] - tr[-1] < 0.05
    with pytest.raises(BFMError) as e:
        simulate_2d(m2, sim_config_2d(n_cells=40, seed=1), founder_cell_2d(), AMAP, cap=8)
    assert e.value.status == 2

    class Bare:
        pass
    bare = Bare()
    bare.ctx = Context(16, max_pairs=1024)
    with pytest.raises(BFMError) as e:
        simulate_2d(bare, sim_config_2d(max_iter=5), founder_cell_2d(), AMAP, cap=16)
    assert e.value.status == 1


def test_zero_length_daughter_stops_with_nan():
    """divideCells clips the length ratio to [0, 1] (simulation.hpp:302-315 of
    ABM-2d), so a daughter can have length 0, where the composite drag l eta
    vanishes and the reference's velocities of that cell are not finite. A
    wide length-ratio distribution makes that happen at the first division:
    the loop stops with BFM_ERR_NAN and hands back the state the failing
    attempt started from."""
    from bfm.core import BFMError
    from bfm.model2d import founder_cell_2d, sim_config_2d, simulate_2d
    print("2-D growth zero-length daughter", flush=True)
    cfg = sim_config_2d(n_cells=8, daughter_length_std=50.0, seed=2)
    m2 = _model(16)
    with pytest.raises(BFMError) as e:
        simulate_2d(m2, cfg, founder_cell_2d(), AMAP, cap=16)
    assert e.value.status == 3
    cells, _, _, st = simulate_2d(m2, cfg, founder_cell_2d(), AMAP, cap=16, allow_nan=True)
    assert st["nan"] and cells.shape[0] == 2 and np.all(np.isfinite(cells))
    assert np.min(cells[:, 9]) == 0.0

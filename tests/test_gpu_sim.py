"""GPU growth / division loop (bfm_simulate) against the oracle's restatement
of runSimulation on the same seeded inputs and division random stream.

Tolerance: the device velocities agree with the oracle's to ~1e-13 relative
(different summation order), the step-size controller is continuous in the
error, and division events are threshold tests on l; so trajectories agree
to 1e-9 (absolute, um and unit vectors) with identical division events,
lineage and step counts as long as no quantity lands within rounding of a
threshold. The accepted dt sequences agree to 1e-9 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_STATE = 1e-9
TOL_DT = 1e-9


def _ctx(cap):
    from bfm import Context
    from bfm.core import make_params
    return Context(cap, max_pairs=32 * cap + 1024, params=make_params())


def _compare(got, want, trace, trace_o, stats, iters_o, steps_o):
    assert got.shape == want.shape
    assert stats["iterations"] == iters_o
    assert stats["rk_steps"] == steps_o
    assert np.array_equal(got[:, 0], want[:, 0])            # ids
    assert np.array_equal(got[:, 20], want[:, 20])          # groups
    for cols in (slice(1, 7), slice(13, 15), slice(15, 17)):
        assert np.max(np.abs(got[:, cols] - want[:, cols])) < TOL_STATE
    assert np.max(np.abs(trace - trace_o) / trace_o) < TOL_DT


@pytest.mark.parametrize("tableau,n_cells,max_iter,zb", [("bogacki_shampine", 16, 0, 0.0),
                                                         ("dormand_prince", 16, 0, 0.0),
                                                         ("heun_euler", 0, 12000, 0.05)])
def test_founder_growth_matches_oracle(orc, tableau, n_cells, max_iter, zb):
    """From the single founder cell of the reference driver through the first
    divisions (the stiff post-division transient included)."""
    from bfm import sim
    from bfm.core import make_params
    cfg = sim.sim_config(n_cells=n_cells, max_iter=max_iter, tableau=tableau,
                         daughter_angle_xy_bound=0.05, daughter_angle_z_bound=zb, seed=42)
    c0 = sim.founder_cell()
    cap = 64
    print(f"founder growth {tableau} n_cells={n_cells} max_iter={max_iter}", flush=True)
    ctx = _ctx(cap)
    got, parents, trace, stats = sim.simulate(ctx, cfg, c0, cap=cap)
    want, trace_o, iters_o, steps_o = orc.simulate(cfg, make_params(), c0, cap=cap)
    assert stats["divisions"] > 0
    _compare(got, want, trace, trace_o, stats, iters_o, steps_o)
    # lineage: every id > 0 has a parent with a smaller id
    assert parents[0] == -1 and np.all(parents[1:] < np.arange(1, len(parents)))


def test_crowded_snapshot_with_mass_division(orc):
    """A 400-cell 3-D cluster where about a third of the cells are past the
    division length at the first iteration: one division batch with
    re-orientation tries, neighbor rebuilds and tilted, surface and bulk cells."""
    from bfm import sim, make_dome
    from bfm.core import make_params
    cells, _ = make_dome(400, seed=7)
    n0 = cells.shape[0]
    n_div0 = int(np.sum(cells[:, 13] > 2.4))
    assert n_div0 > 50
    cfg = sim.sim_config(max_iter=40, tableau="bogacki_shampine", daughter_angle_xy_bound=0.2,
                         daughter_angle_z_bound=0.1, seed=11)
    cap = 2 * n0 + 64
    print(f"crowded snapshot n={n0} dividing={n_div0}", flush=True)
    ctx = _ctx(cap)
    got, parents, trace, stats = sim.simulate(ctx, cfg, cells, cap=cap)
    want, trace_o, iters_o, steps_o = orc.simulate(cfg, make_params(), cells, cap=cap)
    assert stats["divisions"] >= n_div0
    _compare(got, want, trace, trace_o, stats, iters_o, steps_o)


def test_capacity_error_is_reported():
    """A division batch beyond the capacity raises instead of writing past it."""
    from bfm import sim, make_dome
    from bfm.core import BFMError
    cells, _ = make_dome(200, seed=3)
    cfg = sim.sim_config(max_iter=3, tableau="bogacki_shampine", seed=1)
    ctx = _ctx(210)
    with pytest.raises(BFMError) as e:
        sim.simulate(ctx, cfg, cells, cap=210)
    assert e.value.status == 2

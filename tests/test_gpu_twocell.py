"""Two-cell relaxation scenarios of testTwoCellSimulations.cpp on the GPU:
the 27 cases as one batched ensemble (bfm_ensemble_simulate) and a few
through bfm_simulate, each against the oracle's run of the same case and
against the contact distance 2R.

Tolerances as for the growth loop: velocities agree to ~1e-13 relative, and
over 10,000 iterations the state agrees to 4e-12 per iteration; the elapsed
time (sum of accepted dt) to 1e-6 relative."""
import numpy as np
import pytest

import twocell_cases as tc

pytestmark = pytest.mark.gpu

TOL_STATE = 4e-12 * 10000
TOL_CONTACT = 5e-7


def _ctx(cap, adhesion=False):
    from bfm import Context
    return Context(cap, max_pairs=32 * cap + 1024, params=tc.params(adhesion))


def _check(orc, got, name, d, t_end=None, iters=None, gamma=0.0, deq=2 * tc.R):
    c0 = tc.cells(name, d, gamma=gamma)
    want, trace_o, iters_o, _ = orc.simulate(tc.config(), tc.params(gamma > 0), c0, cap=8)
    assert got.shape == (2, 22)
    if iters is not None:
        assert iters == iters_o
    if t_end is not None:
        assert abs(t_end - trace_o[:iters_o].sum()) < 1e-6 * trace_o[:iters_o].sum()
    assert np.max(np.abs(got[:, 1:7] - want[:2, 1:7])) < TOL_STATE, (name, d)
    final = tc.distance(orc, got)
    if d < 2 * tc.R:
        assert abs(final - deq) < TOL_CONTACT, (name, d, final)
    else:
        assert np.array_equal(got, c0)


def test_two_cell_ensemble(orc):
    from bfm import sim
    reps = [tc.cells(name, d) for name, d in tc.CASES]
    cap = 2 * len(reps) + 64
    print(f"two-cell ensemble of {len(reps)} replicates", flush=True)
    out, t_out, it_out, st = sim.ensemble_simulate(_ctx(cap), tc.config(), reps,
                                                   list(range(len(reps))), autotune=False, cap=cap)
    assert st["failed_replicates"] == [] and st["divisions"] == 0
    assert np.all(it_out == 10000)
    for r, (name, d) in enumerate(tc.CASES):
        _check(orc, out[r], name, d, t_end=t_out[r])


@pytest.mark.parametrize("name", ["head_to_head", "shifted_T", "skew_cross"])
def test_two_cell_simulate(orc, name):
    from bfm import sim
    print(f"two-cell bfm_simulate {name} d=1.0", flush=True)
    got, parents, trace, stats = sim.simulate(_ctx(8), tc.config(), tc.cells(name, 1.0), cap=8)
    assert stats["divisions"] == 0
    _check(orc, got, name, 1.0, t_end=stats["t_end"], iters=stats["iterations"])


def test_two_cell_adhesion_ensemble(orc):
    """The 27 cases with JKR adhesion (exact contact radius) at the surface
    energy density of a 1.5 um equilibrium distance, solved by the product's
    host setup (bfm_jkr_optimal_surface_energy_density)."""
    from bfm import sim
    gamma = sim.jkr_optimal_surface_energy_density(tc.R, 0.5, 3900.0, 1.5, 100.0)
    deq = sim.jkr_equilibrium_distance(tc.R, 3900.0, gamma, 1.5 * tc.R, 1.0)
    reps = [tc.cells(name, d, gamma=gamma) for name, d in tc.CASES]
    cap = 2 * len(reps) + 64
    print(f"two-cell adhesion ensemble of {len(reps)} replicates, gamma {gamma}", flush=True)
    out, t_out, it_out, st = sim.ensemble_simulate(_ctx(cap, adhesion=True), tc.config(), reps,
                                                   list(range(len(reps))), autotune=False, cap=cap)
    assert st["failed_replicates"] == [] and np.all(it_out == 10000)
    for r, (name, d) in enumerate(tc.CASES):
        _check(orc, out[r], name, d, t_end=t_out[r], gamma=gamma, deq=deq)

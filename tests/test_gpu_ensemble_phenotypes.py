"""The batched-ensemble workload, reduced: ragged replicates whose cells carry
one of two phenotypes each (group 1: gamma 500, group 2: gamma 50 Pa um, a
seeded 50/50 draw per cell; pairs adhere with min(gamma_i, gamma_j),
mechanics.hpp:1464) advanced together with the learned JKR contact radius
(bfm_ensemble_simulate), checked against

* the oracle's loop (runSimulation restated) with the exact JKR radius, per
  replicate: the learned radius changes each contact force by at most ~1e-4
  relative, so velocities, and the accepted step sizes that the controller
  derives from them, change by that order, and over a few iterations the
  positions and orientations differ from the exact run's by that fraction of
  their displacement (asserted with a 10x margin); and
* independent bfm_simulate runs with the same learned radius: bit-identical.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_REP = 16
ITERS = 6


def _ctx(cap, surrogate):
    import bfm
    from bfm import surrogate as sur
    from bfm.core import make_params, JKR_SURROGATE, JKR_EXACT
    p = make_params(jkr_mode=JKR_SURROGATE if surrogate else JKR_EXACT)
    c = bfm.Context(cap, max_pairs=32 * cap + 1024, params=p)
    if surrogate:
        sur.install(c, sur.load())
    return c


@pytest.fixture(scope="module")
def batch():
    import bfm
    from bfm import sim
    sizes = bfm.ensemble_sizes(N_REP, 40, 400, seed=0)
    reps, groups = bfm.make_ensemble(sizes, seeds=range(N_REP), hl_hi=1.2)
    for g in groups:
        assert abs(int((g == 1).sum()) - int((g == 2).sum())) <= 1
    cfg = sim.sim_config(max_iter=ITERS, tableau="bogacki_shampine", seed=0, iter_update_neighbors=1)
    cap = 2 * int(sizes.sum()) + 64
    print(f"phenotype ensemble: {N_REP} replicates, {int(sizes.sum())} cells", flush=True)
    out, t_out, it_out, st = sim.ensemble_simulate(_ctx(cap, True), cfg, reps, list(range(N_REP)),
                                                   autotune=False, cap=cap)
    assert st["failed_replicates"] == [] and np.all(it_out == ITERS)
    return reps, cfg, out, t_out


def test_phenotype_ensemble_against_exact_oracle(orc, batch):
    from bfm.core import make_params
    reps, cfg, out, t_out = batch
    p_exact = make_params()
    worst = 0.0
    for r, cells in enumerate(reps):
        cfg.seed = r
        ref, trace, iters, _ = orc.simulate(cfg, p_exact, cells, cap=2 * cells.shape[0] + 64)
        assert iters == ITERS and ref.shape == out[r].shape, r
        assert np.array_equal(ref[:, 20:22], out[r][:, 20:22])         # groups and gamma carried
        disp = np.max(np.abs(ref[:, 1:7] - cells[:, 1:7]))
        diff = np.max(np.abs(out[r][:, 1:7] - ref[:, 1:7]))
        assert disp > 0
        worst = max(worst, diff / disp)
        assert abs(t_out[r] - trace.sum()) <= 1e-3 * trace.sum(), (r, t_out[r], trace.sum())
    print(f"learned radius vs exact oracle: worst position/orientation difference "
          f"{worst:.2e} of the displacement", flush=True)
    assert worst <= 1e-3


def test_phenotype_ensemble_equals_independent_runs(batch):
    from bfm import sim
    reps, cfg, out, t_out = batch
    for r in (0, 5, N_REP - 1):
        cfg.seed = r
        cap = 2 * reps[r].shape[0] + 64
        want, _, _, s1 = sim.simulate(_ctx(cap, True), cfg, reps[r], cap=cap)
        assert np.array_equal(out[r], want), r
        assert t_out[r] == s1["t_end"]

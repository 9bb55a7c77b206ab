"""Batched replicates (bfm_ensemble_simulate) against independent bfm_simulate
runs of the same replicates with the same division streams.

The batched loop shares the neighbor build (per-replicate grids, no pairs
across replicates), one kernel per RK attempt and per-replicate dt,
controller and division; with the list rebuilt every iteration in both runs
the arithmetic is the same, so each replicate must come out bit-identical.
The 64- and 48-replicate batches run the controller's steps as kernels of
their own (more than 16 work-groups) while most single runs fold them into
the attempt's last work-group, so the two forms are checked against each
other too."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx(cap):
    from bfm import Context
    from bfm.core import make_params
    return Context(cap, max_pairs=32 * cap + 1024, params=make_params())


def _replicates():
    from bfm import sim, make_dome
    a = sim.founder_cell()
    b, _ = make_dome(150, seed=5)
    c, _ = make_dome(300, seed=9, gamma=500.0)
    d, _ = make_dome(40, seed=2, gamma=50.0)
    return [a, b, c, d]


@pytest.mark.parametrize("tableau", ["bogacki_shampine", "heun_euler"])
def test_ensemble_equals_independent_runs(tableau):
    from bfm import sim
    reps = _replicates()
    seeds = [3, 5, 9, 2]
    cfg = sim.sim_config(max_iter=30, tableau=tableau, iter_update_neighbors=1,
                         daughter_angle_xy_bound=0.2, seed=0)
    cap = 2 * sum(r.shape[0] for r in reps) + 64
    print(f"ensemble of {len(reps)} replicates, {tableau}", flush=True)
    out, t_out, it_out, st = sim.ensemble_simulate(_ctx(cap), cfg, reps, seeds, autotune=False, cap=cap)
    assert st["divisions"] > 0 and st["attempts"] >= st["iterations"] == 30
    for r, cells in enumerate(reps):
        cfg.seed = seeds[r]
        capr = 2 * cells.shape[0] + 64
        want, _, trace, s1 = sim.simulate(_ctx(capr), cfg, cells, cap=capr)
        assert out[r].shape == want.shape, r
        assert it_out[r] == s1["iterations"]
        assert t_out[r] == s1["t_end"]
        assert np.array_equal(out[r], want), (r, np.max(np.abs(out[r] - want)))


def test_ensemble_many_replicates_autotuned():
    """64 uneven replicates with the autotuner on: every replicate advances,
    the tuner picks a valid configuration, and replicates agree bit for bit
    with independent runs under the same rebuild schedule (the list rebuilt
    every iteration; with a longer period the batch rebuilds all replicates
    whenever one divides, so a replicate can see a fresher list than alone,
    which matters only when cells cross the 0.4 um skin between rebuilds, as
    in replicate 37's unrelaxed start). Replicate 47
    draws a daughter length ratio of 4e-5 at its first division (a 4 sigma
    draw of the clipped N(0.5, 0.125)); its rotational drag ~ l^3 makes the
    next step non-finite, as it does for the independent run (and the
    reference's formulation): the ensemble freezes it and flags it."""
    from bfm import sim, make_dome
    rng = np.random.default_rng(0)
    sizes = np.rint(np.exp(rng.uniform(np.log(30), np.log(300), 64))).astype(int)
    reps = [make_dome(int(n), seed=r, gamma=(50.0, 500.0)[r % 2])[0] for r, n in enumerate(sizes)]
    seeds = list(range(64))
    cfg = sim.sim_config(max_iter=12, tableau="bogacki_shampine", seed=0, iter_update_neighbors=1)
    cap = 2 * int(sizes.sum()) + 64
    print(f"ensemble of 64 replicates, {int(sizes.sum())} cells", flush=True)
    out, t_out, it_out, st = sim.ensemble_simulate(_ctx(cap), cfg, reps, seeds, autotune=True, cap=cap)
    assert 0 <= st["tuned_variant"] <= 8 and st["tuned_bin_scale"] in (0.5, 0.75, 1.0, 1.25)
    assert st["failed_replicates"] == [47]
    ok = np.ones(64, bool)
    ok[47] = False
    assert np.all(it_out[ok] == 12) and np.all(t_out[ok] > 0) and it_out[47] < 12
    assert st["n_final"] == sum(o.shape[0] for o in out) >= sizes.sum()
    for r in (0, 37):
        cfg.seed = seeds[r]
        capr = 2 * reps[r].shape[0] + 64
        want, _, _, _ = sim.simulate(_ctx(capr), cfg, reps[r], cap=capr)
        assert np.array_equal(out[r], want)


def test_failed_replicate_matches_independent_failure():
    """The replicate frozen by the ensemble fails in an independent run too."""
    from bfm import sim, make_dome
    from bfm.core import BFMError
    rng = np.random.default_rng(0)
    sizes = np.rint(np.exp(rng.uniform(np.log(30), np.log(300), 64))).astype(int)
    cells = make_dome(int(sizes[47]), seed=47, gamma=500.0)[0]
    cfg = sim.sim_config(max_iter=12, tableau="bogacki_shampine", seed=47)
    cap = 2 * cells.shape[0] + 64
    with pytest.raises(BFMError) as e:
        sim.simulate(_ctx(cap), cfg, cells, cap=cap)
    assert e.value.status == 3


def test_large_batch_multi_slice_error_maxima():
    """618k cells in 48 uneven replicates: the RK update's blocks then cover
    more than one 256-row slice each (chunk = n / 2048 blocks rounded up to
    256), so the per-lane running error maximum that a wavefront keeps while
    its slices stay in one replicate, and its flush when the replicate
    changes, decide every replicate's step-size control. A wrong maximum
    changes the accepted step or the retries, so two lockstep iterations must
    match independent runs of replicates that span several slices, bit for
    bit."""
    import bfm
    from bfm import sim
    sizes = bfm.ensemble_sizes(48, 4000, 30000, seed=3)
    reps, _ = bfm.make_ensemble(sizes, seeds=range(48), hl_hi=1.2)
    n = int(sizes.sum())
    assert n > 2048 * 256   # every block of the update takes two or more slices
    cfg = sim.sim_config(max_iter=2, tableau="bogacki_shampine", seed=0, iter_update_neighbors=1)
    cap = n + 4096
    print(f"multi-slice ensemble: {n} cells", flush=True)
    out, t_out, it_out, st = sim.ensemble_simulate(_ctx(cap), cfg, reps, list(range(48)), autotune=False, cap=cap)
    assert np.all(it_out == 2) and st["failed_replicates"] == []
    for r in (int(np.argmax(sizes)), int(np.argmin(sizes)), 17):
        cfg.seed = r
        capr = reps[r].shape[0] + 1024
        want, _, _, s1 = sim.simulate(_ctx(capr), cfg, reps[r], cap=capr)
        assert t_out[r] == s1["t_end"], r
        assert np.array_equal(out[r], want), (r, np.max(np.abs(out[r] - want)))

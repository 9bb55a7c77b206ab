"""Cell-cell kinetic friction on the GPU (friction_mode 1: cellCellFrictionForces,
ABM-3d/ABM-3d-cpp/include/mechanics.hpp:1031-1175, added to the cell-cell forces
at :1341-1347 and :1650-1656) against the oracle on identical inputs.

Bars: velocities, conservative forces and RK steps within 1e-10 normwise per
column (the friction rows are summed in the reference's row order on both
sides); growth runs as in test_gpu_sim (identical division events and lineage,
state within 4e-12 per iteration); replicates of a batched run bit-identical
to independent runs. Two geometries: a 3-D dome with random stored velocities,
and a monolayer lying on the surface whose nearly horizontal contacts take
the 2-D branch (|d_z| / |d| < 1e-3, contact points below 1.01 R) -- in the RK
stages that test reads the stage positions."""
import numpy as np
import pytest

from gpu_common import LDIV, R, THR, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _with_eta(cells, seed, vel_scale=3.0):
    rng = np.random.default_rng(seed)
    n = cells.shape[0]
    c = np.zeros((n, 23), order="F")
    c[:, :22] = cells[:, :22]
    c[:, 22] = rng.uniform(0.2, 2.0, n)
    c[:, 7:13] = rng.normal(scale=vel_scale, size=(n, 6))
    return c


@pytest.fixture(scope="module")
def dome():
    import bfm
    from oracle import oracle as orc
    cells, _ = bfm.make_dome(6000, seed=41)
    cells = _with_eta(cells, 42)
    nb = orc.get_cell_neighbors(cells, THR, R, LDIV, binned=True)
    return cells, nb


@pytest.fixture(scope="module")
def layer():
    """Horizontal cells on the surface (n_z = 0, the well-conditioned horizontal
    branch of the velocity solve), 60 % resting near 0.75 um and 40 % at
    0.80-0.815 um, across the 1.01 R = 0.808 um contact-point test."""
    from oracle import oracle as orc
    rng = np.random.default_rng(43)
    n = 900
    cells = np.zeros((n, 23), order="F")
    cells[:, 0] = np.arange(n)
    cells[:, 1:3] = rng.uniform(0, 60, (n, 2))
    hi = rng.uniform(0, 1, n) < 0.4
    cells[:, 3] = np.where(hi, rng.uniform(0.80, 0.815, n), 0.75 + rng.uniform(0, 1e-3, n))
    th = rng.uniform(0, 2 * np.pi, n)
    cells[:, 4], cells[:, 5] = np.cos(th), np.sin(th)
    hl = rng.uniform(0.4, 1.2, n)
    cells[:, 13], cells[:, 14] = 2 * hl, hl
    cells[:, 16], cells[:, 17], cells[:, 18], cells[:, 19], cells[:, 20] = 1.5, 0.7, 720, 100, 1
    cells[:, 21] = 100.0
    cells[:, 22] = rng.uniform(0.2, 2.0, n)
    cells[:, 7:13] = rng.normal(scale=50.0, size=(n, 6))
    cells[:, 9] = cells[:, 12] = 0.0
    nb = orc.get_cell_neighbors(cells, THR, R, LDIV)
    d = nb[:, 2:5]
    dist = np.linalg.norm(d, axis=1)
    flat = (dist < 2 * R) & (np.abs(d[:, 2]) / dist < 1e-3)
    assert flat.sum() > 100 and (~flat & (dist < 2 * R)).sum() > 100
    return cells, nb


def _params(**kw):
    from bfm.core import make_params
    return make_params(friction=True, **kw)


@pytest.mark.parametrize("geom", ["dome", "layer"])
@pytest.mark.parametrize("coulomb", [1.0, 0.0])
def test_velocities_and_forces(orc, request, geom, coulomb):
    import bfm
    cells, nb = request.getfixturevalue(geom)
    p = _params(cell_cell_coulomb_coeff=coulomb)
    c = bfm.Context(cells.shape[0], max_pairs=max(nb.shape[0], 16), params=p)
    a2d = (cells[:, 6] < p.nz_threshold).astype(np.int32)
    v = c.get_velocities(cells, nb, a2d)
    v_ref = orc.get_velocities(cells, nb, a2d, None, p)
    print(f"{geom} coulomb={coulomb} vel rel_err={rel_err(v, v_ref):.3e}", flush=True)
    assert rel_err(v, v_ref) <= TOL
    # friction changes the result (the test is not vacuous)
    v_nf = orc.get_velocities(cells, nb, a2d, None, bfm.core.make_params(ncols=23))
    assert rel_err(v_nf, v_ref) > 1e-6
    f = c.get_conservative_forces(cells, nb, a2d)
    assert rel_err(f, orc.conservative_forces(cells, nb, a2d, p)) <= TOL


@pytest.mark.parametrize("geom", ["dome", "layer"])
@pytest.mark.parametrize("name", ["bogacki_shampine", "dormand_prince"])
def test_step_rk(orc, request, geom, name):
    import bfm
    cells, nb = request.getfixturevalue(geom)
    p = _params()
    A, b, bs, _ = bfm.tableau(name)
    c = bfm.Context(cells.shape[0], max_pairs=max(nb.shape[0], 16), params=p)
    dt = 2e-5 if geom == "dome" else 1e-7
    new, err = c.step_rk(A, b, bs, dt, cells, nb)
    new_ref, err_ref = orc.step_rk(A, b, bs, dt, cells, nb, p)
    print(f"{geom} {name} rel_err={rel_err(new[:, 1:13], new_ref[:, 1:13]):.3e}", flush=True)
    assert rel_err(new[:, 1:13], new_ref[:, 1:13]) <= TOL
    assert np.array_equal(new[:, [0] + list(range(13, 23))], new_ref[:, [0] + list(range(13, 23))])
    disp = np.max(np.abs(new_ref[:, 1:7] - cells[:, 1:7]), axis=0)
    assert np.all(np.max(np.abs(err - err_ref), axis=0) <= 1e-10 * disp)


def _founder(eta=0.5):
    from bfm import sim
    c0 = sim.founder_cell(ncols=23)
    c0[0, 22] = eta
    return c0


def test_founder_growth_with_friction_matches_oracle(orc):
    """Founder through its first division with friction between the
    daughters. After the division the controller runs near 1e-7 h, where
    retries are threshold tests on an error of order 1 (one implementation can
    take a retry the other does not, as in the tilted-daughter case of
    test_gpu_sim), so the accepted dt is compared in the median (1e-6) and
    the state to 1e-6 um."""
    from bfm import sim, Context
    cfg = sim.sim_config(max_iter=20000, tableau="bogacki_shampine", seed=42)
    p = _params()
    print("founder growth with friction, 20000 iterations", flush=True)
    ctx = Context(64, max_pairs=32 * 64 + 1024, params=p)
    got, parents, trace, stats = sim.simulate(ctx, cfg, _founder(), cap=64)
    want, trace_o, iters_o, steps_o = orc.simulate(cfg, p, _founder(), cap=64)
    assert stats["divisions"] > 0 and stats["iterations"] == iters_o == 20000
    assert abs(stats["rk_steps"] - steps_o) <= max(2, steps_o // 1000)
    assert got.shape == want.shape and np.array_equal(got[:, [0, 20]], want[:, [0, 20]])
    for cols in (slice(1, 7), slice(13, 15), slice(22, 23)):
        assert np.max(np.abs(got[:, cols] - want[:, cols])) < 1e-6
    assert np.median(np.abs(trace - trace_o) / trace_o) < 1e-6


def test_batched_replicates_with_friction_equal_independent_runs():
    """The captured batched attempt computes the friction rows per stage at
    the stage positions; with the list rebuilt every iteration each replicate
    is bit-identical to its own bfm_simulate run."""
    from bfm import sim, Context
    cfg = sim.sim_config(max_iter=15000, tableau="bogacki_shampine", iter_update_neighbors=1)
    p = _params()
    reps = [_founder(0.3), _founder(0.8), _founder(1.5)]
    seeds = [5, 6, 7]
    print("3 replicates with friction, 15000 iterations", flush=True)
    ctx = Context(128, max_pairs=64 * 128, params=p)
    out, t_out, it_out, st = sim.ensemble_simulate(ctx, cfg, reps, seeds, autotune=False)
    for r, (c0, sd) in enumerate(zip(reps, seeds)):
        cfg.seed = sd
        ctx1 = Context(64, max_pairs=64 * 64, params=p)
        got, _, _, stats = sim.simulate(ctx1, cfg, c0, cap=64)
        assert np.array_equal(out[r], got), r
        assert it_out[r] == stats["iterations"]

"""Anisotropic JKR adhesion on the GPU (adhesion_mode 2, the table path of
forcesAnisotropicJKRLagrange, ABM-3d/ABM-3d-cpp/include/adhesion.hpp:734-1046,
through jkr_data->getJKRForces, mechanics.hpp:1468-1479) against the oracle on
identical inputs and tables.

Each contact's JKR force is a table entry picked by nearest-value lookups on
angles, half-lengths, centerline coordinates, principal radii and the
overlap; the device forms those angles with the CPU's operation order (no FMA
contraction), so both sides pick the same entries and the bars are those of
the isotropic path: velocities, conservative forces and steps within 1e-10
normwise per column, growth runs with identical division events and state
within 4e-12 per iteration, batched replicates bit-identical to independent
runs, and the driver's frames (10 digits) within 5e-10 of the loop called
directly."""
import json

import numpy as np
import pytest

from gpu_common import LDIV, R, THR, random_box, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def tables():
    from bfm.jkr import synthetic_aniso_tables
    return {"fixed": synthetic_aniso_tables(),
            "variable": synthetic_aniso_tables(gammas=np.linspace(0.0, 200.0, 6))}


@pytest.fixture(scope="module")
def dome():
    import bfm
    from oracle import oracle as orc
    cells, _ = bfm.make_dome(6000, seed=61)
    rng = np.random.default_rng(62)
    cells[:, 21] = rng.uniform(20.0, 180.0, cells.shape[0])
    cells[:, 7:13] = rng.normal(scale=3.0, size=(cells.shape[0], 6))
    nb = orc.get_cell_neighbors(cells, THR, R, LDIV, binned=True)
    return cells, nb


@pytest.fixture(scope="module")
def box():
    from oracle import oracle as orc
    cells = random_box(1500, 20.0, 63, zmin=0.3)
    cells[:, 21] = np.random.default_rng(64).uniform(20.0, 180.0, 1500)
    nb = orc.get_cell_neighbors(cells, THR, R, LDIV)
    return cells, nb


def _ctx(cells, nb, t, **kw):
    import bfm
    from bfm.core import make_params
    p = make_params(anisotropic=True, **kw)
    c = bfm.Context(cells.shape[0], max_pairs=max(nb.shape[0], 16), params=p)
    c.set_aniso_jkr_tables(t)
    return c, p


@pytest.mark.parametrize("geom", ["dome", "box"])
@pytest.mark.parametrize("kind", ["fixed", "variable"])
def test_velocities_and_forces(orc, request, tables, geom, kind):
    from bfm.core import make_params
    cells, nb = request.getfixturevalue(geom)
    t = tables[kind]
    c, p = _ctx(cells, nb, t)
    orc.set_aniso_tables(t)
    a2d = (cells[:, 6] < p.nz_threshold).astype(np.int32)
    v = c.get_velocities(cells, nb, a2d)
    v_ref = orc.get_velocities(cells, nb, a2d, None, p)
    print(f"{geom} {kind} vel rel_err={rel_err(v, v_ref):.3e}", flush=True)
    assert rel_err(v, v_ref) <= TOL
    f = c.get_conservative_forces(cells, nb, a2d)
    assert rel_err(f, orc.conservative_forces(cells, nb, a2d, p)) <= TOL
    # the tables change the result against the isotropic exact path
    v_iso = orc.get_velocities(cells, nb, a2d, None, make_params())
    assert rel_err(v_iso, v_ref) > 1e-4


@pytest.mark.parametrize("name", ["bogacki_shampine", "dormand_prince"])
def test_step_rk(orc, dome, tables, name):
    import bfm
    cells, nb = dome
    t = tables["fixed"]
    c, p = _ctx(cells, nb, t)
    orc.set_aniso_tables(t)
    A, b, bs, _ = bfm.tableau(name)
    new, err = c.step_rk(A, b, bs, 2e-5, cells, nb)
    new_ref, err_ref = orc.step_rk(A, b, bs, 2e-5, cells, nb, p)
    print(f"{name} rel_err={rel_err(new[:, 1:13], new_ref[:, 1:13]):.3e}", flush=True)
    assert rel_err(new[:, 1:13], new_ref[:, 1:13]) <= TOL
    disp = np.max(np.abs(new_ref[:, 1:7] - cells[:, 1:7]), axis=0)
    assert np.all(np.max(np.abs(err - err_ref), axis=0) <= 1e-10 * disp)


def test_verlet_step(orc, box, tables):
    """include_constraint = true: the JKR torque form s(-(n.v) n - v)."""
    cells, nb = box
    t = tables["variable"]
    c, p = _ctx(cells, nb, t)
    orc.set_aniso_tables(t)
    new = c.step_verlet(1e-5, 1e-3, cells, nb)
    ref = orc.step_verlet(1e-5, 1e-3, cells, nb, p)
    print(f"verlet rel_err={rel_err(new[:, 1:13], ref[:, 1:13]):.3e}", flush=True)
    assert rel_err(new[:, 1:13], ref[:, 1:13]) <= TOL


def test_founder_growth_matches_oracle(orc, tables):
    from bfm import sim, Context
    from bfm.core import make_params
    t = tables["fixed"]
    cfg = sim.sim_config(n_cells=6, tableau="bogacki_shampine", seed=42)
    p = make_params(anisotropic=True)
    orc.set_aniso_tables(t)
    print("founder growth to 6 cells, anisotropic JKR tables", flush=True)
    ctx = Context(64, max_pairs=32 * 64 + 1024, params=p)
    ctx.set_aniso_jkr_tables(t)
    c0 = sim.founder_cell()
    got, parents, trace, stats = sim.simulate(ctx, cfg, c0, cap=64)
    want, trace_o, iters_o, steps_o = orc.simulate(cfg, p, c0, cap=64)
    assert stats["divisions"] >= 5 and stats["iterations"] == iters_o
    assert got.shape == want.shape and np.array_equal(got[:, [0, 20]], want[:, [0, 20]])
    tol = 4e-12 * iters_o
    for cols in (slice(1, 7), slice(13, 15)):
        scale = max(1.0, np.max(np.abs(want[:, cols])))
        assert np.max(np.abs(got[:, cols] - want[:, cols])) < tol * scale


def test_batched_replicates_equal_independent_runs(tables):
    from bfm import sim, Context
    from bfm.core import make_params
    t = tables["variable"]
    p = make_params(anisotropic=True)
    cfg = sim.sim_config(max_iter=8000, tableau="bogacki_shampine", iter_update_neighbors=1)
    reps = [sim.founder_cell(gamma=g) for g in (40.0, 100.0, 170.0)]
    seeds = [11, 12, 13]
    print("3 replicates, anisotropic tables, 8000 iterations", flush=True)
    ctx = Context(128, max_pairs=64 * 128, params=p)
    ctx.set_aniso_jkr_tables(t)
    out, _, it_out, _ = sim.ensemble_simulate(ctx, cfg, reps, seeds, autotune=False)
    for r, (c0, sd) in enumerate(zip(reps, seeds)):
        cfg.seed = sd
        ctx1 = Context(64, max_pairs=64 * 64, params=p)
        ctx1.set_aniso_jkr_tables(t)
        got, _, _, stats = sim.simulate(ctx1, cfg, c0, cap=64)
        assert np.array_equal(out[r], got), r
        assert it_out[r] == stats["iterations"]


def test_driver_reads_the_table_files(tmp_path, tables):
    """bfm_run with adhesion_mode 2 and precompute_jkr_forces 0 reads the two
    files (simulation.hpp:1405-1441) and runs the same trajectory as the loop
    given the tables read back from them."""
    from bfm import io, sim
    from bfm.core import make_params, Context
    from bfm.jkr import write_aniso_table_files
    from test_gpu_run import BASE
    cf, ff = tmp_path / "curv.txt", tmp_path / "forces.txt"
    write_aniso_table_files(tables["fixed"], cf, ff)
    cfg_json = dict(BASE, adhesion_mode=2, adhesion_precompute_jkr_forces=0,
                    adhesion_curvature_filename=str(cf), adhesion_jkr_forces_filename=str(ff),
                    n_cells=6)
    cfgp = tmp_path / "cfg.json"
    cfgp.write_text(json.dumps(cfg_json))
    prefix = str(tmp_path / "run")
    print("bfm_run founder -> 6 cells, anisotropic tables from files", flush=True)
    st = io.run(cfgp, prefix, 42, max_cells=64)
    final, pf = io.readCells(prefix + "_final.txt")
    assert pf["adhesion_mode"] == "2" and pf["adhesion_precompute_jkr_forces"] == "0"
    t = io.readAnisotropicJKRTables(cf, ff)
    gamma = sim.jkr_optimal_surface_energy_density(0.8, 0.5, 3900.0, BASE["adhesion_eqdist"],
                                                   init_gamma=100.0)
    cfg = sim.sim_config(n_cells=6, tableau="bogacki_shampine", seed=42)
    ctx = Context(64, max_pairs=32 * 64 + 1024, params=make_params(anisotropic=True))
    ctx.set_aniso_jkr_tables(t)
    got, parents, _, stats = sim.simulate(ctx, cfg, sim.founder_cell(gamma=gamma), cap=64)
    assert stats["iterations"] == st["iterations"] and got.shape == final.shape
    scale = np.maximum(np.abs(got), 1e-300)
    assert np.max(np.abs(final - got)[:, 1:20] / scale[:, 1:20]) <= 5e-10
    assert np.array_equal(final[:, [0, 20]], got[:, [0, 20]])

"""The driver (bfm_run): a JSON configuration in the reference's keys, frames
and lineage written the reference's way, and the same trajectory as the
growth loop called directly (bfm_simulate) with the parameters the driver
derives (surface energy density from adhesion_eqdist, founder cell of
abmConstAdhesion.cpp:405-421). Frames hold 10 significant digits, so the
final frame matches the direct run's cells to 5e-10 relative."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASE = {"R": 0.8, "Rcell": 0.5, "L0": 0.4, "E0": 3900.0, "Ecell": 3.9e6, "sigma0": 100.0,
        "eta_ambient": 0.7, "eta_surface": 720.0, "max_stepsize": 1e-4, "min_stepsize": 1e-9,
        "dt_write": 0.25, "iter_update_stepsize": 1, "iter_update_neighbors": 10,
        "max_iter": 0, "n_cells": 8, "max_time": 0.0, "growth_mean": 1.5, "growth_std": 0.15,
        "daughter_length_std": 0.125, "daughter_angle_xy_bound": 0.05,
        "daughter_angle_z_bound": 0.0, "nz_threshold": 1e-8, "max_rxy_noise": 0.0,
        "max_rz_noise": 0.0, "max_nxy_noise": 0.0, "max_nz_noise": 0.0,
        "max_error_allowed": 1e-8, "basal_only": 0, "adhesion_mode": 1,
        "cell_cell_friction_mode": 0, "adhesion_eqdist": 1.4, "integration_mode": 2}


def _frames(prefix):
    d, base = os.path.split(prefix)
    its = sorted(int(f[len(base) + 5:-4]) for f in os.listdir(d)
                 if f.startswith(base + "_iter") and f.endswith(".txt"))
    return its


def test_run_writes_reference_outputs_and_matches_simulate(tmp_path):
    from bfm import io, sim
    from bfm.core import make_params, Context, JKR_EXACT
    cfgp = tmp_path / "cfg.json"
    cfgp.write_text(json.dumps(BASE))
    prefix = str(tmp_path / "run")
    print("bfm_run founder -> 8 cells, exact JKR", flush=True)
    st = io.run(cfgp, prefix, 42, jkr_mode=JKR_EXACT, max_cells=64, verbose=True)
    assert st["divisions"] >= 7

    init, p0 = io.readCells(prefix + "_init.txt")
    assert init.shape == (1, 22) and p0["t_curr"] == "0"
    final, pf = io.readCells(prefix + "_final.txt")
    assert final.shape[0] >= 8 and final.shape[1] == 22
    for key in ("R", "Rcell", "L0", "Ldiv", "E0", "Ecell", "dt_write", "random_seed",
                "adhesion_eqdist", "integration_mode", "neighbor_threshold", "growth_mean1"):
        assert key in pf, key
    assert pf["random_seed"] == "42" and pf["integration_mode"] == "2" and pf["Ldiv"] == "2.4"
    assert float(pf["t_curr"]) == pytest.approx(st["t_end"], rel=1e-9)
    lin = io.readLineage(prefix + "_lineage.txt")
    ids = final[:, 0].astype(int)
    # every division retires the mother's id and gives both daughters new ids
    assert len(lin) == 1 + 2 * st["divisions"] and lin[0] == -1
    assert len(set(ids)) == len(ids) and ids.max() < len(lin) and set(ids).isdisjoint(lin[lin >= 0])
    assert np.all(lin[1:] < np.arange(1, len(lin)))

    # a frame at each dt_write crossing, named by iteration, times increasing
    its = _frames(prefix)
    assert len(its) >= int(st["t_end"] / BASE["dt_write"]) - 1 and its == sorted(set(its))
    ts = [float(io.readCells(f"{prefix}_iter{k}.txt")[1]["t_curr"]) for k in its]
    assert all(b > a for a, b in zip(ts, ts[1:]))
    assert all(int(np.floor(t / BASE["dt_write"] + 1e-9)) == j + 1 for j, t in enumerate(ts))

    # the same run through bfm_simulate
    gamma = sim.jkr_optimal_surface_energy_density(0.8, 0.5, 3900.0, 1.4, init_gamma=100.0)
    c0 = sim.founder_cell(gamma=gamma)
    assert np.allclose(c0, init, rtol=5e-10, atol=0)
    cfg = sim.sim_config(n_cells=8, tableau="bogacki_shampine", seed=42)
    ctx = Context(64, max_pairs=32 * 64 + 1024, params=make_params())
    got, parents, _, stats = sim.simulate(ctx, cfg, c0, cap=64)
    assert stats["iterations"] == st["iterations"] and got.shape == final.shape
    scale = np.maximum(np.abs(got), 1e-300)
    assert np.max(np.abs(final - got)[:, 1:20] / scale[:, 1:20]) <= 5e-10
    assert np.array_equal(final[:, [0, 20]], got[:, [0, 20]])
    assert np.array_equal(parents, lin)


def test_run_resumes_from_its_own_frame(tmp_path):
    """init_filename + lineage_filename (abmConstAdhesion.cpp:348-401): the run
    starts from the saved population and lineage, and new ids continue after
    the saved ones. One saved cell is set just past the division length so
    that it divides in the first iteration. (A resumed run restarts from
    positions rounded to 10 digits, which excites the stiff vertical surface
    mode: the controller then runs near min_stepsize, in the reference as here,
    so the continuation is bounded by max_iter.)"""
    from bfm import io
    from bfm.core import JKR_EXACT, JKR_TABLE_1D
    cfg = dict(BASE, n_cells=4)
    (tmp_path / "a.json").write_text(json.dumps(cfg))
    print("bfm_run founder -> 4 cells, table JKR", flush=True)
    io.run(tmp_path / "a.json", str(tmp_path / "a"), 7, jkr_mode=JKR_TABLE_1D, max_cells=64, verbose=True)
    a, pa = io.readCells(tmp_path / "a_final.txt")
    lin_a = io.readLineage(tmp_path / "a_lineage.txt")
    k = int(np.argmax(a[:, 13]))
    a[k, 13] = 2.4 + 1e-6
    a[k, 14] = a[k, 13] / 2
    io.writeCells(a, pa, tmp_path / "start.txt", {21: 1})
    cfg2 = dict(BASE, n_cells=0, max_iter=50, init_filename=str(tmp_path / "start.txt"),
                lineage_filename=str(tmp_path / "a_lineage.txt"))
    (tmp_path / "b.json").write_text(json.dumps(cfg2))
    print("bfm_run resume for 50 iterations, exact JKR", flush=True)
    st = io.run(tmp_path / "b.json", str(tmp_path / "b"), 8, jkr_mode=JKR_EXACT, max_cells=64)
    assert st["iterations"] == 50 and st["divisions"] == 1
    b_init, _ = io.readCells(tmp_path / "b_init.txt")
    assert np.array_equal(b_init[:, :21], a[:, :21])
    b, _ = io.readCells(tmp_path / "b_final.txt")
    lin = io.readLineage(tmp_path / "b_lineage.txt")
    mother = int(a[k, 0])
    assert np.array_equal(lin[:len(lin_a)], lin_a) and len(lin) == len(lin_a) + 2
    assert list(lin[len(lin_a):]) == [mother, mother]
    ids = set(b[:, 0].astype(int))
    assert ids == (set(a[:, 0].astype(int)) - {mother}) | {len(lin_a), len(lin_a) + 1}

"""Frames, lineage files and the JSON configuration of the 3-D driver
(bfm_write_cells / bfm_read_cells / bfm_write_lineage / bfm_read_lineage /
bfm_load_config), against the formats of ABM-3d/ABM-3d-cpp/include/utils.hpp:
95-258 and the keys of src/abmConstAdhesion.cpp:31-333. Host code only.

The expected texts below are written out literally: writeCells streams each
value with setprecision(10), which the C++ standard defines as printf's %.10g
(so 1e-05 keeps its two-digit exponent and -0.0 prints as -0), the id and
group columns go through static_cast<int> (truncation toward zero), the last
tab of a row is removed and the parameter lines come first, sorted by key
(std::map)."""
import json

import numpy as np
import pytest


def _row(vals):
    c = np.zeros((1, 22), order="F")
    c[0, :len(vals)] = vals
    return c


def test_write_cells_known_text(tmp_path):
    from bfm import io
    cells = np.zeros((2, 22), order="F")
    cells[0, :21] = [3, 0.1, -2.5, 0.7, 1, 0, 0, 1e-5, -0.0, 123456789012.0, 0, 0, 0,
                     1.0, 0.5, 0.0, 1.5, 0.7, 720, 100, 1]
    cells[0, 21] = 99.12345678912
    cells[1, :21] = [4.9, 1.0 / 3.0, 2, 3, 0, 1, 0, 0, 0, 0, 0, 0, 0,
                     2, 1, 0.25, 1.5, 0.7, 720, 100, 2.7]
    cells[1, 21] = 5
    path = tmp_path / "f.txt"
    io.writeCells(cells, {"t_curr": "0.5", "R": "0.8", "adhesion_mode": "1"}, path, {21: 1})
    want = ("# R = 0.8\n# adhesion_mode = 1\n# t_curr = 0.5\n"
            "3\t0.1\t-2.5\t0.7\t1\t0\t0\t1e-05\t-0\t1.23456789e+11\t0\t0\t0\t1\t0.5\t0\t1.5\t0.7"
            "\t720\t100\t1\t99.12345679\n"
            "4\t0.3333333333\t2\t3\t0\t1\t0\t0\t0\t0\t0\t0\t0\t2\t1\t0.25\t1.5\t0.7\t720\t100\t2\t5\n")
    assert path.read_text() == want
    # an extra column that is not listed is left out; type 0 writes an integer
    io.writeCells(cells, {}, path, {})
    assert path.read_text().splitlines()[0].count("\t") == 20
    io.writeCells(cells, {}, path, {21: 0})
    assert path.read_text().splitlines()[0].endswith("\t1\t99")


def test_read_cells_roundtrip_and_params(tmp_path):
    from bfm import io
    rng = np.random.default_rng(3)
    n = 50
    cells = np.asfortranarray(rng.normal(size=(n, 22)) * 10)
    cells[:, 0] = np.arange(n)
    cells[:, 20] = rng.integers(1, 3, n)
    path = tmp_path / "r.txt"
    params = {"b": "2", "a": "x y = z"}
    io.writeCells(cells, params, path, {21: 1})
    got, p = io.readCells(path)
    assert got.shape == (n, 22) and got.flags.f_contiguous
    assert p == {"a": "x y = z", "b": "2"}
    assert np.array_equal(got[:, [0, 20]], cells[:, [0, 20]])
    # 10 significant digits
    rel = np.abs(got - cells) / np.maximum(np.abs(cells), 1e-300)
    assert np.max(rel[:, 1:20]) <= 5e-10 and np.max(rel[:, 21]) <= 5e-10
    # writing what was read gives the same text
    path2 = tmp_path / "r2.txt"
    io.writeCells(got, p, path2, {21: 1})
    assert path2.read_text() == path.read_text()


def test_read_cells_ragged_rows_widen_the_array(tmp_path):
    """readCells grows the column count to the widest row (utils.hpp:127-134)."""
    from bfm import io
    path = tmp_path / "w.txt"
    base = "\t".join(["1"] * 21)
    path.write_text("# k = v\n" + base + "\n" + base + "\t7\t8\n")
    cells, p = io.readCells(path)
    assert cells.shape == (2, 23) and p == {"k": "v"}
    assert cells[1, 21] == 7 and cells[1, 22] == 8 and cells[0, 21] == 0
    path.write_text(base + "\n" + "1\tnot-a-number\n")
    from bfm import BFMError
    with pytest.raises(BFMError):
        io.readCells(path)


def test_lineage_roundtrip_and_order(tmp_path):
    from bfm import io
    par = np.array([-1, 0, 0, 1, 1, 2], dtype=np.int32)
    path = tmp_path / "lin.txt"
    io.writeLineage(par, path)
    assert path.read_text() == "0\t-1\n1\t0\n2\t0\n3\t1\n4\t1\n5\t2\n"
    assert np.array_equal(io.readLineage(path), par)
    path.write_text("2\t0\n0\t-1\n1\t0\n")          # any line order
    assert list(io.readLineage(path)) == [-1, 0, 0]
    path.write_text("0\t-1\n5\t0\n")                # ids must be 0 .. lines-1
    from bfm import BFMError
    with pytest.raises(BFMError):
        io.readLineage(path)


def _config(**over):
    c = {"R": 0.8, "Rcell": 0.5, "L0": 0.4, "E0": 3900.0, "Ecell": 3.9e6, "sigma0": 100.0,
         "eta_ambient": 0.7, "eta_surface": 720.0, "max_stepsize": 1e-4, "min_stepsize": 1e-9,
         "dt_write": 0.01, "iter_update_stepsize": 1, "iter_update_neighbors": 10,
         "max_iter": 0, "n_cells": 16, "max_time": 0.0, "growth_mean": 1.5, "growth_std": 0.15,
         "daughter_length_std": 0.125, "daughter_angle_xy_bound": 0.05,
         "daughter_angle_z_bound": 0.0, "nz_threshold": 1e-8, "max_rxy_noise": 0.0,
         "max_rz_noise": 0.0, "max_nxy_noise": 0.0, "max_nz_noise": 0.0,
         "max_error_allowed": 1e-8, "basal_only": 0, "adhesion_mode": 1,
         "cell_cell_friction_mode": 0, "adhesion_eqdist": 1.2}
    c.update(over)
    return {k: v for k, v in c.items() if v is not None}


def _write_json(path, obj):
    # json.dumps writes 3900.0 as 3900.0 and 16 as 16: the kinds survive
    path.write_text(json.dumps(obj, indent=2))
    return path


def test_load_config_keys_and_defaults(tmp_path):
    from bfm import io
    cfg = io.load_config(_write_json(tmp_path / "c.json", _config())).as_dict()
    assert cfg["R"] == 0.8 and cfg["Ldiv"] == 2 * 0.4 + 2 * 0.8
    assert cfg["neighbor_threshold"] == 0.5 * (2 * 0.8 + 2.4)
    assert cfg["n_cells"] == 16 and cfg["adhesion_mode"] == 1 and cfg["adhesion_eqdist"] == 1.2
    # optional keys keep the driver's defaults (abmConstAdhesion.cpp:99-131, 249-297)
    assert cfg["jkr_imag_tol"] == 1e-8 and cfg["jkr_aberth_tol"] == 1e-8
    assert cfg["n_mesh_overlap"] == 100 and cfg["n_mesh_gamma"] == 100
    assert cfg["integration_mode"] == 1 and cfg["no_surface"] == 0
    assert cfg["cell_surface_coulomb_coeff"] == 1.0 and cfg["init_filename"] == ""
    cfg = io.load_config(_write_json(tmp_path / "c2.json", _config(
        integration_mode=4, adhesion_n_mesh_overlap=50, no_surface=1,
        init_filename="pop.txt", lineage_filename="lin.txt"))).as_dict()
    assert cfg["integration_mode"] == 4 and cfg["n_mesh_overlap"] == 50 and cfg["no_surface"] == 1
    assert cfg["init_filename"] == "pop.txt" and cfg["lineage_filename"] == "lin.txt"
    # an optional key of the wrong kind is ignored, as the caught conversion there
    cfg = io.load_config(_write_json(tmp_path / "c3.json", _config(
        adhesion_n_mesh_overlap=50.0))).as_dict()
    assert cfg["n_mesh_overlap"] == 100


@pytest.mark.parametrize("over,needle", [
    ({"R": 1}, "R"),                              # integer literal for a double key
    ({"n_cells": 16.0}, "n_cells"),               # double for an integer key
    ({"E0": None}, "E0"),                         # required key missing
    ({"adhesion_mode": 3}, "adhesion mode"),
    ({"cell_cell_friction_mode": 2}, "friction mode"),
    ({"adhesion_eqdist": None}, "adhesion_eqdist"),   # required with adhesion on
    ({"basal_only": 1}, "basal_min_overlap"),
    ({"integration_mode": 7}, "integration mode"),
])
def test_load_config_errors(tmp_path, over, needle):
    from bfm import io, BFMError
    with pytest.raises(BFMError, match=needle):
        io.load_config(_write_json(tmp_path / "e.json", _config(**over)))


def test_load_config_json_syntax(tmp_path):
    from bfm import io, BFMError
    p = tmp_path / "s.json"
    text = json.dumps(_config())
    p.write_text(text[:-1] + ', "note": "caf\\u00e9 \\ud83d\\ude00", "x": [1, 2.5e3, true, null]}')
    assert io.load_config(p).R == 0.8
    for bad in (text + " x", text[:-1], text.replace(":", "=", 1), "[" + text + "]"):
        p.write_text(bad)
        with pytest.raises(BFMError):
            io.load_config(p)
    # a repeated key takes the last value
    p.write_text(text[:-1] + ', "R": 0.9}')
    assert io.load_config(p).R == 0.9


def test_run_rejects_unsupported_modes_before_the_device(tmp_path):
    from bfm import io, BFMError
    from bfm.core import BFM_ERR_UNSUPPORTED
    for over in ({"adhesion_mode": 2, "adhesion_precompute_jkr_forces": 1}, {"max_rxy_noise": 0.1},
                 {"basal_only": 1, "basal_min_overlap": 0.0}):
        p = _write_json(tmp_path / "u.json", _config(**over))
        with pytest.raises(BFMError) as e:
            io.run(p, tmp_path / "out", 42)
        assert e.value.status == BFM_ERR_UNSUPPORTED


def test_anisotropic_config_keys(tmp_path):
    """adhesion_mode 2 (abmConstAdhesion.cpp:133-249): the precompute flag is
    required, the table file names when it is 0; the ellipsoid-pipeline keys
    are optional with the driver's defaults."""
    from bfm import io, BFMError
    from bfm.core import BFM_ERR_ARG
    p = _write_json(tmp_path / "a.json", _config(adhesion_mode=2))
    with pytest.raises(BFMError) as e:
        io.load_config(p)
    assert e.value.status == BFM_ERR_ARG and "adhesion_precompute_jkr_forces" in str(e.value)
    p = _write_json(tmp_path / "a.json", _config(adhesion_mode=2, adhesion_precompute_jkr_forces=0))
    with pytest.raises(BFMError):
        io.load_config(p)                  # the file names are then required
    p = _write_json(tmp_path / "a.json", _config(
        adhesion_mode=2, adhesion_precompute_jkr_forces=0, adhesion_curvature_filename="c.txt",
        adhesion_jkr_forces_filename="f.txt", adhesion_n_mesh_theta=40, adhesion_brent_tol=1e-9,
        adhesion_n_tries_bracket=2.0))
    c = io.load_config(p)
    assert (c.adhesion_mode, c.precompute_jkr_forces) == (2, 0)
    assert c.curvature_filename == b"c.txt" and c.jkr_forces_filename == b"f.txt"
    assert c.n_mesh_theta == 40 and c.n_mesh_half_l == 100 and c.brent_tol == 1e-9
    assert c.n_tries_bracket == 5           # a double literal for an integer key keeps the default
    assert c.min_aspect_ratio == 0.01 and c.calibrate_endpoint_radii == 1.0

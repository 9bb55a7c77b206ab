"""The C-ABI library loads and exports every entry point include/bfm.h declares
(no GPU calls), and the workload generator honours its placement rules."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(bfm_\w+)\s*\(", text, re.M)))


def test_library_exports_header_symbols():
    import bfm
    lib = ctypes.CDLL(bfm.library_path())
    names = _declared("bfm.h")
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name


def test_python_binding_covers_header():
    import bfm.core as core
    lib = core.load_library()
    for name in _declared("bfm.h"):
        assert getattr(lib, name).restype is not None


def test_dome_generator_rules():
    import bfm
    from oracle import oracle as orc
    n = 3000
    cells, rd = bfm.make_dome(n, seed=4)
    assert cells.shape == (n, 22) and cells.flags.f_contiguous
    r = cells[:, 1:4]
    assert np.all(np.linalg.norm(r, axis=1) <= rd + 1e-9) and np.all(r[:, 2] >= 0)
    nrm = cells[:, 4:7]
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1) and np.all(nrm[:, 2] >= 0)
    hl = cells[:, 14]
    assert np.all((hl >= 0.4) & (hl <= 1.6)) and np.allclose(cells[:, 13], 2 * hl)
    assert np.all(r[:, 2] - hl * nrm[:, 2] >= 0.5)          # core does not pierce z = 0
    nb = orc.get_cell_neighbors(cells, 2.0, 0.8, 2.4, binned=True)
    assert np.all(np.linalg.norm(nb[:, 2:5], axis=1) >= 1.0 - 1e-12)   # cores disjoint
    again, _ = bfm.make_dome(n, seed=4)
    assert np.array_equal(cells, again)


def test_struct_layout_matches_ctypes():
    """sizeof / offsetof of bfm_params, bfm_sim_config, bfm_sim_stats and bfm2d_params as the
    C++ compiler lays them out (tests/abi_cpp/abi_check --layout) against the
    ctypes mirrors the Python binding passes across the ABI."""
    import ctypes as C
    import json
    import subprocess

    from bfm.core import Params
    from bfm.model2d import Params2D
    from bfm.sim import SimConfig, SimStats
    exe = os.path.join(ROOT, "tests", "abi_cpp", "abi_check")
    out = json.loads(subprocess.run([exe, "--layout"], capture_output=True, text=True, check=True).stdout)
    for S, name in ((Params, "bfm_params"), (SimConfig, "bfm_sim_config"), (SimStats, "bfm_sim_stats"),
                    (Params2D, "bfm2d_params")):
        assert out[f"sizeof({name})"] == C.sizeof(S), name
        for field, _ in S._fields_:
            assert out[f"{name}.{field}"] == getattr(S, field).offset, (name, field)


def test_phenotypes_per_cell():
    """The batched-ensemble workload: two phenotypes drawn 50/50 per cell in
    every replicate, group 1 -> gamma 500, group 2 -> gamma 50."""
    import bfm
    sizes = bfm.ensemble_sizes(8, 40, 400, seed=0)
    reps, groups = bfm.make_ensemble(sizes, seeds=range(8))
    for cells, g in zip(reps, groups):
        assert set(np.unique(g)) <= {1, 2}
        assert abs(int((g == 1).sum()) - int((g == 2).sum())) <= 1
        assert np.array_equal(cells[:, 20], g)
        assert np.array_equal(cells[:, 21], np.where(g == 1, 500.0, 50.0))
    again, g2 = bfm.make_ensemble(sizes, seeds=range(8))
    assert all(np.array_equal(a, b) for a, b in zip(reps, again))

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

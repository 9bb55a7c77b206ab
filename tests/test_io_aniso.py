"""Reading the anisotropic JKR table files (parseCurvatureRadiiTable,
parseReducedJKRForceTable, parseJKRForceTable; ABM-3d/ABM-3d-cpp/include/
simulation.hpp:565-1008) into the dense tables of bfm_set_aniso_jkr_tables,
on literal files whose tables are written out by hand: the grid of each
parameter from the first block, the index state machine over the rows,
combinations the file lacks reading 0, ragged blocks past the grids dropped,
the Rx index keyed by the field's text. Host code only (no GPU)."""
import numpy as np
import pytest

CURV = """0\t0.5\t0\t0.8\t0.8
0\t0.5\t0.5\t0.9\t0.8
0\t1\t0\t1.1\t0.8
0\t1\t0.5\t1.2\t0.8
0.7\t0.5\t0\t1.3\t0.8
0.7\t0.5\t0.5\t1.4\t0.8
0.7\t1\t0\t1.5\t0.8
0.7\t1\t0.5\t1.6\t0.8
"""

CURV_RAGGED = """0\t0.5\t0\t0.8\t0.8
0\t0.5\t0.5\t0.9\t0.8
0\t1\t0\t1.1\t0.8
0.7\t0.5\t0\t1.3\t0.8
0.7\t0.5\t0.5\t1.4\t0.8
0.7\t0.5\t0.9\t9.9\t0.8
0.7\t1\t0\t1.5\t0.8
0.7\t2\t0\t7.7\t0.8
"""

FORCES_REDUCED = """100
1\t1\t0\t0\t10\t0.1
1\t1\t0\t0.3\t11\t0.2
1\t1\t0.5\t0\t12\t0.3
1\t1\t0.5\t0.3\t13\t0.4
1\t2\t0\t0\t20\t0.5
1\t2\t0\t0.3\t21\t0.6
1\t2\t0.5\t0\t22\t0.7
1\t2\t0.5\t0.3\t23\t0.8
2\t2\t0\t0\t30\t0.9
2\t2\t0\t0.3\t31\t1.0
2\t2\t0.5\t0\t32\t1.1
2\t2\t0.5\t0.3\t33\t1.2
"""

FORCES_VARIABLE = """1\t1\t0\t0\t50\t10\t0.1
1\t1\t0\t0\t100\t11\t0.2
1\t1\t0\t0.3\t50\t12\t0.3
1\t1\t0\t0.3\t100\t13\t0.4
"""


def _read(tmp_path, curv, forces, variable=False):
    from bfm import io
    c = tmp_path / "curv.txt"
    f = tmp_path / "forces.txt"
    c.write_text(curv)
    f.write_text(forces)
    return io.readAnisotropicJKRTables(c, f, variable_gamma=variable)


def test_curvature_and_reduced_force_tables(tmp_path):
    t = _read(tmp_path, CURV, FORCES_REDUCED)
    assert t.theta.tolist() == [0.0, 0.7] and t.half_l.tolist() == [0.5, 1.0]
    assert t.coords.tolist() == [0.0, 0.5]
    np.testing.assert_array_equal(t.curvature_rx, [[[0.8, 0.9], [1.1, 1.2]], [[1.3, 1.4], [1.5, 1.6]]])
    assert t.fixed_gamma == 100.0 and t.gamma is None
    assert t.rx.tolist() == [1.0, 2.0] and t.phi.tolist() == [0.0, 0.5] and t.delta.tolist() == [0.0, 0.3]
    assert t.force_radius.shape == (2, 2, 2, 2, 2)
    np.testing.assert_array_equal(t.force_radius[0, 0, :, :, 0], [[10, 11], [12, 13]])
    np.testing.assert_array_equal(t.force_radius[0, 1, :, :, 0], [[20, 21], [22, 23]])
    np.testing.assert_array_equal(t.force_radius[1, 1, :, :, 0], [[30, 31], [32, 33]])
    np.testing.assert_array_equal(t.force_radius[1, 1, :, :, 1], [[0.9, 1.0], [1.1, 1.2]])
    # Rx1 > Rx2 rows are not in the file: the map lookup gives 0
    assert not t.force_radius[1, 0].any()


def test_ragged_curvature_blocks(tmp_path):
    t = _read(tmp_path, CURV_RAGGED, FORCES_REDUCED)
    # grids from the first theta / (theta, half_l) block; rows whose index
    # lies past them cannot be looked up and are dropped; missing rows read 0
    assert t.theta.tolist() == [0.0, 0.7] and t.half_l.tolist() == [0.5, 1.0]
    assert t.coords.tolist() == [0.0, 0.5]
    np.testing.assert_array_equal(t.curvature_rx, [[[0.8, 0.9], [1.1, 0.0]], [[1.3, 1.4], [1.5, 0.0]]])


def test_rx_index_follows_the_field_text(tmp_path):
    forces = "50\n1\t1\t0\t0\t10\t0.1\n1\t2.0\t0\t0\t20\t0.2\n2\t2\t0\t0\t30\t0.3\n"
    t = _read(tmp_path, CURV, forces)
    # "2.0" never appears in the first column: its index is 0, so the row
    # overwrites (Rx1, Rx2) = (0, 0); the new Rx2 value resets phi and delta
    assert t.rx.tolist() == [1.0, 2.0]
    assert t.force_radius[0, 0, 0, 0].tolist() == [20.0, 0.2]
    assert not t.force_radius[0, 1].any()
    assert t.force_radius[1, 1, 0, 0].tolist() == [30.0, 0.3]


def test_variable_gamma_table(tmp_path):
    t = _read(tmp_path, CURV, FORCES_VARIABLE, variable=True)
    assert t.rx.tolist() == [1.0] and t.phi.tolist() == [0.0]
    assert t.delta.tolist() == [0.0, 0.3] and t.gamma.tolist() == [50.0, 100.0]
    assert t.force_radius.shape == (1, 1, 1, 2, 2, 2)
    np.testing.assert_array_equal(t.force_radius[0, 0, 0], [[[10, 0.1], [11, 0.2]], [[12, 0.3], [13, 0.4]]])


@pytest.mark.parametrize("curv,forces", [
    (CURV.replace("0.9\t0.8", "0.9"), FORCES_REDUCED),          # a missing field
    (CURV, FORCES_REDUCED.replace("21\t0.6", "x\t0.6")),        # not a number
    (CURV, ""),                                                  # empty force file
    ("", FORCES_REDUCED),
])
def test_malformed_files_are_errors(tmp_path, curv, forces):
    from bfm import BFMError
    with pytest.raises(BFMError):
        _read(tmp_path, curv, forces)


@pytest.mark.parametrize("gammas", [None, [0.0, 50.0, 120.0]])
def test_written_files_read_back(tmp_path, gammas):
    """The files in the precompute programs' format (10 significant digits)
    read back to the tables rounded to those digits."""
    from bfm import io
    from bfm.jkr import synthetic_aniso_tables, write_aniso_table_files
    t = synthetic_aniso_tables(n_theta=7, n_half_l=5, n_coords=4, n_rx=6, n_phi=3, n_delta=5,
                               gammas=gammas)
    c, f = tmp_path / "c.txt", tmp_path / "f.txt"
    write_aniso_table_files(t, c, f)
    b = io.readAnisotropicJKRTables(c, f, variable_gamma=gammas is not None)
    r10 = np.vectorize(lambda x: float("%.10g" % x))
    for name in ("theta", "half_l", "coords", "rx", "phi", "delta", "curvature_rx", "force_radius"):
        np.testing.assert_array_equal(getattr(b, name), r10(getattr(t, name)), err_msg=name)
    if gammas is None:
        assert b.fixed_gamma == 100.0
    else:
        np.testing.assert_array_equal(b.gamma, gammas)

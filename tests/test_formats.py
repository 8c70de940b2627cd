"""XTC codec and topology reader (host C++ in libmdsctk_knn.so, no GPU)."""
import ctypes
import os

import numpy as np
import pytest

from mdsctk_amd import knn, synth

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
TRP_XTC = os.path.join(GOLD, "trp-cage.xtc")
TRP_OOS = os.path.join(GOLD, "trp-cage-outofsample.xtc")
TRP_PDB = os.path.join(GOLD, "trp-cage.pdb")
# the reference's backbone GRO of the same system (examples/mld/traj/proteins)
TRP_GRO = os.path.join(GOLD, "trp-cage.bb.gro")


def xtc_info(path):
    L = knn.load()
    n, na, pr = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_float()
    rc = L.mdsctk_xtc_info(path.encode(), ctypes.byref(n), ctypes.byref(na), ctypes.byref(pr))
    return rc, n.value, na.value, pr.value


def xtc_read(path, natoms=None, nthreads=4):
    L = knn.load()
    rc, n, na, _ = xtc_info(path)
    assert rc == 0, L.mdsctk_io_last_error()
    natoms = na if natoms is None else natoms
    x = np.zeros((n, natoms, 3), np.float32)
    steps = np.zeros(n, np.int32)
    times = np.zeros(n, np.float32)
    boxes = np.zeros((n, 9), np.float32)
    got = ctypes.c_int64()
    rc = L.mdsctk_xtc_read(path.encode(), x.ctypes.data, n, natoms, ctypes.byref(got),
                           steps.ctypes.data, times.ctypes.data, boxes.ctypes.data, nthreads)
    return rc, x[: got.value], steps, times, boxes


def xtc_write(path, x, precision, steps=None, times=None, boxes=None):
    L = knn.load()
    x = np.ascontiguousarray(x, np.float32)
    p = lambda a, t: None if a is None else np.ascontiguousarray(a, t).ctypes.data
    return L.mdsctk_xtc_write(path.encode(), x.ctypes.data, x.shape[0], x.shape[1], precision,
                              p(steps, np.int32), p(times, np.float32), p(boxes, np.float32))


def topology(path):
    L = knn.load()
    na = ctypes.c_int32()
    rc = L.mdsctk_topology_read(path.encode(), ctypes.byref(na), None, 0)
    if rc < 0:
        return rc, None
    m = np.zeros(na.value, np.float32)
    rc = L.mdsctk_topology_read(path.encode(), ctypes.byref(na), m.ctypes.data, na.value)
    return rc, m


def test_trp_cage_header_facts():
    rc, n, na, pr = xtc_info(TRP_XTC)
    assert (rc, n, na, pr) == (0, 1000, 60, 1000.0)
    rc, x, steps, times, boxes = xtc_read(TRP_XTC)
    assert rc == 0 and x.shape == (1000, 60, 3)
    assert steps[0] == 100000 and abs(times[0] - 200.0) < 1e-3
    assert (np.diff(steps) > 0).all()
    rc, n, na, pr = xtc_info(TRP_OOS)
    assert (rc, n, na) == (0, 10000, 60)


@pytest.mark.parametrize("path", [TRP_XTC, TRP_OOS])
def test_backbone_geometry_invariants(path):
    """Every decoded frame is a physical N-CA-C backbone: mean bond lengths
    N-CA 0.147, CA-C 0.154, C-N 0.134 nm, each within +-0.02 nm."""
    rc, x, *_ = xtc_read(path)
    assert rc == 0
    d = np.linalg.norm(np.diff(x.astype(np.float64), axis=1), axis=2)
    for j, want in enumerate([0.147, 0.154, 0.134]):
        dd = d[:, j::3]
        assert abs(dd.mean() - want) < 0.003
        assert np.abs(dd - want).max() < 0.02
    # coordinates are multiples of 1/precision
    q = x * 1000.0
    assert np.abs(q - np.round(q)).max() < 2e-3


def test_parallel_decode_matches_serial():
    _, a, *_ = xtc_read(TRP_OOS, nthreads=1)
    _, b, *_ = xtc_read(TRP_OOS, nthreads=8)
    assert (a == b).all()


@pytest.mark.parametrize("precision", [1000.0, 1e5])
def test_round_trip(tmp_path, precision):
    x = synth.trajectory(25, 77, 3)
    f = str(tmp_path / "t.xtc")
    steps = np.arange(25, dtype=np.int32) * 10
    times = np.arange(25, dtype=np.float32) * 2.0
    boxes = np.tile(np.eye(3, dtype=np.float32).reshape(9) * 7.0, (25, 1))
    assert xtc_write(f, x, precision, steps, times, boxes) == 0
    rc, y, s2, t2, b2 = xtc_read(f)
    assert rc == 0
    want = np.round(x.astype(np.float64) * precision) / precision
    assert np.abs(y - want).max() <= 1.5 / precision * 1e-3 + 1e-6
    assert np.abs(y - x).max() <= 0.5 / precision + 1e-6
    assert (s2 == steps).all() and (t2 == times).all() and (b2 == boxes).all()


def test_round_trip_large_range_and_few_atoms(tmp_path):
    # a 300 nm span at precision 1e5 needs the per-axis ("large") coding
    x = synth.trajectory(4, 30, 5) * 60.0
    f = str(tmp_path / "big.xtc")
    assert xtc_write(f, x, 1e5) == 0
    rc, y, *_ = xtc_read(f)
    # quantisation (0.5/precision) plus float32 rounding at |x| ~ 300 nm
    assert rc == 0 and (np.abs(y - x) <= 0.5e-5 + 4 * np.spacing(np.abs(x))).all()
    # <= 9 atoms are stored as raw floats
    s = synth.trajectory(3, 7, 6)
    f2 = str(tmp_path / "small.xtc")
    assert xtc_write(f2, s, 1000.0) == 0
    rc, y2, *_ = xtc_read(f2)
    assert rc == 0 and (y2 == s).all()


def test_errors(tmp_path):
    rc, *_ = xtc_info(str(tmp_path / "missing.xtc"))
    assert rc == knn.ERR_IO
    data = open(TRP_XTC, "rb").read()
    bad = tmp_path / "trunc.xtc"
    bad.write_bytes(data[: len(data) // 2 + 7])
    rc, *_ = xtc_info(str(bad))
    assert rc == knn.ERR_FORMAT
    rc, *_ = xtc_read(TRP_XTC, natoms=59)
    assert rc == knn.ERR_NATOMS


def test_topology_masses(tmp_path):
    rc, m = topology(TRP_PDB)
    assert rc == 0 and m.shape == (60,)
    # N, CA, C repeating backbone
    assert np.allclose(m[0::3], 14.0067) and np.allclose(m[1::3], 12.011) and np.allclose(m[2::3], 12.011)
    gro = tmp_path / "t.gro"
    gro.write_text("title\n    4\n"
                   "    1ALA      N    1   0.100   0.200   0.300\n"
                   "    1ALA     CA    2   0.200   0.200   0.300\n"
                   "    1ALA     HA    3   0.300   0.200   0.300\n"
                   "    2NA      NA    4   0.400   0.200   0.300\n"
                   "   1.00000   1.00000   1.00000\n")
    rc, m = topology(str(gro))
    assert rc == 0 and np.allclose(m, [14.0067, 12.011, 1.008, 22.98977])
    rc, _ = topology(str(tmp_path / "x.tpr"))
    assert rc == knn.ERR_FORMAT


def test_reference_gro_and_pdb_give_the_same_masses():
    """The reference's example PDB (examples/trp-cage.pdb) and backbone GRO
    (examples/mld/traj/proteins/trp-cage.bb.gro) list the same 60 atoms in
    the same order: the reader gives both the same masses, one per atom of
    the trajectory."""
    rc_p, m_p = topology(TRP_PDB)
    rc_g, m_g = topology(TRP_GRO)
    assert rc_p == 0 and rc_g == 0
    np.testing.assert_array_equal(m_g, m_p)
    rc, n, na, pr = xtc_info(TRP_XTC)
    assert rc == 0 and na == m_g.shape[0]

"""XTC decode on the GPU (mdsctk_xtc_read_device, one lane per frame)
against the host decoder (mdsctk_xtc_read), bit for bit: the trp-cage
fixtures (GROMACS-written, with small-difference runs), synthetic files at
precision 1e3 and 1e5, frames of at most 9 atoms (raw floats), coordinate
ranges beyond 2^24 quantisation steps (per-axis bit widths), and malformed
or mismatched files (error codes, no output)."""
import ctypes

import numpy as np
import pytest

from mdsctk_amd import knn, synth
from tests.devmem import hip
from tests.test_formats import TRP_OOS, TRP_XTC, xtc_read, xtc_write

pytestmark = pytest.mark.gpu


def read_device(path, natoms):
    L = knn.load()
    d = ctypes.c_void_p()
    n = ctypes.c_int64()
    rc = L.mdsctk_xtc_read_device(path.encode(), natoms, 0, ctypes.byref(d), ctypes.byref(n))
    if rc < 0:
        return rc, None
    x = np.zeros((n.value, natoms, 3), np.float32)
    if n.value:
        assert hip().hipMemcpy(x.ctypes.data_as(ctypes.c_void_p), d, ctypes.c_size_t(x.nbytes), 2) == 0
    assert L.mdsctk_xtc_free_device(d) == 0
    return rc, x


def _same(path):
    rc, x_host, *_ = xtc_read(path)
    assert rc == 0
    rc, x_dev = read_device(path, x_host.shape[1])
    assert rc == 0, knn.load().mdsctk_io_last_error()
    assert x_dev.shape == x_host.shape
    assert x_dev.tobytes() == x_host.tobytes()
    return x_dev


def test_trp_cage_fixtures_decode_identically():
    x = _same(TRP_XTC)
    assert x.shape == (1000, 60, 3)
    y = _same(TRP_OOS)
    assert y.shape == (10000, 60, 3)


@pytest.mark.parametrize("natoms,precision,n", [(1000, 1e5, 700), (304, 1e3, 513), (60, 1e3, 3),
                                                (9, 1e3, 40), (1, 1e3, 5), (150, 1e4, 1)])
def test_synthetic_files_decode_identically(tmp_path, natoms, precision, n):
    f = str(tmp_path / "s.xtc")
    assert xtc_write(f, synth.trajectory(n, natoms, natoms + n), precision) == 0
    _same(f)


def test_wide_coordinate_range(tmp_path):
    """Coordinates spanning more than 2^24 quantisation steps per axis take
    the per-axis bit widths instead of the packed triple."""
    x = synth.trajectory(20, 50, 3)
    x[:, ::7, :] *= 400.0  # +-2000 nm at precision 1e4: 4e7 steps
    f = str(tmp_path / "w.xtc")
    assert xtc_write(f, x, 1e4) == 0
    _same(f)


def test_file_spanning_several_upload_chunks(tmp_path):
    """More than one 64 MiB read chunk: headers straddle chunk boundaries."""
    f = str(tmp_path / "big.xtc")
    x = synth.trajectory(12000, 1000, 5)
    assert xtc_write(f, x, 1e5) == 0
    _same(f)


def test_errors(tmp_path):
    L = knn.load()
    d = ctypes.c_void_p()
    n = ctypes.c_int64()
    f = str(tmp_path / "t.xtc")
    assert xtc_write(f, synth.trajectory(10, 64, 1), 1e3) == 0
    assert L.mdsctk_xtc_read_device(f.encode(), 65, 0, ctypes.byref(d), ctypes.byref(n)) == knn.ERR_NATOMS
    raw = open(f, "rb").read()
    g = str(tmp_path / "cut.xtc")
    open(g, "wb").write(raw[:-10])
    assert L.mdsctk_xtc_read_device(g.encode(), 64, 0, ctypes.byref(d), ctypes.byref(n)) == knn.ERR_FORMAT
    assert not d.value
    assert L.mdsctk_xtc_read_device(b"/nonexistent.xtc", 64, 0, ctypes.byref(d), ctypes.byref(n)) == knn.ERR_IO
    na = ctypes.c_int32()
    assert L.mdsctk_xtc_natoms(f.encode(), ctypes.byref(na)) == 0 and na.value == 64

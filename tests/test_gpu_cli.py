"""The drop-in executables end to end on the trp-cage fixtures: knn_rms
writes the reference's distances.dat / indices.dat (checked against the
FP64 oracle on the decoded frames with the topology's masses), and the
example pipeline knn_rms -> make_sysparse -> decomp_sparse reproduces the
CPU pipeline's spectrum."""
import subprocess

import numpy as np
import pytest

import oracle
import oracle.graph as og
from mdsctk_amd import build
from tests.parity import check_knn
from tests.test_formats import TRP_GRO, TRP_OOS, TRP_PDB, TRP_XTC, topology, xtc_read

pytestmark = pytest.mark.gpu


def run(tool, args, cwd):
    return subprocess.run([tool] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True, timeout=300)


def read_rows(cwd, n, k, d="distances.dat", i="indices.dat"):
    dist = np.fromfile(cwd / d, np.float64)
    idx = np.fromfile(cwd / i, np.int32)
    assert dist.size == n * k and idx.size == n * k
    return dist.reshape(n, k), idx.reshape(n, k)


@pytest.fixture(scope="module")
def trp():
    _, x, *_ = xtc_read(TRP_XTC)
    _, y, *_ = xtc_read(TRP_OOS)
    _, w = topology(TRP_PDB)
    return x, y, w


@pytest.mark.parametrize("prec", ["bf16", "fp32", "exact"])
def test_knn_rms_self_query_files(tmp_path, trp, prec):
    x, _, w = trp
    p = run(build.CLI_PATH, ["-k", "12", "-p", TRP_PDB, "-r", TRP_XTC, "--precision", prec],
            tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    d, i = read_rows(tmp_path, 1000, 12)
    check_knn(x, x, w, d, i, *oracle.knn_exact(x, x, w, 12))


def test_knn_rms_self_query_forms_agree(tmp_path, trp):
    """-f omitted, -f equal to -r (both a self query scoring each
    unordered pair once) and -f naming a copy of the file (a cross query
    through the candidate lists) write identical files."""
    import shutil
    copy = tmp_path / "copy.xtc"
    shutil.copy(TRP_XTC, copy)
    outs = []
    for tag, extra in (("a", []), ("b", ["-f", TRP_XTC]), ("c", ["-f", str(copy)])):
        p = run(build.CLI_PATH, ["-k", "15", "-p", TRP_PDB, "-r", TRP_XTC, "-d", f"d{tag}.dat",
                                 "-i", f"i{tag}.dat"] + extra, tmp_path)
        assert p.returncode == 0, p.stdout + p.stderr
        outs.append(((tmp_path / f"d{tag}.dat").read_bytes(), (tmp_path / f"i{tag}.dat").read_bytes()))
    assert outs[0] == outs[1] == outs[2]


def test_knn_rms_gro_topology_writes_the_same_files(tmp_path):
    """-p naming the reference's backbone GRO of the system writes the same
    files as -p naming its PDB (the same atoms and masses)."""
    outs = []
    for tag, top in (("p", TRP_PDB), ("g", TRP_GRO)):
        p = run(build.CLI_PATH, ["-k", "10", "-p", top, "-r", TRP_XTC, "-d", f"d{tag}.dat",
                                 "-i", f"i{tag}.dat"], tmp_path)
        assert p.returncode == 0, p.stdout + p.stderr
        outs.append(((tmp_path / f"d{tag}.dat").read_bytes(), (tmp_path / f"i{tag}.dat").read_bytes()))
    assert outs[0] == outs[1]


def test_knn_rms_out_of_sample_and_nofit(tmp_path, trp):
    x, y, w = trp
    p = run(build.CLI_PATH, ["-k", "8", "-p", TRP_PDB, "-r", TRP_XTC, "-f", TRP_OOS, "-n", "true",
                             "-d", "d.dat", "-i", "i.dat"], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    d, i = read_rows(tmp_path, y.shape[0], 8, "d.dat", "i.dat")
    # rank 0 is dropped for an external -f too (knn_rms.cpp:246-255)
    check_knn(x, y, w, d, i, *oracle.knn_exact(x, y, w, 8, nofit=True), nofit=True)


def test_knn_rms_full_rows(tmp_path, trp):
    x, _, w = trp
    p = run(build.CLI_PATH, ["-s", "false", "-p", TRP_PDB, "-r", TRP_XTC], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    d, i = read_rows(tmp_path, 1000, 1000)
    assert np.abs(d - oracle.full_exact(x, x, w)).max() < 1e-4
    assert (i == np.arange(1000)[None, :]).all()


def test_example_pipeline_spectrum(tmp_path, trp):
    """knn_rms -k 20 | make_sysparse | decomp_sparse (the cluster_rms example
    flow) against the same flow on the CPU: FP64 oracle kNN, the
    edge-dictionary make_sysparse and ARPACK."""
    x, _, w = trp
    assert run(build.CLI_PATH, ["-k", "20", "-p", TRP_PDB, "-r", TRP_XTC], tmp_path).returncode == 0
    assert run(build.MAKE_SYSPARSE_PATH, ["-k", "20"], tmp_path).returncode == 0
    d, i = read_rows(tmp_path, 1000, 20)
    sigma = float(np.median(d))
    p = run(build.DECOMP_SPARSE_PATH, ["-q", repr(sigma), "-n", "8"], tmp_path)
    assert p.returncode == 0, p.stdout
    ev = np.loadtxt(tmp_path / "eigenvalues.dat")
    d_o, i_o = oracle.knn_exact(x, x, w, 20)
    n, pcol, irow, vals = og.ssm_build(d_o, i_o)
    w_o, _ = og.spectral_decomp(n, pcol, irow, vals, float(np.median(d_o)), 8)
    assert np.abs(ev - w_o).max() < 1e-6


def test_knn_rms_full_rows_stream_in_blocks(tmp_path):
    """--sort false with more than 2^27 output values: the rows are produced
    and written in two blocks (host memory stays bounded); sampled rows on
    either side of the block boundary equal the library's full rows, and
    every index row is the identity."""
    from mdsctk_amd import knn, synth
    from tests.test_formats import xtc_write
    n, natoms = 11600, 20  # 1.3456e8 values > 2^27
    x = synth.trajectory(n, natoms, 44)
    f = str(tmp_path / "t.xtc")
    assert xtc_write(f, x, 1e5) == 0
    pdb = tmp_path / "t.pdb"
    with open(pdb, "w") as fh:
        for a in range(natoms):
            fh.write("ATOM  %5d  CA  ALA A%4d    %8.3f%8.3f%8.3f  1.00  0.00           C\n"
                     % (a + 1, a + 1, 0.0, 0.0, 0.0))
    p = run(build.CLI_PATH, ["-s", "false", "-p", str(pdb), "-r", f], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    d = np.memmap(tmp_path / "distances.dat", np.float64, "r", shape=(n, n))
    i = np.memmap(tmp_path / "indices.dat", np.int32, "r", shape=(n, n))
    rows = [0, 1, 11566, 11567, 11568, 11599]
    from tests.test_formats import topology, xtc_read
    _, xd, *_ = xtc_read(f)
    rc, w = topology(str(pdb))
    assert rc >= 0 and w.shape == (natoms,)
    with knn.KnnContext(0, "bf16") as ctx:
        ctx.set_reference(xd, w)
        d_l, _ = ctx.query(xd[rows], 0, knn.KNN_FULL_ROWS)
    np.testing.assert_array_equal(np.asarray(d[rows]), d_l)
    for r in rows:
        assert (np.asarray(i[r]) == np.arange(n)).all()

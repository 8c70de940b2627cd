"""knn_rms command-line contract that needs no GPU: options and exit codes
(knn_rms.cpp:74-120, 158-178)."""
import os
import subprocess

import numpy as np

from mdsctk_amd import build, synth
from tests.test_formats import TRP_PDB, TRP_XTC, xtc_write

CLI = build.CLI_PATH


def run(args, cwd):
    return subprocess.run([CLI] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True, timeout=60)


def test_help_exits_1(tmp_path):
    p = run(["--help"], tmp_path)
    assert p.returncode == 1 and "--knn" in p.stdout and "--topology-file" in p.stdout


def test_missing_knn_exits_255(tmp_path):
    p = run(["-p", TRP_PDB, "-r", TRP_XTC], tmp_path)
    assert p.returncode == 255 and "--knn not supplied" in p.stdout


def test_unknown_option_exits_1(tmp_path):
    assert run(["--bogus", "3"], tmp_path).returncode == 1


def test_atom_count_mismatch_exits_4(tmp_path):
    f = str(tmp_path / "five.xtc")
    assert xtc_write(f, synth.trajectory(3, 12, 1), 1000.0) == 0
    p = run(["-k", "2", "-p", TRP_PDB, "-r", f], tmp_path)
    assert p.returncode == 4 and "does not match" in p.stdout
    p = run(["-k", "2", "-p", TRP_PDB, "-r", TRP_XTC, "-f", f], tmp_path)
    assert p.returncode == 4


def test_option_spellings_are_accepted(tmp_path):
    # --sort false does not need --knn; prefixes and -xVALUE forms parse.
    p = run(["--sort=false", "-n0", "--topo", TRP_PDB, "-r", "missing.xtc"], tmp_path)
    assert "sort =           0" in p.stdout and p.returncode == 3


# knn_data (knn_data.cpp:66-111): the same contract on .pts data vectors


def run_data(args, cwd):
    return subprocess.run([build.KNN_DATA_PATH] + args, cwd=cwd, stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=60)


def test_knn_data_options_and_exit_codes(tmp_path):
    p = run_data(["-h"], tmp_path)
    assert p.returncode == 1 and "--vector-size" in p.stdout and "--correlation" in p.stdout
    p = run_data(["-v", "8"], tmp_path)
    assert p.returncode == 255 and "--knn not supplied" in p.stdout
    p = run_data(["-k", "3"], tmp_path)
    assert p.returncode == 255 and "--vector-size not supplied" in p.stdout
    p = run_data(["-s", "false"], tmp_path)
    assert p.returncode == 255 and "--knn not supplied" not in p.stdout
    assert run_data(["--bogus", "1"], tmp_path).returncode == 1
    p = run_data(["-k", "3", "-v", "8", "-r", "missing.pts"], tmp_path)
    assert p.returncode == 3 and "reference-file = missing.pts" in p.stdout
    p = run_data(["-k", "3", "-v", "8", "-c", "-r", "missing.pts"], tmp_path)
    assert "correlation =    1" in p.stdout and "fit-file =       missing.pts" in p.stdout

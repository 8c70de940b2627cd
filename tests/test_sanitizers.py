"""Host code under sanitizers (no GPU), built from the library's own sources
with g++, once under AddressSanitizer + UndefinedBehaviorSanitizer and once
under ThreadSanitizer.  Any sanitizer report fails the run (halt_on_error /
a non-zero exit).

  host_driver.cpp   the XTC codec (parallel decode, truncated and corrupted
                    files), the topology reader and the graph stages with
                    the threaded Lanczos solver
  serve_driver.cpp  the serving queue (serve_queue.h) with a host-only
                    device: submit / poll / wait from six client threads,
                    injected launch failures, a stop with requests queued"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mdsctk_amd", "csrc")
SOURCES = [os.path.join(ROOT, "tests", "sanitize", "host_driver.cpp")] + [
    os.path.join(CSRC, f) for f in ("xtc.cpp", "topology.cpp", "spectral.cpp")]
GOLD = os.path.join(ROOT, "tests", "golden")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")


SERVE_SOURCES = [os.path.join(ROOT, "tests", "sanitize", "serve_driver.cpp")]


def _build_and_run(tmp_path, flags, env_extra, sources=SOURCES, name="host_driver"):
    exe = str(tmp_path / name)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", f"-I{ROOT}/include"] + \
        flags + sources + ["-o", exe, "-lpthread"]
    b = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert b.returncode == 0, b.stdout
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    env = dict(os.environ, **env_extra)
    args = [exe, GOLD, str(scratch), "quick"] if name == "host_driver" else [exe, "quick"]
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env,
                       timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "0 failures" in r.stdout
    assert "Sanitizer" not in r.stderr, out[-4000:]


def test_address_and_undefined_behaviour(tmp_path):
    _build_and_run(tmp_path, ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
                   {"ASAN_OPTIONS": "halt_on_error=1:detect_leaks=1",
                    "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"})


def test_threads(tmp_path):
    _build_and_run(tmp_path, ["-fsanitize=thread"], {"TSAN_OPTIONS": "halt_on_error=1"})


def test_serving_queue_address_and_undefined_behaviour(tmp_path):
    _build_and_run(tmp_path, ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
                   {"ASAN_OPTIONS": "halt_on_error=1:detect_leaks=1",
                    "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"},
                   SERVE_SOURCES, "serve_driver")


def test_serving_queue_threads(tmp_path):
    _build_and_run(tmp_path, ["-fsanitize=thread"], {"TSAN_OPTIONS": "halt_on_error=1"},
                   SERVE_SOURCES, "serve_driver")

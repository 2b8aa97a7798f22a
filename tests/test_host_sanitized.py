"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

The solver's host-only C++ (the tile planner pbs_tiles.cpp and the mesh utilities pbs_mesh.cpp)
and the C oracle (oracle/pbs_oracle.c) are compiled from their sources with gcc/g++ into two
standalone drivers (tests/cpp/host_check.cpp, tests/cpp/oracle_check.c), once plainly and once with
-fsanitize=address,undefined, and run.  host_check also checks the tile plan itself: every
corner slot unique, the gather visiting each node's contributions in (tet, corner) order (bitwise
the sequential sums), owned + halo partials equal to the direct sums.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import meshgen as M

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "phybs_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no gcc/g++")


def build_host_check(out, sanitize):
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
           os.path.join(HERE, "cpp", "host_check.cpp"), os.path.join(CSRC, "pbs_tiles.cpp"),
           os.path.join(CSRC, "pbs_mesh.cpp"), "-o", out] + (SAN if sanitize else [])
    subprocess.run(cmd, check=True, capture_output=True)


def build_oracle_check(out, sanitize):
    cmd = ["gcc", "-std=c11", "-O1", "-g", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
           "-I" + os.path.join(ROOT, "oracle"), os.path.join(HERE, "cpp", "oracle_check.c"),
           os.path.join(ROOT, "oracle", "pbs_oracle.c"), "-lm", "-o", out] + (SAN if sanitize else [])
    subprocess.run(cmd, check=True, capture_output=True)


@pytest.mark.parametrize("sanitize", [False, True])
def test_host_check(tmp_path, sanitize):
    exe = str(tmp_path / "host_check")
    build_host_check(exe, sanitize)
    r = subprocess.run([exe, os.path.join(HERE, "golden")], capture_output=True, text=True, env=ENV, timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "host_check: all passed" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


@pytest.mark.parametrize("sanitize", [False, True])
def test_oracle_check(tmp_path, sanitize):
    rest, tets = M.bar(3, 2, 5, 0.05)
    mask = M.clamp_x0(rest)
    x0 = M.twisted_state(rest, 3, 2, 5, 0.05, 40.0, 0.01, 2)
    path = tmp_path / "mesh.bin"
    with open(path, "wb") as f:
        np.array([len(rest), len(tets)], np.int32).tofile(f)
        rest.astype(np.float32).tofile(f)
        tets.astype(np.int32).tofile(f)
        mask.astype(np.int32).tofile(f)
        x0.astype(np.float32).tofile(f)
    exe = str(tmp_path / "oracle_check")
    build_oracle_check(exe, sanitize)
    r = subprocess.run([exe, str(path)], capture_output=True, text=True, env=dict(ENV, OMP_NUM_THREADS="2"),
                       timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "oracle_check: all passed" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr

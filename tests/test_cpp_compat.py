"""The header-only C++ layer (include/phybs.hpp) with the reference's class surface.

CPU: the demo that drives CUDAInterface / DescentOptimizeT the way the reference application does
compiles and links against the library; TimeIntegrationT's host-only surface.  GPU: its trajectories equal, bit for bit, those of the
same calls made through the C ABI from Python.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import phybs_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO_DIR = os.path.join(ROOT, "tests", "cpp")


def test_cpp_layer_compiles_and_links(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    P.lib()  # the library must be built
    out = tmp_path / "compat_demo"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(DEMO_DIR, "compat_demo.cpp"), "-L" + os.path.join(ROOT, "phybs_amd"), "-lphybs",
                    "-o", str(out)], check=True)
    assert out.exists()


def test_time_integration_shim_cpu(tmp_path):
    """TimeIntegrationT: the reference constructors' time steps (TimeIntegration.cpp:6-30), explicit
    Euler (TimeIntegration.cpp:57-67) bit for bit against numpy in FP32, and BackEuler refusing a
    CPU-assembled matrix.  No GPU: the demo's timeint mode uses host calls only."""
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    lib = os.path.join(ROOT, "phybs_amd")
    P.lib()
    exe = tmp_path / "compat_demo"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(DEMO_DIR, "compat_demo.cpp"), "-L" + lib, "-lphybs", "-Wl,-rpath," + lib,
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "timeint", "3"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert np.array_equal(np.array(lines[0].split(), np.float32), np.array([1e-6, 2e-3, 0.03], np.float32))
    assert lines[1].strip() == "1"  # BackEuler threw PBS_ERR_STATE
    got = np.array([[float(v) for v in line.split()] for line in lines[2:]], np.float32)
    rest, _ = P.mesh_bar(4, 4, 16, 0.05)
    x = P.mesh_twist(rest, 4, 4, 16, 0.05, 30.0, 0.0, 1).astype(np.float32)
    n = len(rest)
    i = np.arange(n)
    f = np.zeros((n, 3), np.float32)
    f[:, 2] = np.float32(-9.8) * (np.float32(1) + np.float32(0.01) * (i % 7).astype(np.float32))
    m = (np.float32(0.125) + np.float32(0.001) * (i % 5).astype(np.float32)).astype(np.float32)
    t = np.float32(1.0) / np.float32(30.0)
    v = np.zeros_like(x)
    for _ in range(3):
        v = v + t * f / m[:, None]
        x = x + v * t
    assert np.array_equal(got.view(np.uint32), x.view(np.uint32))


def demo_positions(mode, steps):
    exe = os.path.join(DEMO_DIR, "compat_demo")
    if not os.path.exists(exe):
        raise RuntimeError("tests/cpp/compat_demo is not built (run __graft_entry__.build())")
    r = subprocess.run([exe, mode, str(steps)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return np.array([[float(v) for v in line.split()] for line in r.stdout.splitlines()], np.float32)


def bar():
    rest, tets = P.mesh_bar(4, 4, 16, 0.05)
    x0 = P.mesh_twist(rest, 4, 4, 16, 0.05, 30.0, 0.0, 1)
    mask = (rest[:, 0] == 0).astype(np.int32)
    return rest, tets, mask, x0


@pytest.mark.gpu
def test_cuda_interface_shim_matches_c_abi():
    rest, tets, mask, x0 = bar()
    got = demo_positions("cuda", 3)
    s = P.Solver(rest, tets, mask, P.default_params(youngs=1e5, nu=0.45, density=1000.0))
    s.set_positions(x0)
    ext = np.zeros_like(rest)
    for _ in range(3):
        ref = s.do_descent_opt(1.0 / 30.0, 96, ext)
    s.close()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
def test_back_euler_shim_matches_c_abi():
    rest, tets, mask, x0 = bar()
    got = demo_positions("backeuler", 3)
    s = P.Solver(rest, tets, mask, P.default_params(youngs=1e5, nu=0.45, density=1000.0))
    s.set_positions(x0)
    for _ in range(3):
        _, ref = s.implicit_step(want_positions=True)
    s.close()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
def test_descent_optimize_shim_matches_c_abi():
    rest, tets, mask, x0 = bar()
    got = demo_positions("descent", 3)
    s = P.Solver(rest, tets, mask, P.default_params(semantics=P.SEM_REF_CPU, iterations=96))
    s.set_positions(x0)
    v_last = np.zeros_like(rest)
    v = np.zeros_like(rest)
    for _ in range(3):
        s.set_positions(s.positions())
        s.set_velocities(v, v_last)
        v_last = v
        _, ref = s.step(np.zeros_like(rest), want_positions=True)
        v = s.velocities()[0]
    s.close()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))

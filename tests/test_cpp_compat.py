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


def read_arrays(path):
    out = []
    with open(path, "rb") as f:
        data = f.read()
    i = 0
    while i < len(data):
        tag, cnt = np.frombuffer(data, np.int32, 2, i)
        i += 8
        dt = np.float32 if tag == 0 else np.int32
        out.append(np.frombuffer(data, dt, cnt, i).copy())
        i += 4 * cnt
    return out


@pytest.mark.gpu
def test_tetmesh_and_material_shims(tmp_path):
    """TetMeshT (init*FromFile, rest data) and IsotropicNeohookeanMaterialT (the material calls of
    IsotropicNeohookeanMaterial.h:19-20 and IsotropicMaterial.h:30-57) on the reference's TetGen
    bar, driven from C++: the mesh equals the C ABI's loader bitwise and its rest data the
    oracle's bitwise; the forces, energies and global stiffness matrix equal the C ABI's calls
    bitwise; the per-tet stiffness matrices (all tets, through the C ABI) assemble to the oracle's
    global K; the principal-stretch derivatives match finite differences of the energy."""
    from oracle import pyoracle as O

    prefix = os.path.join(ROOT, "tests", "golden", "model_bar", "bar.1")
    out = str(tmp_path / "material.bin")
    exe = os.path.join(DEMO_DIR, "compat_demo")
    r = subprocess.run([exe, "material", prefix, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    (rest, tets, faces, mass, inv_mass, vol, dminv, grav, pos, f, energy, rowptr, col, kval, diag, kid, k0, k1,
     grad, hess, cons) = read_arrays(out)
    rest, tets, pos, f = rest.reshape(-1, 3), tets.reshape(-1, 4), pos.reshape(-1, 3), f.reshape(-1, 3)
    nodes_ref, tets_ref, faces_ref = P.tetgen_load(prefix, 0.1, faces=True)
    assert np.array_equal(rest.view(np.uint32), nodes_ref.view(np.uint32))
    assert np.array_equal(tets, tets_ref) and np.array_equal(faces.reshape(-1, 3), faces_ref)
    nn, nt = len(rest), len(tets)
    mask = np.zeros(nn, np.int32)
    mask[cons] = 1
    assert np.array_equal(cons, np.nonzero(rest[:, 0] == rest[:, 0].min())[0])
    # rest data: bit-exact against the oracle (TetMesh.cpp:189-216)
    pre = O.precompute(rest, tets, mask, 1000.0, (0.0, -9.8, 0.0), 1e8)
    assert np.array_equal(mass.view(np.uint32), pre["mass"].reshape(-1).view(np.uint32))
    assert np.array_equal(vol.view(np.uint32), pre["vol"].reshape(-1).view(np.uint32))
    dm = dminv.reshape(nt, 3, 3).transpose(0, 2, 1)  # 3 x 3T column-major -> per tet row-major
    assert np.array_equal(dm.reshape(-1).view(np.uint32), pre["dminv"].reshape(-1).view(np.uint32))
    assert np.array_equal(grav.view(np.uint32), pre["gforce"].reshape(-1).view(np.uint32))
    assert np.all(inv_mass[mask == 1] == 0) and np.allclose(inv_mass[mask == 0], 1.0 / mass[mask == 0], rtol=1e-7)
    # forces / energies: the C ABI's calls (REF_CPU clamp) plus gravity, external force, springs
    s = P.Solver(rest, tets, mask, P.default_params(semantics=P.SEM_REF_CPU))
    _, nf, te = s.compute_forces(pos)
    ext = np.zeros_like(rest)
    ext[-1, 1] = np.float32(0.25)
    want = (pre["gforce"].reshape(-1, 3).astype(np.float32) + ext) + nf
    want[mask == 1] = np.float32(1e8) * (rest[mask == 1] - pos[mask == 1])
    assert np.array_equal(f.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(energy.view(np.uint32), te.view(np.uint32))
    # global K at pos: the C ABI's values in FP32, the reference's CSR pattern
    s.set_positions(pos)
    r2, c2, v2 = s.stiffness_matrix(tets)
    assert np.array_equal(rowptr, r2) and np.array_equal(col, c2)
    assert np.array_equal(kval.view(np.uint32), v2.astype(np.float32).view(np.uint32))
    o_rp, o_col, o_diag, o_kid = O.csr_pattern(nn, tets)
    assert np.array_equal(diag, o_diag) and np.array_equal(kid, o_kid)
    # per-tet K: the two the demo printed equal the C ABI's; all tets assembled = the oracle's K
    Kt = s.tet_stiffness(pos, np.arange(nt))
    assert np.array_equal(k0.reshape(12, 12).T.view(np.uint32), Kt[0].view(np.uint32))
    assert np.array_equal(k1.reshape(12, 12).T.view(np.uint32), Kt[-1].view(np.uint32))
    assert np.abs(Kt - Kt.transpose(0, 2, 1)).max() <= 1e-5 * np.abs(Kt).max()  # symmetric
    rs = Kt.reshape(nt, 12, 4, 3).sum(axis=2)  # translation null space: rows sum to zero per axis
    assert np.abs(rs).max() <= 1e-4 * np.abs(Kt).max()
    mu, lam = O.lame(1e5, 0.45)
    _, _, kref, _, _ = O.implicit_system(rest, tets, mask, pre, mu, lam, pos.astype(np.float64),
                                         np.zeros((nn, 3)), 0.03)
    asm = np.zeros(len(col))
    for t in range(nt):  # scatter through the reference's getKIDinCSRval (IsotropicMaterial.cpp:619-643)
        for a in range(4):
            for d in range(3):
                for b in range(4):
                    base = kid[48 * t + 4 * (3 * a + d) + b]
                    asm[base:base + 3] += Kt[t, 3 * a + d, 3 * b:3 * b + 3]
    err = np.abs(asm - kref).max() / np.abs(kref).max()
    assert err <= 1e-5, err
    s.close()
    # principal-stretch derivatives against central differences of the energy (FP64)
    fh = np.array([1.1, 0.95, 0.9], np.float64)

    def psi(sv):
        J = np.prod(sv)
        return 0.5 * mu * (np.sum(sv * sv) - 3) - mu * np.log(J) + 0.5 * lam * np.log(J) ** 2

    eps = 1e-6
    g_fd = np.array([(psi(fh + eps * e) - psi(fh - eps * e)) / (2 * eps) for e in np.eye(3)])
    h_fd = np.array([[(psi(fh + eps * (ei + ej)) - psi(fh + eps * (ei - ej)) - psi(fh - eps * (ei - ej))
                       + psi(fh - eps * (ei + ej))) / (4 * eps * eps) for ej in np.eye(3)] for ei in np.eye(3)])
    assert np.allclose(grad, g_fd, rtol=1e-5, atol=1e-3 * mu)
    assert np.allclose(hess.reshape(3, 3), h_fd, rtol=1e-4, atol=1e-3 * mu)


@pytest.mark.gpu
def test_descent_optimize_uses_its_material():
    """DescentOptimizeT takes the energy and constants from the material object it is handed:
    a stretch material with E = 2e5, nu = 0.4 gives bitwise the C ABI's stretch-material run, and
    not the neo-Hookean one."""
    rest, tets, mask, x0 = bar()
    got = demo_positions("descent_stretch", 3)
    neo = demo_positions("descent", 3)
    s = P.Solver(rest, tets, mask, P.default_params(semantics=P.SEM_REF_CPU, iterations=96, material=P.MAT_STRETCH,
                                                   youngs=2e5, nu=0.4))
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
    assert not np.array_equal(got, neo)

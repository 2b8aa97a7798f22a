"""GPU tests of the backward-Euler step (the reference's GPU implicit path, fvm.cpp:179-183):
the stiffness matrix against the oracle's (which follows the reference's SVD / Xu et al. route,
the GPU evaluates the closed form), and steps against the oracle's linear system solved by a
sparse direct solver (scipy; the reference factorises with a sparse Cholesky)."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import meshgen as M
import phybs_amd as P
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

H = float(np.float32(0.03))


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def case(nx=3, ny=3, nz=10, twist=30.0, seed=1):
    rest, tets = M.bar(nx, ny, nz, 0.05)
    mask = M.clamp_x0(rest)
    x0 = M.twisted_state(rest, nx, ny, nz, 0.05, twist, 0.01, seed)
    v0 = (0.05 * np.random.default_rng(seed).standard_normal(rest.shape)).astype(np.float32)
    v0[mask != 0] = 0
    return rest, tets, mask, x0, v0


def oracle_solve(rest, tets, mask, x, v, material=P.MAT_NEOHOOKEAN, ext=None):
    mu, lam = O.lame(1e5, 0.45)
    pre = O.precompute(rest, tets, mask)
    rowptr, col, _, aval, b = O.implicit_system(rest, tets, mask, pre, mu, lam, x, v, H, ext=ext, material=material)
    n3 = 3 * len(rest)
    A = sp.csr_matrix((aval, col, rowptr), shape=(n3, n3))
    vn = spl.spsolve(A.tocsc(), b).reshape(-1, 3)
    return x.reshape(-1, 3) + H * vn, vn


@pytest.mark.parametrize("precision", [P.FP32, P.FP64])
@pytest.mark.parametrize("material", [P.MAT_NEOHOOKEAN, P.MAT_STRETCH])
def test_stiffness_matrix_matches_oracle(precision, material):
    rest, tets, mask, x0, _ = case()
    say("stiffness", precision, material)
    s = P.Solver(rest, tets, mask, P.default_params(precision=precision, material=material))
    s.set_positions(x0)
    rowptr, col, val = s.stiffness_matrix(tets)
    mu, lam = O.lame(1e5, 0.45)
    pre = O.precompute(rest, tets, mask)
    r2, c2, kval, _, _ = O.implicit_system(rest, tets, mask, pre, mu, lam, x0.astype(np.float64),
                                           np.zeros_like(x0, np.float64), H, material=material)
    assert np.array_equal(rowptr, r2) and np.array_equal(col, c2)
    err = np.abs(val - kval).max() / np.abs(kval).max()
    say("max |K_gpu - K_oracle| / max |K|", err)
    # FP64 arithmetic on FP32 mesh data: the oracle takes the reference's FP32 area vectors AN,
    # the GPU forms -vol Dm^-T from FP32 vol and Dm^-1; the two differ by FP32 rounding
    assert err <= (1e-5 if precision == P.FP32 else 1e-6)
    s.close()


@pytest.mark.parametrize("material", [P.MAT_NEOHOOKEAN, P.MAT_STRETCH])
def test_stiffness_matrix_wide_rows_fp64(material):
    """The reference's TetGen bar (tests/golden/model_bar/bar.1) has nodes in up to 28 block
    columns; in FP64 its widest 64-row slice needs 28 * 9 * 64 * 8 B = 129 KB, beyond the 96 KB of
    LDS the assembly accumulates rows in, so the matrix is accumulated in place (the other
    assembly path).  Both materials against the oracle, with a mild random deformation."""
    import os

    nodes, tets = P.tetgen_load(os.path.join(M.GOLDEN, "model_bar/bar.1"), 0.1)
    mask = (nodes[:, 0] == nodes[:, 0].min()).astype(np.int32)
    rng = np.random.default_rng(3)
    x = (nodes + rng.uniform(-1e-3, 1e-3, nodes.shape)).astype(np.float32)
    rowptr, col, _, _ = P.csr_pattern(len(nodes), tets)
    width = np.diff(rowptr)[::3] // 3  # block columns per node (row 3a holds 3 per block)
    say("bar.1 FP64 stiffness, material", material, "widest row", int(width.max()), "block columns")
    assert width.max() * 9 * 64 * 8 > 96 * 1024
    s = P.Solver(nodes, tets, mask, P.default_params(precision=P.FP64, material=material))
    s.set_positions(x)
    rowptr, col, val = s.stiffness_matrix(tets)
    s.close()
    mu, lam = O.lame(1e5, 0.45)
    pre = O.precompute(nodes, tets, mask)
    r2, c2, kval, _, _ = O.implicit_system(nodes, tets, mask, pre, mu, lam, x.astype(np.float64),
                                           np.zeros_like(x, np.float64), H, material=material)
    assert np.array_equal(rowptr, r2) and np.array_equal(col, c2)
    err = np.abs(val - kval).max() / np.abs(kval).max()
    say("max |K_gpu - K_oracle| / max |K|", err)
    assert err <= 1e-6


@pytest.mark.parametrize("precision", [P.FP32, P.FP64])
def test_one_step_matches_direct_solve(precision):
    rest, tets, mask, x0, v0 = case()
    say("one implicit step", precision)
    s = P.Solver(rest, tets, mask, P.default_params(precision=precision))
    s.set_positions(x0)
    s.set_velocities(v0, v0)
    x = s.positions(f64=True)
    tol = 1e-8 if precision == P.FP32 else 1e-12
    st, xg = s.implicit_step(params=P.implicit_params(cg_tolerance=tol), want_positions=True)
    xr, vr = oracle_solve(rest, tets, mask, x, v0.astype(np.float64))
    ex = np.abs(s.positions(f64=True) - xr).max()
    ev = ex / H / np.abs(vr).max()  # x' = x + h v' on every node
    say("CG its", st.cg_iterations, "residual", st.residual / st.rhs_norm, "rel v err", ev, "x err", ex)
    assert st.converged and st.residual <= tol * st.rhs_norm
    # FP32: the matrix is assembled and stored in FP32; FP64: FP32 rounding of AN, as above
    assert ev <= (5e-5 if precision == P.FP32 else 1e-6)
    assert ex <= (5e-7 if precision == P.FP32 else 5e-9)
    assert np.array_equal(xg, s.positions())
    s.close()


def test_trajectory_fp64_matches_oracle():
    rest, tets, mask, x0, v0 = case(seed=2)
    say("10 implicit steps, FP64, with an external force")
    ext = np.zeros_like(rest)
    ext[rest[:, 2] > 0.45, 1] = 0.5
    s = P.Solver(rest, tets, mask, P.default_params(precision=P.FP64))
    s.set_positions(x0)
    xo, vo = x0.astype(np.float64), np.zeros_like(x0, np.float64)
    ip = P.implicit_params(cg_tolerance=1e-12)
    for k in range(10):
        st = s.implicit_step(ext, params=ip)
        assert st.converged
        xo, vo = oracle_solve(rest, tets, mask, xo, vo, ext=ext.astype(np.float64))
    err = np.abs(s.positions(f64=True) - xo).max()
    moved = np.abs(xo - x0).max()
    say("max |x_gpu - x_oracle|", err, "displacement", moved)
    assert moved > 1e-3 and err <= 1e-8
    s.close()


def test_iteration_cap_and_determinism():
    rest, tets, mask, x0, v0 = case()
    say("CG cap and bitwise repeatability")
    res = []
    for _ in range(2):
        s = P.Solver(rest, tets, mask, P.default_params())
        s.set_positions(x0)
        st = s.implicit_step(params=P.implicit_params(cg_max_iterations=3))
        assert st.cg_iterations == 3 and not st.converged and st.residual < st.residual0
        for _ in range(3):
            s.implicit_step()
        res.append(s.positions())
        s.close()
    assert np.array_equal(res[0].view(np.uint32), res[1].view(np.uint32))


def test_timed_replays_give_the_same_result():
    rest, tets, mask, x0, _ = case()
    say("profiling on/off")
    out = []
    for prof in (0, 1):
        s = P.Solver(rest, tets, mask, P.default_params())
        s.set_positions(x0)
        if prof:
            s.set_profiling(1)
        its = sum(s.implicit_step().cg_iterations for _ in range(2))
        out.append(s.positions())
        if prof:
            kt = s.kernel_times()
            say("kernel times", {k: v for k, v in kt.items() if k.startswith("k_ie")})
            assert kt["k_ie_spmv"][1] == its and kt["k_ie_assemble"][1] == 2
        s.close()
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))


def test_batched_solver_is_rejected():
    rest, tets, mask, _, _ = case()
    b = P.Solver.batch([(rest, tets, mask)] * 2, [1e5, 1e5], [0.45, 0.45], P.default_params())
    with pytest.raises(P.PhybsError) as e:
        b.implicit_step()
    assert e.value.code == P.ERR_STATE
    b.close()


def test_full_size_stiffness_matrix_matches_oracle():
    """The stiffness matrix of the 64x64x256 bar (6,291,456 tets, 144 M nonzeros) in a seeded
    twisted state, FP64, against the oracle's construction: same CSR pattern, values within 1e-6 of
    the largest entry (the bound of the small-bar test)."""
    import time
    nx, ny, nz = 64, 64, 256
    rest, tets = P.mesh_bar(nx, ny, nz, 0.05)
    mask = (rest[:, 0] == 0).astype(np.int32)
    x = P.mesh_twist(rest, nx, ny, nz, 0.05, 30.0, 0.005, 7)
    say("stiffness matrix, 64x64x256 bar")
    t0 = time.perf_counter()
    s = P.Solver(rest, tets, mask, P.default_params(precision=P.FP64))
    s.set_positions(x)
    rowptr, col, val = s.stiffness_matrix(tets)
    s.close()
    t1 = time.perf_counter()
    mu, lam = O.lame(1e5, 0.45)
    pre = O.precompute(rest, tets, mask)
    r2, c2, kval, _, _ = O.implicit_system(rest, tets, mask, pre, mu, lam, x.astype(np.float64),
                                           np.zeros_like(x, np.float64), H)
    t2 = time.perf_counter()
    assert np.array_equal(rowptr, r2) and np.array_equal(col, c2)
    err = np.abs(val - kval).max() / np.abs(kval).max()
    say("nnz", len(col), "max |K_gpu - K_oracle| / max |K|", err, "gpu %.1f s, oracle %.1f s" % (t1 - t0, t2 - t1))
    assert err <= 1e-6

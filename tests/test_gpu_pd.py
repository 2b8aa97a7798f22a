"""GPU tests of projective dynamics (ProjDynamic.cpp): the local step against the oracle's (with
inverted tets, where the reference's sign convention matters), and whole steps against the
oracle's local steps with the global systems solved by a sparse direct solver (scipy; the
reference factorises once with Pardiso)."""
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
    return rest, tets, mask, x0


@pytest.mark.parametrize("precision", [P.FP32, P.FP64])
def test_local_step_matches_oracle(precision):
    rest, tets, mask, x0 = case()
    x = x0.copy()
    for t in tets[::7]:  # invert every 7th tet by reflecting its first node through the opposite face
        x[t[0]] = 2 * x[t[1:]].mean(axis=0) - x[t[0]]
    say("pd local step", precision)
    s = P.Solver(rest, tets, mask, P.default_params(precision=precision))
    got = s.pd_project(x, 1e6)
    pre = O.precompute(rest, tets, mask)
    ref = O.pd_project(x.astype(np.float64), tets, pre, 1e6)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say("max rel err", err)
    assert err <= (2e-5 if precision == P.FP32 else 1e-9)
    s.close()


def oracle_pd_step(rest, tets, mask, pre, x, v, ext=None, k=1e6, rounds=5, damping=0.99, precision=64):
    """doProjDynamics (ProjDynamic.cpp:442-479) with the oracle's local step and a direct solve
    (precision 32: the local step in FP32 and the matrix accumulated in FP32, as the reference)."""
    rowptr, col, g = O.pd_matrix(len(rest), tets, pre, k, H, precision=precision)
    n3 = 3 * len(rest)
    G = sp.csr_matrix((g, col, rowptr), shape=(n3, n3)).tocsc()
    m = pre["mass"].astype(np.float64)[:, None]
    f = pre["gforce"].astype(np.float64) + (0 if ext is None else ext)
    im = np.where(mask[:, None] != 0, 0.0, 1.0 / m)
    srhs = (x + H * v + H * H * im * f) * m / (H * H)
    xi = x.copy()
    for _ in range(rounds):
        b = O.pd_project(xi, tets, pre, k, s=srhs, precision=precision).astype(np.float64)
        xi = spl.spsolve(G, b.reshape(-1)).reshape(-1, 3)
    return xi, (xi - x) / H * damping


@pytest.mark.parametrize("precision", [P.FP32, P.FP64])
def test_steps_match_oracle(precision):
    rest, tets, mask, x0 = case(seed=2)
    ext = np.zeros_like(rest)
    ext[rest[:, 2] > 0.45, 1] = 0.5
    say("3 pd steps", precision)
    s = P.Solver(rest, tets, mask, P.default_params(precision=precision))
    s.set_positions(x0)
    pre = O.precompute(rest, tets, mask)
    xo, vo = x0.astype(np.float64), np.zeros_like(x0, np.float64)
    x32, v32 = xo.copy(), vo.copy()
    tol = 1e-12
    for _ in range(3):
        st, xg = s.pd_step(ext, P.pd_params(cg_tolerance=tol), want_positions=True)
        assert st.converged
        xo, vo = oracle_pd_step(rest, tets, mask, pre, xo, vo, ext.astype(np.float64))
        if precision == P.FP32:
            x32, v32 = oracle_pd_step(rest, tets, mask, pre, x32, v32, ext.astype(np.float64), precision=32)
    err = np.abs(s.positions(f64=True) - xo).max()
    moved = np.abs(xo - x0).max()
    say("CG its (last step)", st.cg_iterations, "max |x_gpu - x_oracle|", err, "displacement", moved)
    assert moved > 1e-3
    if precision == P.FP32:
        # the FP32 local step cancels terms ~20x larger than their sum and the matrix is stored in
        # FP32 (as the reference's): bounded by 3x the FP32 oracle's own distance from FP64
        own = np.abs(x32 - xo).max()
        say("FP32 oracle vs FP64 oracle", own)
        assert err <= 3 * own + 1e-7
    else:
        assert err <= 1e-8  # FP32 mesh data only; 3 steps of 5 global solves at 1e-12
    s.close()


def test_pd_step_is_deterministic_and_keeps_the_descent_usable():
    rest, tets, mask, x0 = case()
    say("pd determinism, then a descent step")
    res = []
    for _ in range(2):
        s = P.Solver(rest, tets, mask, P.default_params())
        s.set_positions(x0)
        for _ in range(2):
            s.pd_step()
        res.append(s.positions())
        st = s.step()
        assert np.isfinite(st.energy) and np.isfinite(s.positions()).all()
        s.close()
    assert np.array_equal(res[0].view(np.uint32), res[1].view(np.uint32))


def test_cached_matrix_follows_stiffness_and_parameters():
    """The PD matrix is built once per (stiffness, time step) and rebuilt when either changes or
    the parameters (masses) change: a step after such a change equals, bit for bit, the same step
    of a fresh solver put into the same state (FP32 state, so the state copies exactly)."""
    rest, tets, mask, x0 = case(seed=3)
    say("pd matrix cache")

    def fresh(state, **kw):
        s = P.Solver(rest, tets, mask, P.default_params(**kw))
        s.set_positions(state[0])
        s.set_velocities(state[1], state[2])
        return s

    a = P.Solver(rest, tets, mask, P.default_params())
    a.set_positions(x0)
    a.pd_step(params=P.pd_params(stiffness=1e6))
    a.pd_step(params=P.pd_params(stiffness=1e6))  # reuses the matrix
    st = (a.positions(), *a.velocities())
    a.pd_step(params=P.pd_params(stiffness=3e6))  # rebuilds it
    b = fresh(st)
    b.pd_step(params=P.pd_params(stiffness=3e6))
    assert np.array_equal(a.positions().view(np.uint32), b.positions().view(np.uint32))
    st = (a.positions(), *a.velocities())
    a.set_params(P.default_params(density=2500.0))  # new masses
    a.pd_step(params=P.pd_params(stiffness=3e6))
    c = fresh(st, density=2500.0)
    c.pd_step(params=P.pd_params(stiffness=3e6))
    assert np.array_equal(a.positions().view(np.uint32), c.positions().view(np.uint32))
    for s in (a, b, c):
        s.close()


def test_full_size_local_step_matches_oracle():
    """The local step on the 64x64x256 bar (6,291,456 tets) in a seeded twisted state with every
    50th tet inverted (the rotation's sign convention, taken by the SVD branch; the other tets take
    the polar iteration): FP64 within 1e-9 relative of the oracle's SVD-based projection."""
    nx, ny, nz = 64, 64, 256
    rest, tets = P.mesh_bar(nx, ny, nz, 0.05)
    mask = (rest[:, 0] == 0).astype(np.int32)
    x = P.mesh_twist(rest, nx, ny, nz, 0.05, 30.0, 0.005, 5).astype(np.float64)
    for t in tets[::50]:  # reflect the first node through the opposite face
        x[t[0]] = 2 * x[t[1:]].mean(axis=0) - x[t[0]]
    x = x.astype(np.float32)
    say("pd local step, 64x64x256 bar")
    s = P.Solver(rest, tets, mask, P.default_params(precision=P.FP64))
    got = s.pd_project(x, 1e6)
    s.close()
    pre = O.precompute(rest, tets, mask)
    ref = O.pd_project(x.astype(np.float64), tets, pre, 1e6)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say("max rel err", err)
    assert err <= 1e-9

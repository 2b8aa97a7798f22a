"""GPU parity: the HIP solver (through the C ABI) against the CPU oracle on the same seeded inputs.

Bars: (a) integers and setup data bit-exact; (b) per-tet forces within 1e-5 relative;
(c) per-node positions after 100 steps within 1e-4 of the bar length; plus determinism,
the rollback path, both reference semantics, FP64 mode and the stretch material.
"""
import os
import sys

import numpy as np
import pytest

import meshgen as M
import phybs_amd as P
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def bar_case(nx=4, ny=4, nz=16, edge=0.05, twist=90.0, jitter=0.01, seed=1):
    rest, tets = M.bar(nx, ny, nz, edge)
    mask = M.clamp_x0(rest)
    x0 = M.twisted_state(rest, nx, ny, nz, edge, twist, jitter, seed) if twist or jitter else rest.copy()
    return rest, tets, mask, x0


def oracle_pre(rest, tets, mask, grav=(0.0, -9.8, 0.0)):
    return O.precompute(rest, tets, mask, 1000.0, grav, 1e8)


# ---------------------------------------------------------------------------------------------
def test_setup_data_bit_exact():
    for name in ["bar4x4x16", "model_bar"]:
        if name == "bar4x4x16":
            rest, tets, mask, _ = bar_case()
        else:
            rest, tets = M.read_tetgen(os.path.join(M.GOLDEN, "model_bar/bar.1"))
            mask = (rest[:, 0] < -2.4).astype(np.int32)
        say("setup", name, len(rest), len(tets))
        s = P.Solver(rest, tets, mask, P.default_params())
        d = s.rest_data()
        pre = oracle_pre(rest, tets, mask)
        for k in ("dminv", "vol", "mass", "gforce"):
            a, b = d[k], pre[k].reshape(d[k].shape)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (name, k, np.abs(a - b).max())
        s.close()


def per_tet_rel_err(a, b):
    a = a.reshape(len(a), -1).astype(np.float64)
    b = b.reshape(len(b), -1).astype(np.float64)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-30)


@pytest.mark.parametrize("precision", [P.FP32, P.FP64])
@pytest.mark.parametrize("material", [P.MAT_NEOHOOKEAN, P.MAT_STRETCH])
def test_forces_match_oracle(precision, material):
    rest, tets, mask, x = bar_case(twist=60.0, jitter=0.02, seed=3)
    say("forces", precision, material)
    s = P.Solver(rest, tets, mask, P.default_params(precision=precision, material=material))
    tf, nf, te = s.compute_forces(x)
    mu, lam = O.lame(1e5, 0.45)
    pre = oracle_pre(rest, tets, mask)
    # oracle in FP64 from the same FP32 positions: the exact value of the formula
    rf, re_ = O.tet_forces(x.astype(np.float64), tets, pre, mu, lam, material, O.CLAMP_GPU, 64)
    err = per_tet_rel_err(tf, rf)
    say("per-tet force rel err max", err.max(), "p99", np.quantile(err, 0.99), "median", np.median(err))
    if precision == P.FP64:
        assert err.max() <= 1e-5
    else:
        # FP32 arithmetic (the reference's precision) has a rounding floor on low-stress tets:
        # on this state the FP32 oracle misses the exact value by 1e-4 at worst and 3e-5 at the
        # 99th percentile, so a pointwise 1e-5 bar holds only in FP64 mode (above).  FP32 bar:
        # median within 1e-5, and the error distribution (99th percentile, maximum) no worse
        # than twice the FP32 oracle's.
        rf32, _ = O.tet_forces(x, tets, pre, mu, lam, material, O.CLAMP_GPU, 32)
        err32 = per_tet_rel_err(rf32, rf)
        say("FP32 oracle vs exact: max", err32.max(), "p99", np.quantile(err32, 0.99), "median", np.median(err32))
        assert np.median(err) <= 1e-5
        assert np.quantile(err, 0.99) <= 2 * np.quantile(err32, 0.99)
        assert err.max() <= 2 * err32.max()
    # node forces: gather of the same per-tet forces
    off, ent = O.incidence(len(rest), tets)
    rnf = np.zeros((len(rest), 3))
    np.add.at(rnf, np.repeat(np.arange(len(rest)), np.diff(off)), rf.reshape(-1, 3)[ent // 3])
    scale = np.abs(rf).max()
    assert np.abs(nf - rnf).max() <= 1e-5 * scale
    # energies
    assert np.abs(te - re_).max() <= 1e-5 * np.abs(re_).max()
    s.close()


@pytest.mark.parametrize("material", [P.MAT_NEOHOOKEAN, P.MAT_STRETCH])
def test_hessian_diag_matches_oracle(material):
    rest, tets, mask, x = bar_case(twist=60.0, jitter=0.02, seed=5)
    mu, lam = O.lame(1e5, 0.45)
    pre = oracle_pre(rest, tets, mask)
    for prec in (P.FP32, P.FP64):
        say("hessian", material, prec)
        s = P.Solver(rest, tets, mask, P.default_params(precision=prec, material=material))
        for pos in (x, rest):
            if material == P.MAT_STRETCH and pos is rest:
                continue  # the oracle's Xu solve is singular for the stretch energy at F = I
            nh, th = s.compute_hessian_diag(pos)
            rh, rth = O.hessian_diag(pos.astype(np.float64), tets, pre, mu, lam, material, 64)
            err = per_tet_rel_err(th, rth)
            say("  tet Hd rel err max", err.max())
            # the reference perturbs repeated singular values by 1e-6 (IsotropicMaterial.cpp:887),
            # which moves the value by O(1e-6) relative; FP32 rounding adds ~1e-6
            assert err.max() <= 2e-5
            assert np.abs(nh - rh).max() <= 2e-5 * np.abs(rh).max()
        s.close()


def run_pair(rest, tets, mask, x0, steps, gparams, oparams, precision):
    s = P.Solver(rest, tets, mask, gparams)
    s.set_positions(x0)
    o = O.Solver(rest, tets, mask, params=oparams, precision=precision)
    o.set_state(x=x0.astype(np.float64))
    gst, ost = [], []
    for i in range(steps):
        gst.append(s.step())
        ost.append(o.step())
    xg = s.positions(f64=True)
    xo = o.state()[0]
    s.close()
    return xg, xo, gst, ost


def test_positions_after_100_steps_c0():
    """REF_GPU semantics, FP32 as the reference: 100 steps of 96 iterations from the twisted bar,
    positions within 1e-4 of the bar length of the FP32 oracle's.

    The reference sums per-node forces with float atomics (CUDAInterface.cu:526-541), so its own
    FP32 trajectory depends on the summation order, and the line search's accept/reject decisions
    amplify that over the steps; the oracle's own order sensitivity (ascending vs descending
    gather, both FP32) is printed beside the error.  The single-pass force and Hessian tests and
    the FP64 tracking test pin the arithmetic itself."""
    rest, tets, mask, x0 = bar_case()
    say("4x4x16 bar 100 steps REF_GPU fp32")
    xg, xo, gst, ost = run_pair(rest, tets, mask, x0, 100, P.default_params(), O.default_params(), 32)
    o = O.Solver(rest, tets, mask, params=O.default_params(gather_order=1), precision=32)
    o.set_state(x=x0.astype(np.float64))
    for _ in range(100):
        o.step()
    L = 16 * 0.05
    noise = np.abs(o.state()[0] - xo).max()
    err = np.abs(xg - xo).max()
    say("max |x_gpu - x_oracle| = %.2e L, oracle order noise %.2e L, rollbacks" % (err / L, noise / L),
        sum(g.rollbacks for g in gst), sum(q.rollbacks for q in ost))
    assert np.isfinite(xg).all()
    assert err <= 1e-4 * L


def test_positions_ref_cpu_semantics():
    """REF_CPU semantics (gravity counted twice in the gradient, once in the energy) rolls back
    hundreds of times per 100 steps, so its accept/reject decisions amplify rounding: the FP32
    oracle against itself in FP64 drifts to 1e-2 L by step 50.  The 100-step bar is checked in
    FP64, where GPU and oracle make the same decisions; FP32 is checked over 10 steps."""
    rest, tets, mask, x0 = bar_case()
    L = 16 * 0.05
    say("4x4x16 bar 100 steps REF_CPU fp64")
    xg, xo, gst, ost = run_pair(rest, tets, mask, x0, 100, P.default_params(semantics=P.SEM_REF_CPU, precision=P.FP64),
                                O.default_params(semantics=O.REF_CPU), 64)
    err = np.abs(xg - xo).max()
    say("fp64 REF_CPU err %.2e L" % (err / L), "rollbacks", sum(g.rollbacks for g in gst), sum(q.rollbacks for q in ost))
    assert err <= 1e-4 * L
    assert sum(g.rollbacks for g in gst) > 50
    say("4x4x16 bar 10 steps REF_CPU fp32")
    xg, xo, _, _ = run_pair(rest, tets, mask, x0, 10, P.default_params(semantics=P.SEM_REF_CPU),
                            O.default_params(semantics=O.REF_CPU), 32)
    assert np.abs(xg - xo).max() <= 1e-4 * L


def test_fp64_mode_tracks_fp64_oracle():
    rest, tets, mask, x0 = bar_case()
    say("4x4x16 bar fp64 30 steps")
    xg, xo, gst, ost = run_pair(rest, tets, mask, x0, 30, P.default_params(precision=P.FP64), O.default_params(), 64)
    err = np.abs(xg - xo).max()
    say("fp64 max err", err)
    assert err <= 1e-7 * 0.8
    assert [g.rollbacks for g in gst] == [q.rollbacks for q in ost]
    assert [g.passes for g in gst] == [q.passes for q in ost]


def test_bitwise_deterministic():
    rest, tets, mask, x0 = bar_case()
    say("determinism")
    outs = []
    for _ in range(2):
        s = P.Solver(rest, tets, mask, P.default_params())
        s.set_positions(x0)
        for i in range(5):
            s.step()
        outs.append(s.positions())
        s.close()
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


@pytest.mark.parametrize("relgrad", [False, True])
def test_graph_replay_matches_eager_launches(relgrad):
    """Descent passes replayed from a captured HIP graph (8 passes per replay, so each step takes
    several replays with host syncs and eager kernels between them) give bitwise the same
    trajectory and loop statistics as launching the kernels one by one; with timing events
    inside every replay, and inside every third replay, as well."""
    rest, tets, mask, x0 = bar_case(twist=30.0)
    say("graph replay vs eager, relgrad", relgrad)
    kw = dict(precision=P.FP64, relgrad_tol=1e-4, max_iterations=5000) if relgrad else dict(alpha0=0.4)
    res = []
    for gp, prof in ((-1, 0), (8, 0), (8, 1), (8, 3)):
        s = P.Solver(rest, tets, mask, P.default_params(graph_passes=gp, **kw))
        s.set_positions(x0)
        s.set_profiling(prof)
        st = [s.step() for _ in range(4)]
        ext = np.zeros_like(rest)
        ext[-5:, 1] = 0.3
        st.append(s.step(ext))
        res.append((s.positions(f64=True), [(q.passes, q.iterations, q.rollbacks, q.g_norm) for q in st],
                    s.kernel_times() if prof else None))
        s.close()
    for xg, stats, _ in res[1:]:
        assert np.array_equal(xg.view(np.uint64), res[0][0].view(np.uint64))
        assert stats == res[0][1]
    kt = res[2][2]
    passes = sum(p for p, _, _, _ in res[2][1])
    assert kt["k_tile"][1] == passes and kt["k_node"][1] == passes and kt["k_control"][1] == passes
    assert kt["k_tile"][0] > 0


def test_rollback_path_matches_oracle():
    # alpha0 = 0.4 on the 30-degree bar overshoots once in step 3 (energy rises at a check)
    rest, tets, mask, x0 = bar_case(twist=30.0)
    say("rollback path")
    gp = P.default_params(alpha0=0.4, precision=P.FP64)
    op = O.default_params(alpha0=0.4)
    xg, xo, gst, ost = run_pair(rest, tets, mask, x0, 3, gp, op, 64)
    say("gpu rollbacks", [g.rollbacks for g in gst], "oracle", [q.rollbacks for q in ost])
    assert sum(q.rollbacks for q in ost) > 0
    assert [g.rollbacks for g in gst] == [q.rollbacks for q in ost]
    assert [g.iterations for g in gst] == [q.iterations for q in ost]
    assert [g.alpha for g in gst] == pytest.approx([q.alpha for q in ost])
    assert np.abs(xg - xo).max() <= 1e-6


def test_relgrad_mode_converges_like_oracle():
    rest, tets, mask, x0 = bar_case()
    say("rel-grad 1e-4")
    gp = P.default_params(precision=P.FP64, relgrad_tol=1e-4, max_iterations=5000)
    op = O.default_params(relgrad_tol=1e-4, iterations=5000)
    xg, xo, gst, ost = run_pair(rest, tets, mask, x0, 3, gp, op, 64)
    for g, q in zip(gst, ost):
        say("iters gpu", g.iterations, "oracle", q.iterations, "g/g0", g.g_norm / g.g0_norm)
        assert g.converged == 1 and g.g_norm <= 1e-4 * g.g0_norm
        assert abs(g.iterations - q.iterations) <= 2
    assert np.abs(xg - xo).max() <= 1e-6


@pytest.mark.parametrize("switch", [1e-3, 2e-4])
def test_mixed_mode_converges_like_fp64_oracle(switch):
    """FP64 state with FP32 per-tet arithmetic until |g|/|g0| <= switch, then FP64: every step
    still reaches rel-grad 1e-4 (judged in FP64), the positions stay within the rel-grad mode's
    own bound of the FP64 oracle, and the iteration count stays close to the FP64 one.  (A
    rollback during the FP32 passes sends the step on in FP64 with a restarted Chebyshev
    schedule, so single steps may take a different number of iterations.)"""
    rest, tets, mask, x0 = bar_case()
    say("mixed mode, switch", switch)
    gp = P.default_params(precision=P.FP64, relgrad_tol=1e-4, max_iterations=5000, mixed_switch=switch)
    op = O.default_params(relgrad_tol=1e-4, iterations=5000)
    xg, xo, gst, ost = run_pair(rest, tets, mask, x0, 3, gp, op, 64)
    for g, q in zip(gst, ost):
        say("iters mixed", g.iterations, "oracle", q.iterations, "rollbacks", g.rollbacks, "g/g0", g.g_norm / g.g0_norm)
        assert g.converged == 1 and g.g_norm <= 1e-4 * g.g0_norm
    im, io = sum(g.iterations for g in gst), sum(q.iterations for q in ost)
    assert abs(im - io) <= 0.25 * io
    err = np.abs(xg - xo).max()
    say("max |x_mixed - x_oracle|", err)
    assert err <= 1e-6


def test_mixed_mode_tracks_fp64_on_a_strongly_twisted_bar():
    """The 12x12x50 bar released from a 90 degree twist, gravity -z: its internal forces are large
    against the step's gradient, so the FP32 passes can reach their rounding floor (|g| ~ 5e-5 |f|)
    before |g|/|g0| = 2e-4; the |f| test then moves the step to FP64 before the line search sees
    the noise.  No rollbacks, iteration counts within 2 % of the FP64 run's, positions within
    1e-6 of the bar length of it."""
    rest, tets, mask, x0 = bar_case(12, 12, 50)
    say("mixed mode on the twisted 12x12x50 bar")
    out = []
    for sw in (0.0, 2e-4):
        gp = P.default_params(precision=P.FP64, relgrad_tol=1e-4, max_iterations=20000, mixed_switch=sw,
                              grav=(0.0, 0.0, -9.8), profile=[(1, 0.96), (7, 0.991), (10, 0.9995)])
        s = P.Solver(rest, tets, mask, gp)
        s.set_positions(x0)
        st = [s.step() for _ in range(12)]
        out.append((s.positions(f64=True), st))
        s.close()
    (x64, s64), (xmx, smx) = out
    say("iterations fp64", [q.iterations for q in s64], "mixed", [q.iterations for q in smx])
    assert all(q.converged and q.rollbacks == 0 for q in smx)
    assert abs(sum(q.iterations for q in smx) - sum(q.iterations for q in s64)) <= 0.02 * sum(q.iterations for q in s64)
    err = np.abs(xmx - x64).max()
    say("max |x_mixed - x_fp64|", err)
    assert err <= 1e-6 * 2.5  # 1e-6 of the bar length (2.5 m)


def test_mixed_mode_rejects_fp32_state():
    rest, tets, mask, _ = bar_case(1, 1, 2, twist=0.0, jitter=0.0)
    with pytest.raises(P.PhybsError):
        P.Solver(rest, tets, mask, P.default_params(precision=P.FP32, mixed_switch=1e-3))


def test_stretch_material_steps():
    rest, tets, mask, x0 = bar_case(twist=30.0)
    say("stretch material 20 steps")
    gp = P.default_params(material=P.MAT_STRETCH)
    s = P.Solver(rest, tets, mask, gp)
    s.set_positions(x0)
    o = O.Solver(rest, tets, mask, params=O.default_params(), precision=32, material=O.MAT_STRETCH)
    o.set_state(x=x0.astype(np.float64))
    for _ in range(20):
        s.step()
        o.step()
    err = np.abs(s.positions() - o.state()[0]).max()
    say("stretch err", err)
    assert err <= 1e-4 * 0.8
    s.close()


def test_do_descent_opt_compat():
    rest, tets, mask, x0 = bar_case()
    say("doDescentOpt compat")
    a = P.Solver(rest, tets, mask, P.default_params())
    b = P.Solver(rest, tets, mask, P.default_params())
    a.set_positions(x0)
    b.set_positions(x0)
    ext = np.zeros_like(x0)
    ext[-1] = (0.0, 0.5, 0.0)
    for _ in range(3):
        ya = a.do_descent_opt(0.5, 96, ext)  # h is ignored as in the reference
        _, yb = b.step(ext, want_positions=True)
    assert np.array_equal(ya, yb)
    a.close()
    b.close()


def test_external_force_matches_oracle():
    rest, tets, mask, x0 = bar_case(twist=0.0, jitter=0.0)
    say("external force")
    ext = np.zeros_like(rest, dtype=np.float64)
    ext[rest[:, 0] > 0.19] = (0.0, 0.0, 2.0)
    s = P.Solver(rest, tets, mask, P.default_params())
    o = O.Solver(rest, tets, mask, precision=32)
    for _ in range(10):
        s.step(ext.astype(np.float32))
        o.step(ext)
    assert np.abs(s.positions() - o.state()[0]).max() <= 1e-4 * 0.8
    s.close()


def test_dragged_face_ring_matches_oracle():
    """The interactive input path on a TetGen mesh: a ray picks a boundary face, dragging it sets
    the ring's external forces (host C ABI), and the GPU steps under them like the oracle."""
    nodes, tets, faces = P.tetgen_load(os.path.join(M.GOLDEN, "model_bar/bar.1"), 0.1, faces=True)
    mask = (nodes[:, 0] == nodes[:, 0].min()).astype(np.int32)
    off, ring = P.face_ring(len(nodes), faces)
    far = int(np.argmax(nodes[faces].mean(axis=1)[:, 0]))  # a face at the free end
    f = faces[far]
    n = np.cross(nodes[f[2]] - nodes[f[0]], nodes[f[1]] - nodes[f[0]])
    n /= np.linalg.norm(n)
    c = nodes[f].mean(axis=0)
    pick = P.pick_face(nodes, faces, c + 0.3 * n, -n)
    say("picked face", pick, "of", len(faces), "ring", off[pick + 1] - off[pick])
    assert pick >= 0
    mass = oracle_pre(nodes, tets, mask)["mass"]
    ext = P.drag_face_ring(nodes, mass, faces, off, ring, pick, np.float32([0.0, 0.02, 0.01]))
    assert np.count_nonzero(np.abs(ext).sum(axis=1)) == off[pick + 1] - off[pick]
    s = P.Solver(nodes, tets, mask, P.default_params())
    o = O.Solver(nodes, tets, mask, precision=32)
    for _ in range(10):
        s.step(ext)
        o.step(ext.astype(np.float64))
    L = float(nodes[:, 0].max() - nodes[:, 0].min())
    err = np.abs(s.positions() - o.state()[0]).max()
    say("max |x_gpu - x_oracle| / L", err / L)
    assert err <= 1e-4 * L
    s.close()


def test_inverted_rest_tet_is_rejected():
    rest, tets, mask, _ = bar_case(1, 1, 1, twist=0.0, jitter=0.0)
    bad = tets.copy()
    bad[0, [2, 3]] = bad[0, [3, 2]]
    say("inverted rest tet")
    with pytest.raises(P.PhybsError) as e:
        P.Solver(rest, bad, mask, P.default_params())
    assert e.value.code == P.ERR_MESH


def single_tet():
    rest = np.array([[0, 0, 0], [0.05, 0, 0], [0, 0.05, 0], [0, 0, 0.05]], np.float32)
    tets = np.array([[0, 1, 2, 3]], np.int32)
    return rest, tets, M.clamp_x0(rest)  # nodes 0, 2, 3 lie on x = 0: one free node


def test_single_tet_tracks_fp64_oracle():
    """The smallest mesh: one tet, one free node, pulled by gravity from a displaced start; FP64
    against the FP64 oracle over 20 steps.  (Once the node rests, successive energies differ only
    by rounding, so the line search's accept/rollback decisions, and with them the pass counts,
    may differ from the oracle's; the positions may not.)"""
    rest, tets, mask = single_tet()
    x0 = rest.copy()
    x0[1] += (0.004, -0.003, 0.002)
    say("single tet fp64 20 steps")
    xg, xo, gst, ost = run_pair(rest, tets, mask, x0, 20, P.default_params(precision=P.FP64),
                                O.default_params(), 64)
    err = np.abs(xg - xo).max()
    say("single tet max err", err, "moved", np.abs(xo - x0).max())
    assert np.abs(xo - x0).max() > 1e-4
    assert err <= 1e-10
    assert [g.passes for g in gst[:5]] == [q.passes for q in ost[:5]]  # while the node still moves


@pytest.mark.parametrize("relgrad", [False, True])
def test_all_nodes_fixed_leaves_the_mesh_in_place(relgrad):
    """Every node clamped: nothing moves, the step ends (rel-grad: at once, |g0| = 0)."""
    rest, tets, _, x0 = bar_case(2, 2, 3)
    mask = np.ones(len(rest), np.int32)
    kw = dict(precision=P.FP64, relgrad_tol=1e-4, max_iterations=500) if relgrad else {}
    say("all nodes fixed, relgrad", relgrad)
    s = P.Solver(rest, tets, mask, P.default_params(**kw))
    s.set_positions(x0)
    for _ in range(3):
        st = s.step()
        assert np.isfinite(st.energy)
    assert np.array_equal(s.positions(), x0)
    if relgrad:
        assert st.converged == 1 and st.iterations == 0
    s.close()



def test_twisted_12x12x50_bar_forces_and_fp64_steps():
    """The 12x12x50 bar (43,200 tets) released from a 90 degree twist with 1 % jitter: per-tet
    forces in FP64 within 1e-5 of the oracle's, then 10 FP64 steps of 96 iterations within 1e-7 of
    the bar length (2.5 m) of the FP64 oracle's, with identical rollbacks and pass counts."""
    rest, tets, mask, x0 = bar_case(12, 12, 50, twist=90.0, jitter=0.01, seed=1)
    say("12x12x50 bar: forces and 10 FP64 steps")
    s = P.Solver(rest, tets, mask, P.default_params(precision=P.FP64))
    tf, _, _ = s.compute_forces(x0)
    s.close()
    mu, lam = O.lame(1e5, 0.45)
    pre = oracle_pre(rest, tets, mask)
    rf, _ = O.tet_forces(x0.astype(np.float64), tets, pre, mu, lam, O.MAT_NEOHOOKEAN, O.CLAMP_GPU, 64)
    err = per_tet_rel_err(tf, rf)
    say("per-tet force rel err max", err.max())
    assert err.max() <= 1e-5
    threads = min(16, len(os.sched_getaffinity(0)))
    s = P.Solver(rest, tets, mask, P.default_params(precision=P.FP64))
    s.set_positions(x0)
    o = O.Solver(rest, tets, mask, params=O.default_params(), precision=64, threads=threads)
    o.set_state(x=x0.astype(np.float64))
    gst = [s.step() for _ in range(10)]
    ost = [o.step() for _ in range(10)]
    err = np.abs(s.positions(f64=True) - o.state()[0]).max()
    say("10 steps max |x_gpu - x_oracle|", err, "passes", [g.passes for g in gst])
    assert err <= 1e-7 * 2.5
    assert [g.rollbacks for g in gst] == [q.rollbacks for q in ost]
    assert [g.passes for g in gst] == [q.passes for q in ost]
    s.close()


@pytest.mark.parametrize("material", [P.MAT_NEOHOOKEAN, P.MAT_STRETCH])
def test_full_size_bar_forces_match_oracle(material):
    """The 64x64x256 bar of the benchmark (6,291,456 tets) in a seeded twisted state, both
    materials (the stretch material's forces via the polar rotation, the oracle's via the SVD):
    every per-tet force of the FP64 path within 1e-5 of the oracle's, and the total elastic energy
    (the per-tet values are returned in FP32) within 1e-6 relative."""
    nx, ny, nz = 64, 64, 256
    rest, tets = P.mesh_bar(nx, ny, nz, 0.05)
    mask = (rest[:, 0] == 0).astype(np.int32)
    x = P.mesh_twist(rest, nx, ny, nz, 0.05, 30.0, 0.005, 3)
    say("64x64x256 bar forces, material", material)
    s = P.Solver(rest, tets, mask, P.default_params(precision=P.FP64, material=material))
    tf, _, te = s.compute_forces(x)
    s.close()
    mu, lam = O.lame(1e5, 0.45)
    pre = O.precompute(rest, tets, mask, 1000.0, (0.0, -9.8, 0.0), 1e8)
    rf, re_ = O.tet_forces(x.astype(np.float64), tets, pre, mu, lam, material, O.CLAMP_GPU, 64)
    err = per_tet_rel_err(tf, rf)
    eg, eo = te.astype(np.float64).sum(), re_.astype(np.float64).sum()
    say("per-tet force rel err max", err.max(), "energy", eg, eo)
    assert err.max() <= 1e-5
    assert abs(eg - eo) <= 1e-6 * abs(eo)


def test_full_size_relgrad_step_judged_by_the_oracle():
    """One FP64 rel-grad 1e-4 step of the benchmark's 64x64x256 bar from rest: the oracle, given
    the same step start, evaluates the incremental-potential gradient at the plain guess and at the
    GPU's result; its norms agree with the solver's |g0| and |g| to 1e-6 and the GPU's result meets
    rel-grad 1e-4 by the oracle's own measure."""
    nx, ny, nz = 64, 64, 256
    rest, tets = P.mesh_bar(nx, ny, nz, 0.05)
    mask = (rest[:, 0] == 0).astype(np.int32)
    grav, prof = (0.0, 0.0, -9.8), [(1, 0.96), (7, 0.991), (10, 0.9995)]
    say("64x64x256 bar: one rel-grad step judged by the oracle")
    s = P.Solver(rest, tets, mask, P.default_params(precision=P.FP64, relgrad_tol=1e-4, max_iterations=40000,
                                                   dt=1.0 / 30.0, grav=grav, profile=prof))
    st, _ = s.step(want_positions=True)
    xg = s.positions(f64=True)
    s.close()
    assert st.converged == 1
    threads = min(16, len(os.sched_getaffinity(0)))
    o = O.Solver(rest, tets, mask, params=O.default_params(iterations=1, dt=1.0 / 30.0, grav=grav, profile=prof),
                 precision=64, threads=threads)
    o.set_state(x=rest.astype(np.float64))
    o.step()  # sets the step start (x~ = x + h v) the gradients are taken against
    g0 = o.gradient_norm(rest.astype(np.float64))
    g = o.gradient_norm(xg)
    say("oracle |g0|", g0, "gpu", st.g0_norm, "oracle |g|", g, "gpu", st.g_norm, "iterations", st.iterations)
    assert abs(g0 - st.g0_norm) <= 1e-6 * g0
    assert abs(g - st.g_norm) <= 1e-6 * g0
    assert g <= 1e-4 * g0

"""GPU parity on the benchmark configurations themselves, against the CPU oracle:

* the 64x64x256 bar (6,291,456 tets): per-node elastic forces, stiffness diagonal and elastic
  energy of one tile pass of the descent (the owned and halo partials summed as the node pass sums
  them), FP64 and FP32;
* the 12x12x50 bar (43,200 tets) with a 90 degree twist and 1 % jitter: 500 FP32 steps of 96
  iterations, the configuration's full length;
* the 256-bar batch (5.7 M tets in one batched solver, seeded sizes and materials): instances
  against the oracle stepping each instance alone, FP64 and FP32 fixed iterations and FP64 rel-grad.
"""
import os
import sys

import numpy as np
import pytest

import meshgen as M
import phybs_amd as P
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

BENCH_PROFILE = [(1, 0.96), (7, 0.991), (10, 0.9995)]
GRAV_Z = (0.0, 0.0, -9.8)


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def threads():
    return min(16, len(os.sched_getaffinity(0)))


def node_sum_scale(tf, tets, nn):
    """Per node, the sum of the magnitudes of the corner forces it receives (the scale of the
    rounding of a per-node sum)."""
    s = np.zeros(nn)
    np.add.at(s, tets.reshape(-1), np.linalg.norm(tf.reshape(-1, 3), axis=1))
    return s


@pytest.fixture(scope="module")
def bar64():
    nx, ny, nz = 64, 64, 256
    rest, tets = P.mesh_bar(nx, ny, nz, 0.05)
    mask = (rest[:, 0] == 0).astype(np.int32)
    x = P.mesh_twist(rest, nx, ny, nz, 0.05, 30.0, 0.005, 3)
    mu, lam = O.lame(1e5, 0.45)
    pre = O.precompute(rest, tets, mask, 1000.0, (0.0, -9.8, 0.0), 1e8)
    rf, re_ = O.tet_forces(x.astype(np.float64), tets, pre, mu, lam, O.MAT_NEOHOOKEAN, O.CLAMP_GPU, 64)
    off, ent = O.incidence(len(rest), tets)
    rnf = np.zeros((len(rest), 3))
    np.add.at(rnf, np.repeat(np.arange(len(rest)), np.diff(off)), rf.reshape(-1, 3)[ent // 3])
    scale = node_sum_scale(rf, tets, len(rest))
    return dict(rest=rest, tets=tets, mask=mask, x=x, mu=mu, lam=lam, pre=pre, rf=rf, re=re_, rnf=rnf, scale=scale)


def test_bar64_tile_pass_node_forces_fp64(bar64):
    """FP64: every node's force within 1e-5 of the magnitude of the terms it sums (the oracle sums
    the per-tet forces in ascending tet order), the stiffness diagonal within 2e-5 of the oracle's
    Xu et al. construction, the pass's FP64 energy within 1e-13 of the oracle's (both evaluate the
    logarithm of J in FP64, the device with its own series, the oracle with the C library)."""
    d = bar64
    say("64x64x256 tile pass, FP64")
    s = P.Solver(d["rest"], d["tets"], d["mask"], P.default_params(precision=P.FP64))
    nf, nh, e = s.tile_forces(d["x"], hessian=True)
    s.close()
    err = np.linalg.norm(nf - d["rnf"], axis=1) / np.maximum(d["scale"], 1e-30)
    say("per-node force err / sum|terms|: max %.2e median %.2e" % (err.max(), np.median(err)))
    assert err.max() <= 1e-5
    eo = float(np.sum(d["re"]))
    say("energy gpu %.15e oracle %.15e, relative difference %.2e" % (e, eo, abs(e - eo) / abs(eo)))
    assert abs(e - eo) <= 1e-13 * abs(eo)
    rh, _ = O.hessian_diag(d["x"].astype(np.float64), d["tets"], d["pre"], d["mu"], d["lam"], O.MAT_NEOHOOKEAN, 64)
    herr = np.abs(nh - rh).max() / np.abs(rh).max()
    say("stiffness diagonal max err / max %.2e" % herr)
    assert herr <= 2e-5


def test_bar64_tile_pass_node_forces_fp32(bar64):
    """FP32 (the reference's arithmetic): per-node forces against the FP64 oracle; the error
    distribution no worse than twice the FP32 oracle's own (same formula, FP32, ascending order),
    median within 1e-5 of the magnitude of the summed terms."""
    d = bar64
    say("64x64x256 tile pass, FP32")
    s = P.Solver(d["rest"], d["tets"], d["mask"], P.default_params())
    nf, _, e = s.tile_forces(d["x"])
    s.close()
    rf32, _ = O.tet_forces(d["x"], d["tets"], d["pre"], d["mu"], d["lam"], O.MAT_NEOHOOKEAN, O.CLAMP_GPU, 32)
    off, ent = O.incidence(len(d["rest"]), d["tets"])
    r32 = np.zeros((len(d["rest"]), 3), np.float32)
    # the FP32 oracle's per-node sums in FP32, ascending tet order
    rows = rf32.reshape(-1, 3)[ent // 3]
    for k in range(int(np.diff(off).max())):
        sel = np.diff(off) > k
        r32[sel] += rows[off[:-1][sel] + k]
    err = np.linalg.norm(nf - d["rnf"], axis=1) / np.maximum(d["scale"], 1e-30)
    err32 = np.linalg.norm(r32.astype(np.float64) - d["rnf"], axis=1) / np.maximum(d["scale"], 1e-30)
    say("per-node err / sum|terms|: gpu max %.2e p99 %.2e median %.2e; FP32 oracle max %.2e p99 %.2e" %
        (err.max(), np.quantile(err, 0.99), np.median(err), err32.max(), np.quantile(err32, 0.99)))
    assert np.median(err) <= 1e-5
    assert np.quantile(err, 0.99) <= 2 * np.quantile(err32, 0.99)
    assert err.max() <= 2 * err32.max()
    eo = float(np.sum(d["re"]))
    say("energy gpu %.9e oracle %.9e" % (e, eo))
    assert abs(e - eo) <= 1e-6 * abs(eo)


def c1_case():
    rest, tets = M.bar(12, 12, 50, 0.05)
    return rest, tets, M.clamp_x0(rest), M.twisted_state(rest, 12, 12, 50, 0.05, 90.0, 0.01, 1)


def test_bar12x12x50_500_fp32_steps():
    """The 12x12x50 bar released from a 90 degree twist with 1 % jitter: 500 FP32 steps of 96
    iterations (dt = 1/30, gravity -z), REF_GPU semantics, against the FP32 oracle.

    After 100 steps the positions agree within 1e-4 of the bar length (2.5 m).  Beyond that the
    FP32 trajectory of this configuration depends on the summation order itself: the reference
    sums node forces with float atomics (CUDAInterface.cu:526-541), and the FP32 oracle against
    itself with the opposite per-node order drifts to ~5e-4 L by step 200 as the bar settles and
    the line search rolls back hundreds of times.  Over steps 100-500 the GPU's distance from the
    oracle therefore stays within 3 times the largest such order sensitivity seen at the same
    checkpoints; the FP64 test below holds the arithmetic over the same 500 steps."""
    rest, tets, mask, x0 = c1_case()
    L = 2.5
    say("12x12x50 bar, 500 FP32 steps")
    s = P.Solver(rest, tets, mask, P.default_params(dt=1.0 / 30.0, grav=GRAV_Z, profile=BENCH_PROFILE))
    s.set_positions(x0)
    orcs = []
    for order in (0, 1):
        o = O.Solver(rest, tets, mask, params=O.default_params(dt=1.0 / 30.0, grav=GRAV_Z, profile=BENCH_PROFILE,
                                                               gather_order=order), precision=32, threads=threads())
        o.set_state(x=x0.astype(np.float64))
        orcs.append(o)
    rb = [0, 0]
    errs, noises = {}, {}
    for k in range(1, 501):
        rb[0] += s.step().rollbacks
        rb[1] += orcs[0].step().rollbacks
        orcs[1].step()
        if k in (1, 10, 50, 100, 150, 200, 250, 300, 400, 500):
            xg, xo, xr = s.positions(f64=True), orcs[0].state()[0], orcs[1].state()[0]
            assert np.isfinite(xg).all()
            errs[k], noises[k] = np.abs(xg - xo).max() / L, np.abs(xr - xo).max() / L
            say("step %3d: |x_gpu - x_oracle| %.2e L, oracle order noise %.2e L, rollbacks gpu %d oracle %d" %
                (k, errs[k], noises[k], rb[0], rb[1]))
    s.close()
    assert errs[100] <= 1e-4
    late = [k for k in errs if k >= 100]
    assert max(errs[k] for k in late) <= 3.0 * max(noises[k] for k in late)


def test_bar12x12x50_500_fp64_steps():
    """The same 500 steps in FP64 against the FP64 oracle: positions within 1e-6 of the bar length
    at every checkpoint, and the same rollbacks while the bar still moves (the first 250 steps).
    Once it has come to rest, successive energies differ only by rounding, so the line search's
    accept/rollback decisions depend on the order of the energy sums (the GPU's fixed-order FP64
    reduction against the oracle's sequential sum) and the rollback counts part."""
    rest, tets, mask, x0 = c1_case()
    L = 2.5
    say("12x12x50 bar, 500 FP64 steps")
    s = P.Solver(rest, tets, mask, P.default_params(dt=1.0 / 30.0, grav=GRAV_Z, profile=BENCH_PROFILE,
                                                   precision=P.FP64))
    s.set_positions(x0)
    o = O.Solver(rest, tets, mask, params=O.default_params(dt=1.0 / 30.0, grav=GRAV_Z, profile=BENCH_PROFILE),
                 precision=64, threads=threads())
    o.set_state(x=x0.astype(np.float64))
    rb = [0, 0]
    worst = 0.0
    for k in range(1, 501):
        rb[0] += s.step().rollbacks
        rb[1] += o.step().rollbacks
        if k % 50 == 0:
            err = np.abs(s.positions(f64=True) - o.state()[0]).max() / L
            worst = max(worst, err)
            say("step %3d: |x_gpu - x_oracle| %.2e L, rollbacks gpu %d oracle %d" % (k, err, rb[0], rb[1]))
        if k == 250:
            assert rb[0] == rb[1]
    s.close()
    assert worst <= 1e-6


def batch_instances():
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench

    return bench.batch_instances(P)


def pick_instances(inst):
    """Smallest and largest bar, softest and stiffest material."""
    nts = [len(i[1]) for i in inst]
    Es = [i[4] for i in inst]
    return sorted({int(np.argmin(nts)), int(np.argmax(nts)), int(np.argmin(Es)), int(np.argmax(Es))})


@pytest.fixture(scope="module")
def batch256():
    inst = batch_instances()
    say("256-bar batch:", sum(len(i[1]) for i in inst), "tets")
    return inst


def run_batch(inst, params, steps, picks):
    b = P.Solver.batch([(i[0], i[1], i[2]) for i in inst], [i[4] for i in inst], [i[5] for i in inst], params)
    b.set_positions(np.concatenate([i[3] for i in inst]))
    stats = {k: [] for k in picks}
    allst = []
    for _ in range(steps):
        b.step()
        allst.append([b.instance_stats(i) for i in range(len(inst))])
        for k in picks:
            q = allst[-1][k]
            stats[k].append((q.passes, q.iterations, q.rollbacks, q.converged))
    x = b.positions(f64=True)
    b.close()
    offs = np.cumsum([0] + [len(i[0]) for i in inst])
    return {k: x[offs[k]:offs[k + 1]] for k in picks}, stats, allst


def run_oracle(i, params, steps, precision):
    rest, tets, mask, x0, E, nu = i
    o = O.Solver(rest, tets, mask, E=E, nu=nu, params=params, precision=precision, threads=threads())
    o.set_state(x=x0.astype(np.float64))
    st = [o.step() for _ in range(steps)]
    return o.state()[0], [(q.passes, q.iterations, q.rollbacks, q.converged) for q in st]


@pytest.mark.parametrize("precision", [P.FP64, P.FP32])
def test_batch256_instances_match_oracle(batch256, precision):
    """The 256-bar batch (96 descent iterations per step, dt = 1/30, gravity -z) in one batched solver, 2
    steps; the smallest, largest, softest and stiffest instance against the oracle stepping that
    instance alone: FP64 within 1e-7 of the bar's length with identical pass, iteration and
    rollback counts; FP32 within 1e-4 of the bar's length."""
    inst = batch256
    picks = pick_instances(inst)
    kw = dict(dt=1.0 / 30.0, grav=GRAV_Z, profile=BENCH_PROFILE, iterations=96, precision=precision)
    say("batch256", precision, "instances", picks, [(len(inst[k][1]), round(inst[k][4])) for k in picks])
    xs, stats, _ = run_batch(inst, P.default_params(**kw), 2, picks)
    op = O.default_params(dt=1.0 / 30.0, grav=GRAV_Z, profile=BENCH_PROFILE, iterations=96)
    for k in picks:
        xo, so = run_oracle(inst[k], op, 2, 64 if precision == P.FP64 else 32)
        L = float(inst[k][0][:, 2].max())
        err = np.abs(xs[k] - xo).max()
        say("instance %d (%d tets, E %.3g): |x_gpu - x_oracle| %.2e L, passes gpu %s oracle %s" %
            (k, len(inst[k][1]), inst[k][4], err / L, [q[0] for q in stats[k]], [q[0] for q in so]))
        if precision == P.FP64:
            assert err <= 1e-7 * L
            assert [q[:3] for q in stats[k]] == [q[:3] for q in so]
        else:
            assert err <= 1e-4 * L


def test_batch256_relgrad_every_instance_converges(batch256):
    """The same batch in FP64 rel-grad 1e-4: one step; every instance converges; the picked
    instances' iteration counts within 2 of the oracle's and positions within 1e-6 of the bar's
    length."""
    inst = batch256
    picks = pick_instances(inst)
    kw = dict(dt=1.0 / 30.0, grav=GRAV_Z, profile=BENCH_PROFILE, precision=P.FP64, relgrad_tol=1e-4,
              max_iterations=40000)
    say("batch256 rel-grad, instances", picks)
    xs, stats, allst = run_batch(inst, P.default_params(**kw), 1, picks)
    conv = [q.converged for q in allst[0]]
    say("converged", sum(conv), "of", len(conv), "iterations min/max",
        min(q.iterations for q in allst[0]), max(q.iterations for q in allst[0]))
    assert all(c == 1 for c in conv)
    assert all(q.g_norm <= 1e-4 * q.g0_norm for q in allst[0])
    op = O.default_params(dt=1.0 / 30.0, grav=GRAV_Z, profile=BENCH_PROFILE, iterations=40000, relgrad_tol=1e-4)
    for k in picks:
        xo, so = run_oracle(inst[k], op, 1, 64)
        L = float(inst[k][0][:, 2].max())
        err = np.abs(xs[k] - xo).max()
        say("instance %d: iterations gpu %d oracle %d, |dx| %.2e L" % (k, stats[k][0][1], so[0][1], err / L))
        assert so[0][3] == 1
        assert abs(stats[k][0][1] - so[0][1]) <= 2
        assert err <= 1e-6 * L

"""GPU tests of the batched solver (independent meshes stepped together) and of the autotuner:
a batch gives bitwise the results of separate solvers with the same tile size."""
import sys

import numpy as np
import pytest

import meshgen as M
import phybs_amd as P

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def instances(seed=0, k=5):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        nx, ny, nz = (int(v) for v in (rng.integers(2, 6), rng.integers(2, 6), rng.integers(4, 14)))
        rest, tets = M.bar(nx, ny, nz, 0.05)
        mask = M.clamp_x0(rest)
        x0 = M.twisted_state(rest, nx, ny, nz, 0.05, float(rng.uniform(0, 90)), 0.01, 10 + i)
        E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
        nu = float(rng.uniform(0.3, 0.49))
        out.append((rest, tets, mask, x0, E, nu))
    return out


@pytest.mark.parametrize("mode", ["fixed_f32", "relgrad_f64"])
def test_batch_equals_separate_solvers(mode):
    inst = instances()
    kw = dict(tile_tets=64, grav=(0.0, 0.0, -9.8))
    if mode == "relgrad_f64":
        kw.update(precision=P.FP64, relgrad_tol=1e-4, max_iterations=5000)
    say("batch of", len(inst), mode, [len(t) for _, t, _, _, _, _ in inst])
    sep_x, sep_st = [], []
    for rest, tets, mask, x0, E, nu in inst:
        s = P.Solver(rest, tets, mask, P.default_params(youngs=E, nu=nu, **kw))
        s.set_positions(x0)
        st = [s.step() for _ in range(4)]
        sep_x.append(s.positions(f64=True))
        sep_st.append([(q.iterations, q.rollbacks, q.passes) for q in st])
        s.close()
    b = P.Solver.batch([(r, t, m) for r, t, m, _, _, _ in inst], [i[4] for i in inst], [i[5] for i in inst],
                       P.default_params(**kw))
    b.set_positions(np.concatenate([i[3] for i in inst]))
    per = [[] for _ in inst]
    for _ in range(4):
        agg = b.step()
        for i in range(len(inst)):
            q = b.instance_stats(i)
            per[i].append((q.iterations, q.rollbacks, q.passes))
        assert agg.passes == max(p[-1][2] for p in per)
    xb = b.positions(f64=True)
    b.close()
    off = 0
    for i, x in enumerate(sep_x):
        assert per[i] == sep_st[i], (i, per[i], sep_st[i])
        assert np.array_equal(xb[off:off + len(x)].view(np.uint64), x.view(np.uint64)), i
        off += len(x)


def test_ragged_batch_with_a_single_tet_instance_equals_separate_solvers():
    """Instances of 1, 6 and 96 tets side by side (tiles of one instance never mix with
    another's): each instance's trajectory is bitwise that of its own solver."""
    one = (np.array([[0, 0, 0], [0.05, 0, 0], [0, 0.05, 0], [0, 0, 0.05]], np.float32), np.array([[0, 1, 2, 3]], np.int32))
    inst = [one, M.bar(1, 1, 1, 0.05), M.bar(2, 2, 4, 0.05)]
    masks = [M.clamp_x0(r) for r, _ in inst]
    x0s = [r + np.float32(0.002) * np.sin(np.arange(r.size, dtype=np.float32)).reshape(r.shape) * (1 - m[:, None])
           for (r, _), m in zip(inst, masks)]
    kw = dict(tile_tets=64, grav=(0.0, 0.0, -9.8))
    say("ragged batch", [len(t) for _, t in inst])
    sep = []
    for (r, t), m, x0 in zip(inst, masks, x0s):
        s = P.Solver(r, t, m, P.default_params(**kw))
        s.set_positions(x0)
        for _ in range(3):
            s.step()
        sep.append(s.positions())
        s.close()
    b = P.Solver.batch([(r, t, m) for (r, t), m in zip(inst, masks)], [1e5] * 3, [0.45] * 3, P.default_params(**kw))
    b.set_positions(np.concatenate(x0s))
    for _ in range(3):
        b.step()
    xb = b.positions()
    b.close()
    off = 0
    for i, x in enumerate(sep):
        assert np.array_equal(xb[off:off + len(x)].view(np.uint32), x.view(np.uint32)), i
        off += len(x)


def test_batch_mixed_mode_tracks_fp64():
    """Mixed mode on a batch: the switch to FP64 is taken for all instances together (the largest
    |g|/|g0| among the active ones decides); every instance still converges to 1e-4 on FP64 passes.
    Two iterates that both meet the tolerance differ by the tolerance's own slack, which is
    largest on the softest bars (E down to 1e4 Pa): the end positions of every instance agree with
    the FP64 batch's to 1e-3 of that instance's motion over the three steps."""
    inst = instances(seed=5, k=4)
    out = []
    for sw in (0.0, 2e-4):
        kw = dict(tile_tets=64, grav=(0.0, 0.0, -9.8), precision=P.FP64, relgrad_tol=1e-4, max_iterations=5000,
                  mixed_switch=sw)
        b = P.Solver.batch([(r, t, m) for r, t, m, _, _, _ in inst], [i[4] for i in inst], [i[5] for i in inst],
                           P.default_params(**kw))
        b.set_positions(np.concatenate([i[3] for i in inst]))
        conv = []
        for _ in range(3):
            b.step()
            conv.append([b.instance_stats(i) for i in range(len(inst))])
        out.append(b.positions(f64=True))
        b.close()
        for row in conv:
            for q in row:
                assert q.converged == 1 and q.g_norm <= 1e-4 * q.g0_norm
    off = 0
    for r, t, m, x0, E, nu in inst:
        sl = slice(off, off + len(r))
        motion = np.abs(out[0][sl] - x0).max()
        diff = np.abs(out[0][sl] - out[1][sl]).max()
        say("instance E", E, "motion", motion, "mixed vs fp64", diff)
        assert diff <= 1e-3 * motion
        off += len(r)


def test_batch_graph_replay_matches_eager():
    inst = instances(seed=3, k=4)
    res = []
    for gp in (-1, 8):
        b = P.Solver.batch([(r, t, m) for r, t, m, _, _, _ in inst], [i[4] for i in inst], [i[5] for i in inst],
                           P.default_params(tile_tets=64, graph_passes=gp, alpha0=0.3))
        b.set_positions(np.concatenate([i[3] for i in inst]))
        for _ in range(3):
            b.step()
        res.append(b.positions())
        b.close()
    assert np.array_equal(res[0].view(np.uint32), res[1].view(np.uint32))


def test_autotuner_picks_a_timed_tile_size():
    rest, tets = M.bar(8, 8, 32, 0.05)
    mask = M.clamp_x0(rest)
    x0 = M.twisted_state(rest, 8, 8, 32, 0.05, 30.0, 0.01, 1)
    s = P.Solver(rest, tets, mask, P.default_params(tile_tets=-1), nodes=x0)
    rep = s.autotune_report()
    p = s.get_params()
    say("autotuner us/pass", rep, "chosen", p.tile_tets)
    assert len(rep) >= 2 and p.tile_tets in rep and rep[p.tile_tets] == min(rep.values())
    for _ in range(2):
        s.step()
    x_tuned = s.positions()
    s.close()
    t = P.Solver(rest, tets, mask, P.default_params(tile_tets=p.tile_tets), nodes=x0)
    for _ in range(2):
        t.step()
    assert np.array_equal(t.positions().view(np.uint32), x_tuned.view(np.uint32))
    t.close()


@pytest.mark.parametrize("mode", ["fixed_f32", "relgrad_f64"])
def test_batch_with_per_instance_tiles_equals_separate_solvers(mode):
    """Each instance at its own tile size (16 to 512 tets per tile) in one batch: each instance's
    trajectory and loop statistics are bitwise those of its own solver at that tile size."""
    inst = instances(seed=4, k=6)
    tiles = [16, 64, 512, 128, 32, 256]
    kw = dict(grav=(0.0, 0.0, -9.8))
    if mode == "relgrad_f64":
        kw.update(precision=P.FP64, relgrad_tol=1e-4, max_iterations=5000)
    say("per-instance tiles", mode, tiles, [len(t) for _, t, _, _, _, _ in inst])
    sep_x, sep_st = [], []
    for (rest, tets, mask, x0, E, nu), tt in zip(inst, tiles):
        s = P.Solver(rest, tets, mask, P.default_params(youngs=E, nu=nu, tile_tets=tt, **kw))
        s.set_positions(x0)
        st = [s.step() for _ in range(3)]
        sep_x.append(s.positions(f64=True))
        sep_st.append([(q.iterations, q.rollbacks, q.passes) for q in st])
        s.close()
    b = P.Solver.batch([(r, t, m) for r, t, m, _, _, _ in inst], [i[4] for i in inst], [i[5] for i in inst],
                       P.default_params(**kw), tile_tets=tiles)
    assert list(b.instance_tile_tets()) == tiles
    b.set_positions(np.concatenate([i[3] for i in inst]))
    per = [[] for _ in inst]
    for _ in range(3):
        b.step()
        for i in range(len(inst)):
            q = b.instance_stats(i)
            per[i].append((q.iterations, q.rollbacks, q.passes))
    xb = b.positions(f64=True)
    b.close()
    off = 0
    for i, x in enumerate(sep_x):
        assert per[i] == sep_st[i], (i, per[i], sep_st[i])
        assert np.array_equal(xb[off:off + len(x)].view(np.uint64), x.view(np.uint64)), i
        off += len(x)


def test_autotuner_per_instance_tile_sizes():
    """A batch of 12 bars of 96 to 3,456 tets autotuned (tile_tets = -1): every uniform candidate
    and, when the size classes disagree, the per-instance mix are timed; instances of one class
    share a size; the kept configuration is the fastest timed; the tuned batch steps bitwise like
    a batch built with the same per-instance sizes."""
    rng = np.random.default_rng(21)
    meshes, x0s, Es, nus = [], [], [], []
    for i in range(12):
        nx, ny, nz = int(rng.integers(2, 7)), int(rng.integers(2, 7)), int(rng.integers(4, 17))
        rest, tets = M.bar(nx, ny, nz, 0.05)
        meshes.append((rest, tets, M.clamp_x0(rest)))
        x0s.append(M.twisted_state(rest, nx, ny, nz, 0.05, 20.0, 0.01, 30 + i))
        Es.append(float(np.exp(rng.uniform(np.log(1e4), np.log(1e6)))))
        nus.append(float(rng.uniform(0.3, 0.49)))
    x0 = np.concatenate(x0s)
    kw = dict(grav=(0.0, 0.0, -9.8))
    s = P.Solver.batch(meshes, Es, nus, P.default_params(tile_tets=-1, **kw), nodes=x0)
    rep = s.autotune_report()
    tt = s.instance_tile_tets()
    say("autotune", rep, "per instance", list(tt), "tets", [len(m[1]) for m in meshes])
    uniform = {k: v for k, v in rep.items() if k != "per_instance"}
    assert len(uniform) >= 2 and all(k in (64, 128, 256, 512) for k in uniform)
    assert all(t in (64, 128, 256, 512) for t in tt)
    order = np.argsort([len(m[1]) for m in meshes], kind="stable")
    for c in range(4):
        assert len(set(tt[order[3 * c:3 * c + 3]])) == 1
    best_uniform = min(uniform, key=uniform.get)
    if "per_instance" in rep and rep["per_instance"] < uniform[best_uniform]:
        say("kept: per-instance sizes")
    else:
        assert all(t == best_uniform for t in tt)
    for _ in range(2):
        s.step()
    x_tuned = s.positions(f64=True)
    s.close()
    b = P.Solver.batch(meshes, Es, nus, P.default_params(**kw), nodes=x0, tile_tets=list(tt))
    for _ in range(2):
        b.step()
    assert np.array_equal(b.positions(f64=True).view(np.uint64), x_tuned.view(np.uint64))
    b.close()


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_state_snapshot_resumes_a_batch_bitwise(tmp_path, prec):
    """Save (x, v, v_last) of a batch after 2 steps, step 3 more; a new batch on the same meshes
    loads the snapshot and steps 3 more: bitwise the same positions, velocities and loop
    statistics.  A snapshot of other meshes is refused and leaves the state as it was."""
    inst = instances(seed=6, k=4)
    kw = dict(tile_tets=64, grav=(0.0, 0.0, -9.8))
    if prec == "f64":
        kw.update(precision=P.FP64, relgrad_tol=1e-4, max_iterations=5000)
    meshes = [(r, t, m) for r, t, m, _, _, _ in inst]
    Es, nus = [i[4] for i in inst], [i[5] for i in inst]
    x0 = np.concatenate([i[3] for i in inst])
    path = str(tmp_path / "state.bin")
    a = P.Solver.batch(meshes, Es, nus, P.default_params(**kw), nodes=x0)
    for _ in range(2):
        a.step()
    a.save_state(path)
    sa = [(q.iterations, q.rollbacks, q.passes) for q in (a.step() for _ in range(3))]
    xa, va = a.positions(f64=True), a.velocities()
    a.close()
    say("snapshot", prec, "bytes", (tmp_path / "state.bin").stat().st_size)
    b = P.Solver.batch(meshes, Es, nus, P.default_params(**kw), nodes=x0)
    b.load_state(path)
    sb = [(q.iterations, q.rollbacks, q.passes) for q in (b.step() for _ in range(3))]
    xb, vb = b.positions(f64=True), b.velocities()
    assert sa == sb
    assert np.array_equal(xa.view(np.uint64), xb.view(np.uint64))
    for u, w in zip(va, vb):
        assert np.array_equal(u.view(np.uint32), w.view(np.uint32))
    other = P.Solver.batch(meshes[:3], Es[:3], nus[:3], P.default_params(**kw))
    other.save_state(str(tmp_path / "other.bin"))
    other.close()
    with pytest.raises(RuntimeError, match="snapshot"):
        b.load_state(str(tmp_path / "other.bin"))
    assert np.array_equal(b.positions(f64=True).view(np.uint64), xb.view(np.uint64))
    b.close()

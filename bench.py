"""Benchmark: implicit-Euler steps/s of the MI355X descent solver at rel-grad 1e-4.

Default workload (one GPU): the 64x64x256 bar of 6,291,456 tets (Freudenthal split, edge 0.05 m,
neo-Hookean E=1e5 Pa, nu=0.45, rho=1000, gravity -z, clamp x=0), dt = 1/30, iterating each step
until |g_k| <= 1e-4 |g_0| (FP64 state and per-tet arithmetic).  Prints one JSON line.

In rel-grad mode the line also reports the learned warm start against the plain guess: the
network is trained on the GPU on another seeded trajectory of the same bar (and of the
12x12x50 bar, 43,200 tets, for the second comparison) and then steps the benchmark trajectory.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload bar64x64x256|bar64x64x256-stretch|bar12x12x50|bar4x4x16|batch256] [--mode relgrad|fixed]
                  [--no-learned] [--no-cpu-baseline] [--no-implicit] [--no-batch] [--tile-tets N]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (each scaled by the
# same counter's reading for a plain elementwise kernel of known size) on the 64x64x256 bar:
# (kernel, bytes per position component) -> bytes.  These are not measured by this script (a
# counter run serialises every dispatch); the bench line names the profile they come from.  All
# four were measured with 256-tet tiles (FP64 neo-Hookean solvers now default to 224).
TRAFFIC_PROFILE = "profiles/k_tile_k_node_bar64x64x256_relgrad_f64.md"
MEASURED_TRAFFIC = {("k_tile", 8): 237.6e6, ("k_tile", 4): 188.7e6, ("k_node", 8): 202.4e6, ("k_node", 4): 118.5e6}

WORKLOADS = {
    # name: (nx, ny, nz, material, twist_deg, jitter, fixed iterations)
    "bar4x4x16": (4, 4, 16, 0, 0.0, 0.0, 96),
    "bar12x12x50": (12, 12, 50, 0, 90.0, 0.01, 96),
    "bar64x64x256": (64, 64, 256, 0, 0.0, 0.0, 64),
    "bar64x64x256-stretch": (64, 64, 256, 1, 0.0, 0.0, 64),
    # 256 independent bars in one batched solver (sizes and materials from batch_instances)
    "batch256": (0, 0, 0, 0, 0.0, 0.0, 96),
}


def batch_instances(P, count=256, seed=3):
    """Seeded bars nx, ny in [4, 16], nz in [8, 64], E log-uniform in [1e4, 1e6] Pa, nu uniform in
    [0.3, 0.49], each clamped at x = 0 and starting at rest."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nx, ny = int(rng.integers(4, 17)), int(rng.integers(4, 17))
        nz = int(rng.integers(8, 65))
        rest, tets = P.mesh_bar(nx, ny, nz, 0.05)
        mask = (rest[:, 0] == 0).astype(np.int32)
        x0 = rest
        E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
        nu = float(rng.uniform(0.3, 0.49))
        out.append((rest, tets, mask, x0, E, nu))
    return out


def algorithmic_bytes(nt, nn, rsz):
    """Compulsory HBM bytes per launch of the two kernels of a descent pass (rsz = 4 or 8 bytes
    per position component), as the reference's algorithm needs them.
    k_tile: per tet its 4 node indices (int32) and rest data (9 FP32 Dm^-1 entries + FP32 volume)
            = 56 B; per node the position read and the force written (2 x 3 rsz).  (This tile pass
            reads 16 B per tet and recomputes the rest data from the rest positions, so its
            measured traffic is below this figure.)
    k_node: per node the reads of x, x~, x_last and the force (4 x 3 rsz), Hessian diagonal
            (3 x FP32), mass (FP32) and mask (int32), and the write of x_next (3 rsz)."""
    t = 56.0 * nt + 6.0 * rsz * nn
    return {"k_tile": t, "k_tile_mixed": t, "k_node": (15.0 * rsz + 20.0) * nn}


def host_threads():
    """Threads for the CPU baseline: the cores this process may run on, at most 16 (one GPU's
    share of the host)."""
    return min(16, len(os.sched_getaffinity(0)))


def cpu_baseline(rest, tets, mask, material, grav, gpu_fixed_sps=None, relgrad=False):
    """The CPU restatement of the reference's descent path (the oracle; the reference's own CPU
    code cannot be built here) timed on the host on one full step of the 64x64x256 bar's own
    configuration: 64 FP32 descent iterations from rest (Hessian refreshes at iterations 0 and 32
    included), on all host threads; nothing extrapolated.  With relgrad, also one full FP64
    rel-grad 1e-4 step (about 2 minutes on 16 threads, so off by default)."""
    from oracle import pyoracle as O

    cores = host_threads()
    prof = [(1, 0.96), (7, 0.991), (10, 0.9995)]
    p = O.default_params(iterations=64, grav=grav, dt=1.0 / 30.0, profile=prof)
    o = O.Solver(rest, tets, mask, params=p, precision=32, material=material, threads=cores)
    o.set_state(x=rest.astype(np.float64))
    t0 = time.perf_counter()
    st = o.step()
    el = time.perf_counter() - t0
    del o
    out = {"value": 1.0 / el, "unit": "steps/s (64 FP32 descent iterations per step)", "cores": cores, "kind": "port",
           "sample": f"full step: one step of 64 FP32 descent iterations ({st.iterations} run) of the same "
                     f"{len(tets)}-tet bar from rest on {cores} host threads in {el:.1f} s",
           "descent_iterations_per_s": st.iterations / el}
    if gpu_fixed_sps:
        out["gpu_same_workload_steps_per_s"] = gpu_fixed_sps
        out["gpu_over_cpu"] = gpu_fixed_sps * el
    if relgrad:
        p = O.default_params(iterations=40000, relgrad_tol=1e-4, grav=grav, dt=1.0 / 30.0, profile=prof)
        o = O.Solver(rest, tets, mask, params=p, precision=64, material=material, threads=cores)
        o.set_state(x=rest.astype(np.float64))
        t0 = time.perf_counter()
        st = o.step()
        el = time.perf_counter() - t0
        out["relgrad_full_step"] = {"steps_per_s": 1.0 / el, "iterations": st.iterations, "seconds": el,
                                    "converged": st.converged, "precision": "f64"}
    return out


def cpu_baseline_batch(inst, iters, grav, sample=8):
    """The oracle (FP32, OpenMP) on the first `sample` instances, one step of `iters` iterations
    each; the batch step time is extrapolated by tet count."""
    from oracle import pyoracle as O

    cores = host_threads()
    t_sum, tets_sum = 0.0, 0
    for rest, tets, mask, x0, E, nu in inst[:sample]:
        p = O.default_params(iterations=iters, grav=grav, dt=1.0 / 30.0, profile=[(1, 0.96), (7, 0.991), (10, 0.9995)])
        o = O.Solver(rest, tets, mask, E=E, nu=nu, params=p, precision=32, threads=cores)
        o.set_state(x=x0.astype(np.float64))
        t0 = time.perf_counter()
        o.step()
        t_sum += time.perf_counter() - t0
        tets_sum += len(tets)
    total = sum(len(i[1]) for i in inst)
    step_s = t_sum * total / tets_sum
    return {"value": 1.0 / step_s, "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"{sample} of {len(inst)} instances, one {iters}-iteration FP32 step each in {t_sum:.1f} s "
                      f"({tets_sum} tets), scaled by tet count to the whole batch"}


# the learned correction is smoothed by 20 averaging sweeps over the node graph and taken at its
# full length; the gradient test then decides between it and the plain guess
WARM_OPTIONS = dict(sweeps=20, betas=(1.0,))


def impulse(rest, rng, k):
    """Seeded external impulses on 20 random nodes every 10th step of a training trajectory."""
    if k % 10 != 5:
        return None
    ext = np.zeros_like(rest)
    ext[rng.integers(0, len(rest), 20)] = rng.normal(0, 50.0, (20, 3))
    return ext


def collect_trajectory(P, rest, tets, mask, x0, params, steps, seed, stride, capacity):
    """The plain solver steps a seeded trajectory (impulses every 10th step) and keeps (features,
    converged step) pairs of every stride-th free node per step.  Returns the solver."""
    s = P.Solver(rest, tets, mask, params)
    s.set_positions(x0)
    s.learn_collect(capacity, stride)
    rng = np.random.default_rng(seed)
    for k in range(steps):
        s.step(impulse(rest, rng, k))
    return s


def train_pooled_network(P, params, bar_dims, epochs=30, seed=7):
    """The warm start's training set and network: pairs generated by the plain solver on seeded
    trajectories of the 64x64x256 bar, the 12x12x50 bar and the batch's bars, none of them the
    evaluation trajectories (those start from rest without jitter, and from a 90 degree twist with
    seed 11):
      * the 64x64x256 bar from rest under gravity with 0.2 % jitter (seed 7), 30 steps, every 4th
        free node per step;
      * the 12x12x50 bar released from 90 degree twists with 1 % jitter (seeds 2 and 3), 60 steps
        each, every 2nd free node;
      * 6 bars of the 256-bar batch's distribution (seed 5) from rest, 40 steps each.
    All with impulses on 20 random nodes every 10th step.  One network (Adam, 30 epochs) on the
    pooled samples, evaluated on both the 64x64x256 and the 12x12x50 bar.  A second network trained
    the same way on the 64x64x256 samples alone (the benchmark bar's own regime, same held-out
    trajectory) is returned beside it: pooling the smaller bars' samples costs the large bar part
    of its gain.  Returns (pooled network, bar network, info)."""
    t0 = time.perf_counter()
    rest, tets = P.mesh_bar(*bar_dims, 0.05)
    mask = (rest[:, 0] == 0).astype(np.int32)
    x0 = P.mesh_twist(rest, *bar_dims, 0.05, 0.0, 0.002, seed)
    s = collect_trajectory(P, rest, tets, mask, x0, params, 30, seed, 4, 12_000_000)
    counts = {"bar64x64x256": s.learn_samples()}
    progress("warm-start samples of the %dx%dx%d bar collected" % tuple(bar_dims))
    s.learn_fit_scales()
    s.learn_init_model(seed)
    loss_bar = s.learn_train(epochs, batch=8192, lr=2e-3, seed=seed)
    model_bar = s.learn_get_model()

    def pool(r, t, m, x, prm, steps, sd, stride):
        q = collect_trajectory(P, r, t, m, x, prm, steps, sd, stride, 2_000_000)
        f, tg = q.learn_get_samples()
        q.close()
        s.learn_add_samples(f, tg)
        return len(f)

    r1, t1 = P.mesh_bar(12, 12, 50, 0.05)
    m1 = (r1[:, 0] == 0).astype(np.int32)
    counts["bar12x12x50"] = sum(pool(r1, t1, m1, P.mesh_twist(r1, 12, 12, 50, 0.05, 90.0, 0.01, sd), params, 60, sd, 2)
                                for sd in (2, 3))
    n3 = 0
    for i, (r, t, m, x, E, nu) in enumerate(batch_instances(P, count=6, seed=5)):
        prm = P.default_params(**{k: getattr(params, k) for k in ("dt", "precision", "relgrad_tol", "max_iterations")},
                               grav=tuple(params.grav), youngs=E, nu=nu,
                               profile=[(params.profile_k[j], params.profile_rho[j]) for j in range(params.profile_n)])
        n3 += pool(r, t, m, x, prm, 40, 20 + i, 1)
    counts["batch_bars"] = n3
    t1c = time.perf_counter()
    s.learn_fit_scales()
    s.learn_init_model(seed)
    loss = s.learn_train(epochs, batch=8192, lr=2e-3, seed=seed)
    t2 = time.perf_counter()
    model = s.learn_get_model()
    info = {"samples": int(s.learn_samples()), "samples_by_mesh": counts, "epochs": epochs,
            "loss_first_last": [float(loss[0]), float(loss[-1])],
            "bar_network_loss_first_last": [float(loss_bar[0]), float(loss_bar[-1])],
            "collect_and_bar_network_s": t1c - t0, "train_s": t2 - t1c,
            "trajectories": "64x64x256 from rest, 0.2% jitter, seed 7, 30 steps, every 4th node; 12x12x50 from "
                            "90 deg twists, 1% jitter, seeds 2 and 3, 60 steps, every 2nd node; 6 batch bars "
                            "(seed 5) from rest, 40 steps; impulses every 10th step"}
    s.close()
    return model, model_bar, info


def timed_steps(P, rest, tets, mask, x0, params, warmup, steps, model=None):
    s = P.Solver(rest, tets, mask, params)
    s.set_positions(x0)
    if model is not None:
        s.learn_set_model(*model)
        s.learn_set_options(WARM_OPTIONS["sweeps"], WARM_OPTIONS["betas"])
        s.learn_enable(True)
    for _ in range(warmup):
        s.step()
    t0 = time.perf_counter()
    st = [s.step() for _ in range(steps)]
    dt = time.perf_counter() - t0
    s.close()
    return {"steps_per_s": steps / dt, "iterations_per_step": float(np.mean([q.iterations for q in st])),
            "guess_used_fraction": float(np.mean([q.used_learned for q in st])),
            "g_learned_over_g_plain": [q.g_learned / q.g0_norm for q in st] if model is not None else None,
            "converged_steps": int(sum(q.converged for q in st)), "rollbacks": int(sum(q.rollbacks for q in st))}


def batch_report(P, grav, warmup=1, steps=3):
    """The 256-bar batch (batch_instances) in one batched solver at 96 FP32 descent iterations per
    step, tile size picked by the autotuner: steps/s, time per descent pass (the kernels of all
    instances), and the autotuner's timings."""
    inst = batch_instances(P)
    prm = P.default_params(dt=1.0 / 30.0, grav=grav, material=0, iterations=96,
                           profile=[(1, 0.96), (7, 0.991), (10, 0.9995)], tile_tets=-1)
    s = P.Solver.batch([(i[0], i[1], i[2]) for i in inst], [i[4] for i in inst], [i[5] for i in inst], prm)
    s.set_positions(np.concatenate([i[3] for i in inst]))
    for _ in range(warmup):
        s.step()
    t0 = time.perf_counter()
    st = [s.step() for _ in range(steps)]
    el = time.perf_counter() - t0
    out = {"steps_per_s": steps / el, "us_per_pass": 1e6 * el / max(sum(q.passes for q in st), 1),
           "instances": len(inst), "tets": int(sum(len(i[1]) for i in inst)),
           "nodes": int(sum(len(i[0]) for i in inst)), "tile_tets": s.get_params().tile_tets,
           "autotune_us_per_pass": s.autotune_report(), "steps": steps, "warmup": warmup}
    s.close()
    return out


def progress(msg):
    """One line per finished part of the run, on stderr (stdout carries only the JSON line)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def warm_start_report(P, args, params, rest, tets, mask, x0, dims, plain_sps, plain_its, other=None):
    """Learned warm start vs plain with the pooled network (train_pooled_network):
    on the benchmark bar (the timed steps' own protocol: warm-up, then the timed steps) and on the
    12x12x50 bar released from a 90 degree twist (seed 11; 1 warm-up step, 30 timed).  other =
    (params, plain result) of the other arithmetic (mixed / FP64 throughout): the same network."""
    out = {}
    model, model_bar, info = train_pooled_network(P, params, dims)
    out["training"] = info
    progress("warm-start networks trained")
    lw = timed_steps(P, rest, tets, mask, x0, params, args.warmup, args.steps, model)
    out["bench_bar"] = {"plain_steps_per_s": plain_sps, "plain_iterations_per_step": plain_its,
                        "learned": lw, "speedup": lw["steps_per_s"] / plain_sps}
    lb = timed_steps(P, rest, tets, mask, x0, params, args.warmup, args.steps, model_bar)
    out["bench_bar_bar_network"] = {"learned": lb, "speedup": lb["steps_per_s"] / plain_sps}
    if other is not None:
        op, oplain = other
        lo = timed_steps(P, rest, tets, mask, x0, op, args.warmup, args.steps, model)
        out["bench_bar_other_arithmetic"] = {"mixed_switch": op.mixed_switch, "plain_steps_per_s": oplain["steps_per_s"],
                                             "learned": lo, "speedup": lo["steps_per_s"] / oplain["steps_per_s"]}
    d1 = (12, 12, 50)
    r1, t1 = P.mesh_bar(*d1, 0.05)
    m1 = (r1[:, 0] == 0).astype(np.int32)
    x1 = P.mesh_twist(r1, *d1, 0.05, 90.0, 0.01, 11)
    p1 = timed_steps(P, r1, t1, m1, x1, params, 1, 30)
    l1 = timed_steps(P, r1, t1, m1, x1, params, 1, 30, model)
    out["bar_12x12x50"] = {"tets": len(t1), "plain": p1, "learned": l1, "speedup": l1["steps_per_s"] / p1["steps_per_s"]}
    out["options"] = {"smoothing_sweeps": WARM_OPTIONS["sweeps"], "steps_tried": list(WARM_OPTIONS["betas"])}
    return out


def fixed_iteration_report(P, rest, tets, mask, grav, steps=10):
    """The 6.3M-tet bar at 64 FP32 descent iterations per step from rest, for both materials
    (neo-Hookean and the stretch-based energy): steps/s, descent iterations/s and the force kernel
    against HBM bandwidth."""
    out = {}
    nt, nn = len(tets), len(rest)
    for name, mat in (("neo_hookean", 0), ("stretch", 1)):
        prm = P.default_params(dt=1.0 / 30.0, grav=grav, material=mat, iterations=64,
                               profile=[(1, 0.96), (7, 0.991), (10, 0.9995)])
        s = P.Solver(rest, tets, mask, prm)
        s.step()
        s.set_profiling(8)
        s.reset_kernel_times()
        t0 = time.perf_counter()
        st = [s.step() for _ in range(steps)]
        el = time.perf_counter() - t0
        kt = s.kernel_times()
        s.close()
        alg = algorithmic_bytes(nt, nn, 4)
        ach = alg["k_tile"] / (kt["k_tile"][0] * 1e-3) / 1e9
        out[name] = {"steps_per_s": steps / el, "descent_iterations_per_s": sum(q.iterations for q in st) / el,
                     "kernel_avg_ms": {k: kt[k][0] for k in ("k_tile", "k_node", "k_control")},
                     "k_tile_achieved_gbs": ach, "k_tile_frac": ach / HBM_PEAK_GBS}
    return out


def implicit_report(P, rest, tets, mask, material, grav, steps=3):
    """Backward-Euler steps (the reference's GPU implicit path) on the benchmark bar in FP32 from
    rest under gravity: steps/s, CG iterations, and the SpMV kernel against HBM bandwidth.  SpMV
    algorithmic bytes per launch: every stored 3x3 block once (9 FP32 values + a 4-byte column
    index) plus 2 vectors of 3 doubles per node (p read, A p written)."""
    prm = P.default_params(dt=1.0 / 30.0, grav=grav, material=material)
    s = P.Solver(rest, tets, mask, prm)
    ip = P.implicit_params()
    s.implicit_step(params=ip)  # warm-up: builds the matrix layout and the CG graphs
    s.set_profiling(1)
    s.reset_kernel_times()
    t0 = time.perf_counter()
    st = [s.implicit_step(params=ip) for _ in range(steps)]
    el = time.perf_counter() - t0
    kt = s.kernel_times()
    s.close()
    t32 = np.ascontiguousarray(tets, np.int32)
    nblocks = P.lib().pbs_csr_pattern(len(rest), len(t32), t32.ctypes.data, None, None, None, None) // 9
    nn = len(rest)
    spmv_bytes = nblocks * (9 * 4 + 4) + 2 * 24 * nn
    upd_bytes = 8 * 24 * nn
    spmv_ms, upd_ms = kt["k_ie_spmv"][0], kt["k_ie_update"][0]
    ach = spmv_bytes / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else 0.0
    return {
        "steps": steps, "steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps, "dtype": "f32 matrix, f64 CG",
        "cg_tolerance": ip.cg_tolerance, "cg_iterations_per_step": float(np.mean([q.cg_iterations for q in st])),
        "converged_steps": int(sum(q.converged for q in st)),
        "kernel_avg_ms": {k: v[0] for k, v in kt.items() if k.startswith("k_ie")},
        "spmv_roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": spmv_bytes,
                          "blocks": nblocks},
        "update_achieved_gbs": upd_bytes / (upd_ms * 1e-3) / 1e9 if upd_ms > 0 else 0.0,
    }


def pd_report(P, rest, tets, mask, grav, steps=2):
    """Projective-dynamics steps (5 local/global rounds each) on the benchmark bar in FP32 from rest
    under gravity: steps/s, CG iterations and the local step against HBM bandwidth.  Local-step
    algorithmic bytes per launch: 56 per tet (ids, Dm^-1, volume) + 72 per node (the FP64 iterate
    and momentum term read, the right-hand side written)."""
    prm = P.default_params(dt=1.0 / 30.0, grav=grav)
    s = P.Solver(rest, tets, mask, prm)
    pp = P.pd_params()
    s.pd_step(params=pp)  # warm-up: matrix layout and graphs
    s.set_profiling(1)
    s.reset_kernel_times()
    t0 = time.perf_counter()
    st = [s.pd_step(params=pp) for _ in range(steps)]
    el = time.perf_counter() - t0
    kt = s.kernel_times()
    s.close()
    nt, nn = len(tets), len(rest)
    loc_bytes = 56 * nt + 72 * nn
    loc_ms, spmv_ms = kt["k_pd_local"][0], kt["k_ie_spmv"][0]
    t32 = np.ascontiguousarray(tets, np.int32)
    nblocks = P.lib().pbs_csr_pattern(nn, len(t32), t32.ctypes.data, None, None, None, None) // 9
    spmv_bytes = nblocks * 8 + 2 * 24 * nn  # one FP32 value + column per block, p read, A p written
    return {
        "steps": steps, "steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps, "dtype": "f32 matrix, f64 CG",
        "rounds_per_step": pp.iterations, "cg_tolerance": pp.cg_tolerance,
        "cg_iterations_per_step": float(np.mean([q.cg_iterations for q in st])),
        "converged_steps": int(sum(q.converged for q in st)),
        "kernel_avg_ms": {"k_pd_local": loc_ms, "k_ie_spmv (scalar)": spmv_ms, "k_ie_update": kt["k_ie_update"][0]},
        "local_step_achieved_gbs": loc_bytes / (loc_ms * 1e-3) / 1e9 if loc_ms > 0 else 0.0,
        "spmv_achieved_gbs": spmv_bytes / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else 0.0,
    }


def implicit_cpu_baseline(P, grav):
    """CPU reference point for one backward-Euler step on a 12x12x50 bar (43,200 tets): the
    oracle's FP64 system assembly plus scipy's sparse direct solve (the reference factorises with
    a sparse Cholesky), timed on the host."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spl

    from oracle import pyoracle as O
    rest, tets = P.mesh_bar(12, 12, 50, 0.05)
    mask = (rest[:, 0] == 0).astype(np.int32)
    mu, lam = O.lame(1e5, 0.45)
    pre = O.precompute(rest, tets, mask, 1000.0, grav, 1e8)
    x = rest.astype(np.float64)
    v = np.zeros_like(x)
    t0 = time.perf_counter()
    rowptr, col, _, aval, b = O.implicit_system(rest, tets, mask, pre, mu, lam, x, v, 1.0 / 30.0)
    n3 = 3 * len(rest)
    spl.spsolve(sps.csr_matrix((aval, col, rowptr), shape=(n3, n3)).tocsc(), b)
    el = time.perf_counter() - t0
    return {"value": 1.0 / el, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": "one backward-Euler step of a 12x12x50 bar (43,200 tets): oracle FP64 assembly + "
                      "scipy sparse direct solve"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="bar64x64x256", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="relgrad", choices=["relgrad", "fixed"])
    ap.add_argument("--precision", default="auto", choices=["auto", "f32", "f64"],
                    help="state and per-tet arithmetic (auto: f64 for rel-grad, f32 for fixed iterations)")
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--mixed-switch", type=float, default=0.0,
                    help="FP64 rel-grad: FP32 per-tet arithmetic while |g|/|g0| is above this, FP64 below "
                         "(0: FP64 throughout)")
    ap.add_argument("--max-iterations", type=int, default=40000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-relgrad", action="store_true",
                    help="also time one full FP64 rel-grad step of the CPU path (minutes)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start comparison")
    ap.add_argument("--no-implicit", action="store_true", help="skip the backward-Euler report")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-bar batch report of the default run")
    ap.add_argument("--tile-tets", type=int, default=None,
                    help="tile size of the measured solver (default: from the mesh size; the autotuner for batch256)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read + write) for counter calibration")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.empty(n, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        a.fill_(1.0)  # elementwise fill kernel: writes 4n bytes
        b.fill_(2.0)
        c = a + b  # elementwise add kernel: reads 8n bytes, writes 4n bytes
        torch.cuda.synchronize()
        # the add kernel's streaming rate (HIP events, best of 20), the practical ceiling for
        # kernels that read and write distinct arrays once
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        best = float("inf")
        for _ in range(20):
            e0.record()
            torch.add(a, b, out=c)
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        print(json.dumps({"fill_bytes_written": 4 * n, "add_bytes_read": 8 * n, "add_bytes_written": 4 * n,
                          "add_best_ms": best, "add_gbs": 12 * n / (best * 1e-3) / 1e9, "check": float(c[0])}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the descent solver is a single-GPU path; "
              "multi-GPU runs are not supported.", file=sys.stderr)
        sys.exit(2)

    import phybs_amd as P

    nx, ny, nz, material, twist, jitter, fixed_iters = WORKLOADS[args.workload]
    edge = 0.05
    batch = args.workload == "batch256"
    if batch:
        inst = batch_instances(P)
        rest = np.concatenate([i[0] for i in inst])
        tets = np.concatenate([i[1] for i in inst])
        mask = np.concatenate([i[2] for i in inst])
        x0 = np.concatenate([i[3] for i in inst])
    else:
        rest, tets = P.mesh_bar(nx, ny, nz, edge)
        mask = (rest[:, 0] == 0).astype(np.int32)
        x0 = P.mesh_twist(rest, nx, ny, nz, edge, twist, jitter, 1) if (twist or jitter) else rest
    grav = (0.0, 0.0, -9.8)
    # the 256-bar batch is defined with 96 FP32 iterations per step; in rel-grad mode most tiles
    # of a pass would belong to finished instances and the per-launch roofline would be meaningless
    relg = args.mode == "relgrad" and not batch
    f64 = args.precision == "f64" or (args.precision == "auto" and relg)
    pkw = dict(dt=1.0 / 30.0, grav=grav, material=material, precision=P.FP64 if f64 else P.FP32,
               iterations=fixed_iters, relgrad_tol=args.tol if relg else 0.0, max_iterations=args.max_iterations,
               profile=[(1, 0.96), (7, 0.991), (10, 0.9995)],
               tile_tets=args.tile_tets if args.tile_tets is not None else (-1 if batch else 0))
    params = P.default_params(**pkw, mixed_switch=args.mixed_switch if (relg and f64) else 0.0)
    if batch:
        s = P.Solver.batch([(i[0], i[1], i[2]) for i in inst], [i[4] for i in inst], [i[5] for i in inst], params)
    else:
        s = P.Solver(rest, tets, mask, params)
    s.set_positions(x0)
    for _ in range(args.warmup):
        s.step()
    s.set_profiling(8)  # HIP events around the kernels of every 8th graph replay (32 passes each)
    s.reset_kernel_times()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(s.step())
    t1 = time.perf_counter()
    elapsed = t1 - t0
    kt = s.kernel_times()
    tile_used = s.get_params().tile_tets
    iters = sum(st.iterations for st in stats)
    passes = sum(st.passes for st in stats)
    nt, nn = len(tets), len(rest)
    rsz = 8 if f64 else 4
    alg = algorithmic_bytes(nt, nn, rsz)
    dom = max((k for k in ("k_tile", "k_tile_mixed", "k_node") if k in kt), key=lambda k: kt[k][0] * kt[k][1])
    avg_ms = kt[dom][0]
    achieved = alg[dom] / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    steps_per_s = args.steps / elapsed
    out = {
        # rel-grad mode: value = the plain descent's steps/s; the learned warm start's steps/s and
        # speed-ups (6.3M and 43K tets) are the learned_* fields, GB/s the roofline objects
        "metric": ("steps/s at rel-grad 1e-4, 43K & 6.3M tets: learned warm-start vs plain; GB/s" if relg
                   else "steps/s (fixed iterations)"),
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64 state, f32 per-tet arithmetic above the switch" if params.mixed_switch > 0 else "f64")
                 if f64 else "f32",
        "data": "synthetic",
        "config": {"workload": (f"256 bars (nx, ny 4-16, nz 8-64; E 1e4-1e6 Pa; nu 0.3-0.49) in one batched "
                                f"solver, {nt} tets, {nn} nodes, " if batch else
                                f"{args.workload}: {nx}x{ny}x{nz} bar, {nt} tets, {nn} nodes, ")
                               + f"{'stretch' if material else 'neo-Hookean'}, "
                               + (f"rel-grad {args.tol:g}" if relg else f"{fixed_iters} iterations/step"),
                   "dt": 1.0 / 30.0, "tets": nt, "nodes": nn, "mixed_switch": params.mixed_switch},
        "iterations_per_step": iters / args.steps,
        "descent_iterations_per_s": iters / elapsed,
        # whole descent pass against its compulsory bytes, 56 T + (12 r + 20) N with r bytes per
        # position component: rest data and connectivity per tet; x, x~, x_last read, x_next
        # written, Jacobi diagonal, mass and mask per node (the per-node forces k_tile hands to
        # k_node are this design's staging, not compulsory)
        "pass_roofline": {"bytes_per_pass": 56.0 * nt + (12.0 * rsz + 20.0) * nn,
                          "us_per_pass": 1e6 * elapsed / max(passes, 1),
                          "achieved_gbs": (56.0 * nt + (12.0 * rsz + 20.0) * nn) * max(passes, 1) / elapsed / 1e9,
                          "frac": (56.0 * nt + (12.0 * rsz + 20.0) * nn) * max(passes, 1) / elapsed / 1e9 / HBM_PEAK_GBS},
        "rollbacks": sum(st.rollbacks for st in stats),
        "converged_steps": sum(st.converged for st in stats),
        "rel_grad_final": [st.g_norm / st.g0_norm if st.g0_norm > 0 else None for st in stats],
        "kernel_avg_ms": {k: v[0] for k, v in kt.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC.get((dom, rsz)) if args.workload == "bar64x64x256" else None,
                     "traffic_source": (TRAFFIC_PROFILE + " (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE with 256-tet tiles; "
                                        "not this run, tiles of %d tets)" % tile_used)
                     if args.workload == "bar64x64x256" and (dom, rsz) in MEASURED_TRAFFIC else None,
                     "kernel": dom,
                     "algorithmic_bytes_per_launch": alg[dom]},
        "learned_warm_start": None,
    }
    if batch:
        out["autotune_us_per_pass"] = s.autotune_report() if params.tile_tets < 0 else None
        out["tile_tets"] = s.get_params().tile_tets
    s.close()
    progress("timed steps done")
    if relg and f64 and not batch:
        # the same steps with the other arithmetic: mixed mode (FP32 per-tet arithmetic down to
        # |g|/|g0| = 2e-4) when the main run was FP64 throughout, and the reverse
        sw = 0.0 if params.mixed_switch > 0 else 2e-4
        other_params = P.default_params(**pkw, mixed_switch=sw)
        r = timed_steps(P, rest, tets, mask, x0, other_params, args.warmup, args.steps)
        r["mixed_switch"] = sw
        out["other_arithmetic"] = r
        progress("other arithmetic done")
    if relg and not args.no_learned and not batch:
        other = (other_params, out["other_arithmetic"]) if f64 else None
        lw = warm_start_report(P, args, params, rest, tets, mask, x0, (nx, ny, nz), steps_per_s,
                               iters / args.steps, other)
        out["learned_warm_start"] = lw
        out["learned_steps_per_s"] = lw["bench_bar"]["learned"]["steps_per_s"]
        out["learned_speedup"] = lw["bench_bar"]["speedup"]
        out["learned_speedup_43k"] = lw["bar_12x12x50"]["speedup"]
        out["learned_speedup_bar_network"] = lw["bench_bar_bar_network"]["speedup"]
        out["learned_guess_used_fraction"] = lw["bench_bar"]["learned"]["guess_used_fraction"]
        progress("warm start done")
    if not batch and args.workload == "bar64x64x256":
        out["fixed_64_iterations_f32"] = fixed_iteration_report(P, rest, tets, mask, grav)
        progress("fixed-iteration runs done")
        if not args.no_batch:
            out["batch256"] = batch_report(P, grav)
            progress("batch done")
    if not args.no_implicit and not batch:
        out["implicit_euler"] = implicit_report(P, rest, tets, mask, material, grav)
        out["projective_dynamics"] = pd_report(P, rest, tets, mask, grav)
        progress("implicit and projective-dynamics steps done")
        if not args.no_cpu_baseline:
            out["implicit_euler"]["cpu_baseline"] = implicit_cpu_baseline(P, grav)
    if not args.no_cpu_baseline and batch:
        out["cpu_baseline"] = cpu_baseline_batch(inst, fixed_iters, grav)
    elif not args.no_cpu_baseline:
        fixed = out.get("fixed_64_iterations_f32", {}).get("neo_hookean" if material == 0 else "stretch", {})
        out["cpu_baseline"] = cpu_baseline(rest, tets, mask, material, grav, fixed.get("steps_per_s"),
                                           relgrad=args.cpu_relgrad)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

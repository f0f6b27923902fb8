"""Benchmark of the implicit cloth step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload large|plain]

Workload (default): one 512x512-vertex cloth over a 4x4 square, positions
perturbed by N(0, 1e-3*spacing) from seed 0, top edge pinned, ks 50, kd 0.25,
mass 1, gravity -9.81, dt 1/120, CG fixed at 30 iterations, synthetic data.
A step is one implicit step (assembly + 30 CG iterations + update) with the
state resident in HBM.  Prints one JSON line:

  value        vertex-steps/s = N * K / (time of K steps), HIP events on the
               stream the kernels run on
  roofline     for the dominant kernel (cg_spmv): algorithmic bytes per
               launch (88N + 32P, the spmv share of the per-iteration model
               160N + 32P) / its average launch time inside the step's launch
               sequence (an event attached to every dispatch of the same
               steps, oc_cloth_profile_steps), against 8 TB/s; the same over
               300 back-to-back launches beside it; traffic = HBM bytes per
               launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
               summary, and its fraction beside the model's
  config       the kernel split per step, and the same workload with the
               single-reduction CG
  batch        4096 uneven cloths, 200 steps, one graph-replayed launch per
               step, against one-cloth-per-launch forms; the roofline of the
               launch after the timed steps
  batch_adjoint, adjoint
               batched (512 cloths, 60 steps) and single-cloth (128x128, 120
               steps) forward rollouts and their adjoints
  entries      Verlet, PBD and grid Verlet at 512x512, the IE demo's
               StepPhysics at 21x21
  cpu_baseline the FP32 CPU restatement of the same step (oracle) on all
               host threads (OpenMP) and on one, each on a bounded sample of
               the same workload
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "vertex-steps/s implicit-CG step + % HBM roofline; batched fwd+adjoint vs plain"


def spmv_bytes(N: int, P: int) -> int:
    """Algorithmic bytes of one cg_spmv launch: symmetric blocks 24N+24P,
    topology 8P+4N, d in / q out 24N, direction update d = r + beta d 36N."""
    return 88 * N + 32 * P


def step_bytes(N: int, S: int, P: int, K: int) -> int:
    """Algorithmic bytes of one step: (61N + 12S + 32P) + K(160N + 32P) + 60N."""
    return (61 * N + 12 * S + 32 * P) + K * (160 * N + 32 * P) + 60 * N


def read_profile_value(path: str, key: str):
    """A `key = value` line of a committed profile summary, or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key):
                    return float(line.split("=")[1].split()[0])
    except OSError:
        pass
    return None


def copy_probe(mib: int) -> int:
    """One device-to-device copy of mib MiB (hipMemcpy, run by the HIP
    runtime as a copy kernel that reads and writes mib MiB once each);
    returns the bytes copied."""
    from opencloth_amd.hiptime import device_copy

    return device_copy(mib << 20)


def bench_entries(steps: int = 100) -> dict:
    """The other entry points on 512x512 cloths (HIP events around `steps`
    steps after 5 warm-up steps), with simple algorithmic byte models:
      verlet  the Verlet demo's constants on the 512x512 grid (perturbed,
              top edge pinned, damping -0.0125, ground clamp), one fused
              forces + integration launch per step; bytes 36N + 12S (x and
              x_last read, next positions written, spring constants once)
      pbd     the PBD demo's constraints and constants on a 512x512 grid
              (mass 1/N, 2 solver iterations), one launch per colour; bytes
              ns(72 D + 108 B) + 132N (distance: two particles and inverse
              masses read, indices, rest length and k' read, two particles
              written; bending likewise with three; prediction, damping sums
              and integration 132N)
      verlet_grid  the GPU demos' Verlet kernel (12-neighbour stencil, rest
              lengths from the spacing) on a 512x512 grid with the demo's
              constants, one launch per step; bytes 49N (x and x_last read,
              both written, one pin byte)
      implicit_step_physics  the implicit demo's full StepPhysics (21x21,
              step + collider + Provot), launch-bound."""
    from opencloth_amd import scenes
    from opencloth_amd.cloth import ImplicitCloth, PBDCloth, VerletCloth, VerletGrid
    from opencloth_amd.hiptime import Event

    def timed(run, n):
        run(5)
        e0, e1 = Event(), Event()
        e0.record()
        run(n)
        e1.record()
        e1.synchronize()
        return e0.elapsed_ms(e1) / n

    out = {}
    n = 512
    X0 = scenes.grid_positions(n, n, 4.0, 5.0)
    springs = scenes.grid_springs(n, n, X0, ks=(50.75, 50.75, 50.95), kd=(-0.25, -0.25, -0.25))
    X = scenes.perturb(X0, 1e-3 * 4.0 / (n - 1), 0)
    pinned = np.zeros(n * n, np.uint8)
    pinned[:n] = 1
    vc = VerletCloth(n, n, X, springs, pinned, mass=1.0, gravity=(0.0, -0.00981, 0.0),
                     damping=-0.0125, dt=1.0 / 60.0, ground_clamp=True)
    ms = timed(lambda k: vc.step(k), steps)
    N, S = n * n, len(springs)
    b = 36 * N + 12 * S
    out["verlet"] = {"workload": f"{n}x{n} Verlet step, {S} springs", "ms_per_step": ms,
                     "vertex_steps_per_s": N / (ms * 1e-3), "algorithmic_bytes": b,
                     "roofline_frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    vc.close()

    sc = scenes.pbd_large_scene(n)
    pc = PBDCloth.from_scene(sc)
    nd_col, nb_col = pc.num_colors()
    ms = timed(lambda k: pc.step(k), steps)
    ns = sc.params.solver_iterations
    b = ns * (72 * len(sc.distance) + 108 * len(sc.bending)) + 132 * N
    out["pbd"] = {"workload": f"{n}x{n} PBD step, {len(sc.distance)} distance and {len(sc.bending)} "
                              f"bending constraints, {ns} solver iterations",
                  "ms_per_step": ms, "vertex_steps_per_s": N / (ms * 1e-3),
                  "colours": [nd_col, nb_col],
                  "launches_per_step": 5 + ns * (nd_col + nb_col + int(sc.params.ground)) + 1,
                  "algorithmic_bytes": b, "roofline_frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    pc.close()

    sc = scenes.vgrid_large_scene(n)
    vg = VerletGrid.from_scene(sc)
    ms = timed(lambda k: vg.step(k), steps)
    b = 49 * N
    out["verlet_grid"] = {"workload": f"{n}x{n} grid Verlet step (12-neighbour stencil)",
                          "ms_per_step": ms, "vertex_steps_per_s": N / (ms * 1e-3),
                          "algorithmic_bytes": b, "roofline_frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    vg.close()

    sc = scenes.demo_scene()
    cl = ImplicitCloth.from_scene(sc)
    ell, slack = sc.extra["ellipsoid"], sc.extra["provot_slack"]
    ms = timed(lambda k: cl.step_physics(k, ell, slack), steps)
    cl.close()
    # the same with the CG in one launch that stops at convergence (same bits)
    p = dataclasses.replace(sc.params, cg_persistent=True)
    cl = ImplicitCloth(sc.nx, sc.ny, sc.X, sc.springs, sc.pinned, p, sc.V)
    ms_p = timed(lambda k: cl.step_physics(k, ell, slack), steps)
    cl.close()
    out["implicit_step_physics"] = {"workload": "21x21 implicit demo StepPhysics (step, collider, Provot)",
                                    "ms_per_step": ms, "ms_per_step_persistent_cg": ms_p}
    return out


def cpu_threads() -> int:
    """Host threads for the CPU baseline: the affinity mask, capped by
    OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() else n


def cpu_baseline(sc, budget_s: float = 12.0, single_budget_s: float = 6.0) -> dict:
    """FP32 CPU restatement timed on whole steps of the same workload: on all
    host threads (OpenMP loops over vertices, springs and tiles, tile-tree
    dot products) for about budget_s seconds, and on one thread with
    sequential dot products, as the reference sums (IE:219-225), for about
    single_budget_s seconds."""
    from oracle import Oracle, make_params

    o = Oracle(32)
    sysm = o.system(sc.n_vertices, sc.springs, sc.pinned)

    def run(threads, prm, budget):
        """3 warm-up steps, then timed steps until `budget` seconds; the
        rate is taken at the median step time."""
        o.set_threads(threads)
        X = o.arr(sc.X).copy(); V = o.arr(sc.V).copy(); DV = np.zeros_like(X)
        for _ in range(3):
            sysm.step(X, V, DV, prm)
        times = []
        t0 = time.perf_counter()
        while True:
            t1 = time.perf_counter()
            sysm.step(X, V, DV, prm)
            times.append(time.perf_counter() - t1)
            el = time.perf_counter() - t0
            if el >= budget or len(times) >= 400:
                break
        o.set_threads(1)
        return sc.n_vertices / float(np.median(times)), len(times), el

    from opencloth_amd._lib import tile_shape

    nth = cpu_threads()
    tx, ty = tile_shape()
    v_mt, n_mt, el_mt = run(nth, make_params(sc.params, nx=sc.nx, ny=sc.ny, tx=tx, ty=ty, dot_mode=1),
                            budget_s)
    v_st, n_st, el_st = run(1, make_params(sc.params), single_budget_s)

    # the reference's own IntegrateImplicit, bug for bug (spring-indexed
    # block-diagonal A, IE:684-700), on the same inputs, one thread
    o.set_threads(1)
    X = o.arr(sc.X).copy(); V = o.arr(sc.V).copy(); Vn = np.zeros_like(X)
    p = sc.params
    n_lit, t0 = 0, time.perf_counter()
    while True:
        sysm.ie_literal_step(X, V, Vn, p.mass, p.gravity, p.dt, i_max=10, eps=1e-3)
        n_lit += 1
        el_lit = time.perf_counter() - t0
        if el_lit >= 2.0 or n_lit >= 100:
            break
    return {
        "value": v_mt,
        "unit": "vertex-steps/s",
        "cores": nth,
        "kind": "port",
        "sample": f"median of {n_mt} timed steps of the same {sc.nx}x{sc.ny} workload after 3 warm-up "
                  f"steps on {nth} threads (FP32 restatement, OpenMP, tile-tree dots, {el_mt:.1f} s); "
                  f"single thread with sequential dots: median of {n_st} steps, {el_st:.1f} s",
        "single_thread_value": v_st,
        "reference_literal": {
            "value": sc.n_vertices * n_lit / el_lit,
            "unit": "vertex-steps/s",
            "cores": 1,
            "sample": f"{n_lit} steps of the reference's own ComputeForces + IntegrateImplicit "
                      f"(IE:645-700) restated bug for bug (spring-indexed block-diagonal A, CG 10 "
                      f"iterations, EPS 1e-3) on the same {sc.nx}x{sc.ny} inputs, {el_lit:.1f} s; "
                      f"a different (incorrect) linear system, for the record",
        },
    }


def bench_batch(steps: int, warmup: int, n_cloths: int = 4096, plain_steps: int = 1, plain: bool = True,
                schedule=None) -> dict:
    """Batched forward step: n_cloths cloths with sides uniform in [16,128],
    ks log-uniform [20,200], kd uniform [0.1,0.5], 2-8 random top-edge pins,
    CG <= 40 iterations with per-cloth relative tolerance 1e-4, dt 1/60, one
    persistent launch per step replayed from a HIP graph.  The plain form
    steps the same cloths one cloth per launch sequence (oc_cloth_step)."""
    from opencloth_amd import scenes
    from opencloth_amd.batch import ClothBatch
    from opencloth_amd.cloth import ImplicitCloth
    from opencloth_amd.hiptime import Event

    t0 = time.perf_counter()
    sc = scenes.batch_scene(n_cloths=n_cloths)
    t1 = time.perf_counter()
    b = ClothBatch.from_scene(sc)
    t2 = time.perf_counter()
    # the schedule is tuned on the state two steps in (where the timed steps
    # run), then the cloths go back to rest
    b.step(2)
    if schedule:  # a fixed schedule (reproducible counter runs) instead of the autotuner
        b.set_schedule(*schedule)
        sched = schedule
    else:
        sched = b.autotune()
    b.set_state(sc.X, sc.V)
    b.step(warmup)
    b.synchronize()
    e0, e1 = Event(), Event()
    e0.record()
    b.step(steps)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_ms(e1)
    # the roofline launch: one more step timed alone; its algorithmic bytes
    # come from its own CG iteration counts (the committed counters are of
    # this launch: the batch leg's last k_batch_step dispatch)
    r0, r1 = Event(), Event()
    r0.record()
    b.step(1)
    r1.record()
    r1.synchronize()
    launch_s = r0.elapsed_ms(r1) * 1e-3
    it, st = b.status()
    b.close()
    N = int(sc.X.shape[0])
    # algorithmic bytes of the roofline launch: the single-cloth model per
    # cloth with its own CG iteration count (P = S - (nx + ny): one duplicate
    # bend per grid row and column)
    nxs, nys, ss = sc.nx.astype(np.int64), sc.ny.astype(np.int64), sc.n_springs.astype(np.int64)
    ns = nxs * nys
    ps = ss - (nxs + nys)
    step_b = int(np.sum(61 * ns + 12 * ss + 32 * ps + it.astype(np.int64) * (160 * ns + 32 * ps) + 60 * ns))

    def plain_form(persistent: bool) -> float:
        """ms per step of the same cloths stepped one cloth per launch
        sequence (oc_cloth_step); persistent: each solve is one cooperative
        launch that stops at convergence, so no CG launch is issued after a
        cloth has converged."""
        cls_ = []
        for c in range(sc.n_cloths):
            cs = sc.cloth(c)
            p = dataclasses.replace(cs.params, cg_persistent=persistent)
            cls_.append(ImplicitCloth(cs.nx, cs.ny, cs.X, cs.springs, cs.pinned, p, cs.V))
        for cl in cls_:
            cl.step(1)
        for cl in cls_:
            cl.synchronize()
        p0, p1 = Event(), Event()
        p0.record()
        for _ in range(plain_steps):
            for cl in cls_:
                cl.step(1)
        p1.record()
        p1.synchronize()
        for cl in cls_:
            cl.close()
        return p0.elapsed_ms(p1) / plain_steps

    ms_plain = plain_form(False) if plain else float("nan")
    ms_plain_p = plain_form(True) if plain else float("nan")
    traffic = read_profile_value(os.path.join(ROOT, "profiles", "batch_step_4096_cloths_pmc.txt"),
                                 "batch_step_hbm_bytes_per_launch")
    return {
        "workload": f"{sc.n_cloths} cloths, sides U[16,128], {N} vertices, {len(sc.springs)} springs, "
                    "CG <= 40 it, rel tol 1e-4, dt 1/60",
        "vertex_steps_per_s": N * steps / (ms * 1e-3),
        "ms_per_step": ms / steps,
        "steps": steps,
        "schedule": {"block_threads": sched[0], "vertices_per_thread": sched[1]},
        "cg_iters_mean": float(it.mean()), "cg_iters_max": int(it.max()),
        "status_nonzero": int((st != 0).sum()),
        "step_algorithmic_bytes": step_b,
        "step_roofline_frac": step_b / launch_s / 1e9 / HBM_PEAK_GBS,
        "roofline": {
            "bound": "hbm", "kernel": "k_batch_step (forward), the step after the timed steps",
            "achieved": step_b / launch_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": step_b / launch_s / 1e9 / HBM_PEAK_GBS,
            "algorithmic_bytes_per_launch": step_b, "avg_launch_us": launch_s * 1e6,
            "traffic": traffic,
            "traffic_GBps": None if traffic is None else traffic / launch_s / 1e9,
            "traffic_frac": None if traffic is None else traffic / launch_s / 1e9 / HBM_PEAK_GBS,
        },
        "plain_ms_per_step": ms_plain,
        "plain_vertex_steps_per_s": N / (ms_plain * 1e-3),
        "speedup_vs_plain": ms_plain / (ms / steps),
        "plain_persistent_ms_per_step": ms_plain_p,
        "speedup_vs_plain_persistent": ms_plain_p / (ms / steps),
        "setup_s": {"scene": t1 - t0, "create": t2 - t1},
    }


def bench_adjoint(steps: int = 120, every: int = 16) -> dict:
    """Differentiable rollout of one cloth: 128x128, top edge pinned, dt 1/60,
    CG fixed at 40 iterations, per-spring ks; loss = squared distance of the
    final positions to a smooth target; dL/dks, dL/dx0, dL/dv0 by the
    hand-written adjoint with a checkpoint every 16 steps (each adjoint step
    recomputes its forward step from the checkpoints)."""
    from opencloth_amd import scenes
    from opencloth_amd.cloth import ImplicitCloth
    from opencloth_amd.hiptime import Event

    sc = scenes.diff_scene(steps=steps)
    target = sc.extra["target"]
    cl = ImplicitCloth.from_scene(sc)
    cl.rollout_forward(steps, every)  # warm-up
    cl.rollout_backward(2.0 * (cl.X - target))
    cl.set_state(sc.X, sc.V)
    e0, e1, e2 = Event(), Event(), Event()
    e0.record()
    cl.rollout_forward(steps, every)
    e1.record()
    gx = 2.0 * (cl.X - target)
    gks, _, gv0 = cl.rollout_backward(gx)
    e2.record()
    e2.synchronize()
    ms_f, ms = e0.elapsed_ms(e1), e0.elapsed_ms(e2)
    it, st = cl.last_status()
    N = cl.N
    cl.close()
    return {
        "workload": f"{sc.nx}x{sc.ny} cloth, {len(sc.springs)} springs, {steps}-step rollout, "
                    f"checkpoint every {every}, CG 40 iterations both directions, dt 1/60",
        "ms_fwd_adj": ms,
        "ms_forward": ms_f,
        "ms_backward": ms - ms_f,
        "vertex_steps_per_s_fwd_adj": N * steps / (ms * 1e-3),
        "last_status": [it, st],
        "grad_finite": bool(np.isfinite(gks).all() and np.isfinite(gv0).all()),
    }


def bench_batch_adjoint(n_cloths: int = 512, steps: int = 60, every: int = 10, plain: bool = True, schedule=None,
                        sample=(0, 255, 511)) -> dict:
    """Batched forward + adjoint: n_cloths cloths, sides uniform in [32,96],
    per-spring ks log-uniform [20,200], top edges pinned, 60-step rollouts to
    per-cloth smooth targets, CG <= 40 iterations (relative tol 1e-6) in
    both directions, checkpoints every 10 steps.  The plain form runs the same
    rollouts one cloth at a time through oc_cloth_rollout_forward/backward.
    Gradient error: FP32 batch against the FP64 oracle adjoint on sampled
    cloths (relative 2-norm of dL/dks and dL/dv0)."""
    from opencloth_amd import scenes
    from opencloth_amd.batch import ClothBatch
    from opencloth_amd.cloth import ImplicitCloth
    from opencloth_amd.hiptime import Event

    sc = scenes.batch_diff_scene(n_cloths=n_cloths, steps=steps)
    target = sc.extra["target"]
    b = ClothBatch.from_scene(sc)
    if schedule:
        b.set_schedule(*schedule)
        sched = schedule
    else:
        sched = b.autotune()
    b.rollout_forward(steps, every)  # warm-up
    XT, _ = b.get_state()
    gx = (2.0 * (XT - target)).astype(np.float32)
    b.rollout_backward(gx, n_springs=len(sc.springs))
    b.set_state(sc.X, sc.V)
    e0, e1 = Event(), Event()
    e0.record()
    b.rollout_forward(steps, every)
    XT, _ = b.get_state()
    gx = (2.0 * (XT - target)).astype(np.float32)
    gks, _, gv0 = b.rollout_backward(gx, n_springs=len(sc.springs))
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_ms(e1)
    b.close()

    def plain_form(persistent: bool) -> float:
        """ms of the same rollouts and adjoints one cloth at a time
        (oc_cloth_rollout_forward/backward); persistent: every solve is one
        cooperative launch that stops at convergence."""
        cls_ = []
        for c in range(sc.n_cloths):
            cs = sc.cloth(c)
            p = dataclasses.replace(cs.params, cg_persistent=persistent)
            cls_.append(ImplicitCloth(cs.nx, cs.ny, cs.X, cs.springs, cs.pinned, p, cs.V))
        for cl in cls_[:2]:
            cl.rollout_forward(steps, every)
            cl.rollout_backward(np.zeros((cl.N, 3), np.float32))
        p0, p1 = Event(), Event()
        p0.record()
        for c, cl in enumerate(cls_):
            v0, n = int(sc.voff[c]), cl.N
            cl.set_state(sc.X[v0:v0 + n], sc.V[v0:v0 + n])
            cl.rollout_forward(steps, every)
            xt = cl.X
            cl.rollout_backward(2.0 * (xt - target[v0:v0 + n]))
        p1.record()
        p1.synchronize()
        for cl in cls_:
            cl.close()
        return p0.elapsed_ms(p1)

    ms_plain = plain_form(False) if plain else float("nan")
    ms_plain_p = plain_form(True) if plain else float("nan")

    from oracle import Oracle, make_params

    o = Oracle(64)
    errs = []
    for c in sample:
        cs = sc.cloth(c)
        v0, s0, n = int(sc.voff[c]), int(sc.soff[c]), cs.n_vertices
        s = o.system(n, cs.springs, cs.pinned)
        prm = make_params(cs.params)
        XTr, _, states = s.rollout_forward(cs.X, cs.V, prm, steps)
        rks, _, rv0 = s.rollout_grad(cs.X, cs.V, prm, steps, 2 * (XTr - target[v0:v0 + n]), states=states)
        errs.append({"cloth": int(c), "side": int(cs.nx),
                     "ks_rel_err": float(np.linalg.norm(gks[s0:s0 + len(cs.springs)] - rks) / np.linalg.norm(rks)),
                     "v0_rel_err": float(np.linalg.norm(gv0[v0:v0 + n] - rv0) / np.linalg.norm(rv0))})
    N = int(sc.X.shape[0])
    return {
        "workload": f"{sc.n_cloths} cloths, sides U[32,96], {N} vertices, {len(sc.springs)} springs, "
                    f"{steps}-step rollouts, checkpoint every {every}, CG <= 40 it rel tol 1e-6",
        "ms_fwd_adj": ms,
        "vertex_steps_per_s_fwd_adj": N * steps / (ms * 1e-3),
        "plain_ms_fwd_adj": ms_plain,
        "speedup_vs_plain": ms_plain / ms,
        "plain_persistent_ms_fwd_adj": ms_plain_p,
        "speedup_vs_plain_persistent": ms_plain_p / ms,
        "schedule": {"block_threads": sched[0], "vertices_per_thread": sched[1]},
        "grad_error_vs_fp64": errs,
        # the adjoint launch (k_batch_step<..., true>) from the committed
        # kernel trace and PMC summary of bench.py's batch_adjoint leg
        "adjoint_kernel": adjoint_kernel_profile(),
    }


def adjoint_kernel_profile() -> dict:
    path = os.path.join(ROOT, "profiles", "batch_adjoint_512_cloths_pmc.txt")
    us = read_profile_value(path, "batch_adjoint_avg_launch_us")
    traffic = read_profile_value(path, "batch_adjoint_hbm_bytes_per_launch")
    if us is None or traffic is None:
        return {"avg_launch_us": us, "traffic": traffic}
    return {"avg_launch_us": us, "traffic": traffic, "traffic_GBps": traffic / (us * 1e-6) / 1e9,
            "traffic_frac": traffic / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, "source": path[len(ROOT) + 1:]}


LEGS = ("large", "batch", "batch_adjoint", "adjoint", "entries", "cpu")


def kernel_split(cl, steps: int, reps: int = 200) -> dict:
    """Per-step kernel time of an n-step oc_cloth_step call, measured inside
    the real launch sequence: oc_cloth_profile_steps runs the same steps with
    a start/stop event attached to every kernel dispatch (no packets between
    launches) and sums each class.  Beside it, each class's duration in
    back-to-back launches of that kernel alone (oc_cloth_time_kernel), which
    reuses its inputs from cache and reads lower."""
    prof = cl.profile_steps(steps)
    out = {k: v[0] / steps for k, v in prof.items()}
    out["sum"] = sum(out.values())
    out["launches_per_step"] = {k: v[1] / steps for k, v in prof.items()}
    out["in_sequence_launch_us"] = {k: (1e3 * v[0] / v[1] if v[1] else None) for k, v in prof.items()}
    out["back_to_back_launch_us"] = {k: cl.time_kernel(k, reps) for k in cl.KERNELS}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="large", choices=["large", "plain"])
    ap.add_argument("--legs", default=",".join(LEGS),
                    help="comma list of legs to run: " + ", ".join(LEGS) + " (the full line needs all)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched legs")
    ap.add_argument("--no-entries", action="store_true",
                    help="skip the Verlet / PBD / StepPhysics timings")
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--batch-schedule", default=None, metavar="NT,VPT",
                    type=lambda t: tuple(int(x) for x in t.split(",")),
                    help="fixed batch schedule (block threads, vertices per thread) instead of "
                         "the autotuner, for counter runs that must repeat one launch")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the batched legs' one-cloth-per-launch denominators (profiling runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the timing, run one device-to-device copy of MIB MiB (a kernel of "
                         "known HBM traffic, the scale for FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only; a single cloth would shard into row strips with "
              "2-row halo exchanges and two scalar all-reduces per CG iteration, a batch by "
              "instance, and neither is built here", file=sys.stderr)
        sys.exit(2)
    legs = [x for x in args.legs.split(",") if x]
    bad = [x for x in legs if x not in LEGS]
    if bad:
        ap.error(f"unknown legs {bad}")
    if args.no_cpu:
        legs = [x for x in legs if x != "cpu"]
    if args.no_batch:
        legs = [x for x in legs if x not in ("batch", "batch_adjoint")]
    if args.no_entries:
        legs = [x for x in legs if x != "entries"]

    from opencloth_amd import scenes
    from opencloth_amd.cloth import ImplicitCloth
    from opencloth_amd.hiptime import Event

    sc = scenes.large_scene() if args.workload == "large" else scenes.plain_scene()
    out = {"metric": METRIC}
    if "large" in legs:
        cl = ImplicitCloth.from_scene(sc)
        N, S, P, K = sc.n_vertices, cl.S, cl.P, sc.params.cg_max_iters

        cl.step(args.warmup)
        cl.synchronize()
        e0, e1 = Event(), Event()
        e0.record()
        cl.step(args.steps)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_ms(e1)
        it, st = cl.last_status()
        sticky = cl.sticky_status()
        ms_per_step = ms / args.steps

        # the same steps again with every dispatch timed (same state path)
        split = kernel_split(cl, args.steps)
        # the dominant kernel's launch time inside the step's launch sequence
        spmv_avg_s = split["in_sequence_launch_us"]["cg_spmv"] * 1e-6
        spmv_b2b_s = cl.time_cg_spmv(300) * 1e-6
        achieved = spmv_bytes(N, P) / spmv_avg_s / 1e9
        traffic = read_profile_value(os.path.join(ROOT, "profiles", "cg_spmv_512x512_pmc.txt"),
                                     "cg_spmv_hbm_bytes_per_launch")
        out.update({
            "value": N * args.steps / (ms * 1e-3),
            "unit": "vertex-steps/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": ms_per_step,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic",
            "config": {
                "workload": f"implicit step, {sc.nx}x{sc.ny} cloth, "
                            + ("top edge pinned, CG fixed 30 iterations, dt 1/120"
                               if args.workload == "large" else "2 corners pinned, CG <=10 it tol 1e-3, dt 1/60"),
                "vertices": N, "springs": S, "pairs": P, "cg_iters": K,
                "step_algorithmic_bytes": step_bytes(N, S, P, K),
                "step_roofline_frac": step_bytes(N, S, P, K) / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "last_cg_iters": it, "last_status": st, "sticky_status": sticky,
                "kernel_ms_per_step": split,
            },
            "roofline": {
                "bound": "hbm",
                "kernel": "cg_spmv",
                "achieved": achieved,
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic,
                "algorithmic_bytes_per_launch": spmv_bytes(N, P),
                "avg_launch_us": spmv_avg_s * 1e6,
                "avg_launch_us_back_to_back": spmv_b2b_s * 1e6,
                "frac_back_to_back": spmv_bytes(N, P) / spmv_b2b_s / 1e9 / HBM_PEAK_GBS,
                "traffic_GBps": None if traffic is None else traffic / spmv_avg_s / 1e9,
                "traffic_frac": None if traffic is None else traffic / spmv_avg_s / 1e9 / HBM_PEAK_GBS,
            },
        })
        cl.close()
        # the same workload with the single-reduction CG (one launch per CG
        # iteration; Chronopoulos-Gear recurrence, other rounding)
        p1 = dataclasses.replace(sc.params, cg_single_reduction=True)
        cl1 = ImplicitCloth(sc.nx, sc.ny, sc.X, sc.springs, sc.pinned, p1, sc.V)
        cl1.step(args.warmup)
        cl1.synchronize()
        e0.record()
        cl1.step(args.steps)
        e1.record()
        e1.synchronize()
        ms1 = e0.elapsed_ms(e1) / args.steps
        out["config"]["cg_single_reduction"] = {
            "ms_per_step": ms1, "vertex_steps_per_s": N / (ms1 * 1e-3),
            "speedup_vs_standard": ms_per_step / ms1,
            "last_status": cl1.last_status(), "kernel_ms_per_step": kernel_split(cl1, args.steps)}
        cl1.close()
    if "batch" in legs:
        out["batch"] = bench_batch(args.batch_steps, 2, plain=not args.no_plain, schedule=args.batch_schedule)
    if "batch_adjoint" in legs:
        out["batch_adjoint"] = bench_batch_adjoint(plain=not args.no_plain, schedule=args.batch_schedule)
    if "adjoint" in legs:
        out["adjoint"] = bench_adjoint()
    if "entries" in legs:
        out["entries"] = bench_entries()
    if "cpu" in legs:
        out["cpu_baseline"] = cpu_baseline(sc)
    if args.copy_probe > 0:
        out["copy_probe_bytes"] = copy_probe(args.copy_probe)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

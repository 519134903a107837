#!/usr/bin/env python3
"""Benchmark: body-steps/s of b2World::Step on the random pile, one MI355X.

Workload: the random pile of 100k dynamic bodies (50% circles r~U[0.25,0.75],
50% convex polygons of 3-8 points on a circle of R~U[0.3,0.8]), density 1,
friction 0.6, restitution 0, dropped from a 300x600 m grid into a static U
container; dt 1/60, 8 velocity / 3 position iterations, sleeping and
continuous collision off on both sides, 1000 steps from a fresh load. A step
is one b2World::Step over the whole pile. With --steps K the last K steps of
those 1000 are timed (all state already in HBM) and the steps before them run
untimed; `full_window` reports the rate over all 1000.

Also reported on the same JSON line:
  roofline      the dominant kernel (velocity solve) against HBM bandwidth:
                algorithmic bytes per launch from the byte model
                (132 B per constraint + 48 B per manifold point per velocity
                iteration) over its mean execution time, measured in a second
                pass over the whole workload (fresh pile, the same 1000 steps)
                with HIP events that its dispatch packets record on the step
                stream (hipExtLaunchKernelGGL); traffic is the HBM bytes per
                launch measured by rocprofv3 counters; stream_us_per_launch is
                the time per launch seen on the stream, dispatch gaps included;
  kernel_rooflines
                the same figures for the next kernels (position solve,
                broad-phase query, narrowphase), each from its own byte model;
  cpu_baseline  the reference b2World::Step (upstream Box2D 2.3.2 built from
                the reference's sources, oracle/_ref) on one host core over the
                same window: started from the device's state before the timed
                steps, bounded by --cpu-seconds;
  state_at_end / reference_at_end / matched_penetration
                penetration, FP64 complementarity residual and kinetic energy
                of the device state after the last step, the reference's at the
                same step (committed fixture), and the penetration bound.
Side sections (pyramid, learned warm start, batched worlds, stress,
continuous collision) each record an "error" field instead of failing the line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import box2dpp_amd  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the main kernels on the default workload (100k
# pile, seed 1, steps 1-1000, 8/3 iterations), from rocprofv3 FETCH_SIZE and
# WRITE_SIZE averaged over each kernel's launches and scaled by the counters'
# readings for a plain 256 MiB copy (x2.000 and x1.000), keyed by the step
# schedule the counters ran with (the fusion of trailing colours changes
# which colours the solver kernels see), with the library's one colour
# demotion round per step; the summaries with the commands are under
# profiles/. Another schedule or demotion setting reports traffic as null.
DEFAULT_TRAFFIC_BYTES = {(0, 1, 256, 256, 64, 2048, 0, 0): {
    "k_solve_velocity": 8765133.0, "k_solve_position": 7220736.0, "k_collide": 569667277.0,
    "k_pairs_query": 75295642.0}}
# The pile's step schedule as the autotuner picks it on this workload (step
# 500 of a warm-up copy): fixed so the HBM traffic above applies; the batched
# and stress worlds are tuned on every run.
PILE_SCHEDULE = "0,1,256,256,64,2048,0,0"
DT = 1.0 / 60.0
# b2g_schedule: island labelling method, collide order, workgroup sizes of
# the solver, narrowphase and broad-phase query kernels, trailing-colour
# fusion, cluster size of the clustered solve
SCHEDULE = np.dtype([("islands", "<i4"), ("collideOrder", "<i4"), ("solverBlock", "<i4"), ("collideBlock", "<i4"),
                     ("queryBlock", "<i4"), ("tailColours", "<i4"), ("clusterBodies", "<i4"),
                     ("persistentSolve", "<i4")])
_VEC2 = np.dtype([("x", "<f4"), ("y", "<f4")])
_MPOINT = np.dtype([("lp", _VEC2), ("ni", "<f4"), ("ti", "<f4"), ("id", "<u4")])
# b2g_manifold (include/b2g.h), 64 bytes
MANIFOLD = np.dtype([("points", _MPOINT, (2,)), ("ln", _VEC2), ("lpt", _VEC2), ("type", "<i4"), ("count", "<i4")])
# the schedule the benchmark's worlds run with (set by the autotuner)
_schedule = None
# colour demotion rounds of the pile's worlds (None: the library's default)
_demotion = None


class PhaseProfile(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 24), ("launches", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("ms", ctypes.c_double)]


class StepStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "bodies", "fixtures", "contacts", "constraints", "points", "colours", "colouringRounds", "newPairs",
        "positionIterations", "beginEvents", "endEvents", "overflow")] + [
        ("minSeparation", ctypes.c_float), ("cellSize", ctypes.c_float), ("freshPoints", ctypes.c_int32),
        ("fallback", ctypes.c_int32), ("residual", ctypes.c_float), ("toiEvents", ctypes.c_int32),
        ("queries", ctypes.c_int32), ("scanned", ctypes.c_int32), ("wideConstraints", ctypes.c_int32)]


def sched_dict(sc):
    return None if sc is None else {k: int(sc[0][k]) for k in SCHEDULE.names}


def progress(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    sys.stderr.write(f"[bench {time.strftime('%H:%M:%S')}] {msg}\n")
    sys.stderr.flush()


def apply_schedule(lib, h):
    if _schedule is not None:
        rc = lib.b2pp_world_set_schedule(h, _schedule.ctypes.data_as(ctypes.c_void_p))
        if rc != 0:
            raise RuntimeError(f"schedule rejected: {lib.b2pp_world_last_error(h).decode()}")


def make_world(lib, n, seed):
    h = ctypes.c_void_p(lib.b2pp_world_create(0.0, -10.0))
    lib.b2pp_world_set_flags(h, 0, 1, 0, 0)
    rc = lib.b2pp_world_load_scene(h, b"pile", n, ctypes.c_uint(seed))
    if rc != 0:
        raise RuntimeError(f"scene load failed: {lib.b2pp_world_last_error(h)}")
    apply_schedule(lib, h)
    if _demotion is not None:
        b2g = box2dpp_amd.b2g()
        ctx = ctypes.c_void_p(lib.b2pp_world_device_context(h))
        _check(b2g, ctx, b2g.b2g_set_colour_demotion(ctx, _demotion))
    lib.b2pp_world_sync(h)
    return h


def autotune(lib, h, args, at):
    """Schedule chosen by the autotuner on world h after `at` steps."""
    out = np.zeros(1, SCHEDULE)
    step(lib, h, at, args.vel, args.pos)
    progress(f"autotuning at step {at} ({20 * args.tune_rounds} steps)")
    rc = lib.b2pp_world_autotune(h, ctypes.c_float(DT), args.vel, args.pos, args.tune_rounds,
                                 out.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError(f"autotune failed rc={rc}: {lib.b2pp_world_last_error(h).decode()}")
    return out


def step(lib, h, k, vel, pos):
    rc = lib.b2pp_world_step(h, ctypes.c_float(DT), vel, pos, k)
    if rc != 0:
        raise RuntimeError(f"step failed rc={rc}: {lib.b2pp_world_last_error(h).decode()}")


def stats_of(lib, h):
    st = StepStats()
    lib.b2pp_world_stats(h, ctypes.byref(st))
    return st


def _ref_lib():
    """The reference build (upstream Box2D 2.3.2 from the reference's own
    sources, oracle/_ref) or, where it was not built, the CPU restatement."""
    ref = os.path.join(ROOT, "oracle", "_ref", "libb2ref.so")
    if os.path.exists(ref):
        lib, pre, kind = ctypes.CDLL(ref), "ref_", "reference"
    else:
        lib, pre, kind = ctypes.CDLL(os.path.join(ROOT, "oracle", "_build", "liborc.so")), "orc_", "port"
    f = lambda name: getattr(lib, pre + name)  # noqa: E731
    f("world_create").restype = ctypes.c_void_p
    f("world_create").argtypes = [ctypes.c_float, ctypes.c_float]
    return f, kind


def cpu_reference_window(state, n, seed, first, k, vel, pos, budget_s):
    """Reference b2World::Step on one core over the same window as the timed
    device steps: the reference is started from the device's state before
    step `first` (poses and velocities of every body; its contacts are created
    by one untimed step, without carried impulses), then up to k steps are
    timed within the budget."""
    f, kind = _ref_lib()
    if kind != "reference":
        return cpu_reference_prefix(n, seed, vel, pos, budget_s)
    h = ctypes.c_void_p(f("world_create")(0.0, -10.0))
    f("world_set_flags")(h, 0, 1, 0, 0)
    st = np.ascontiguousarray(state, np.float32)
    rc = f("world_load_scene_state")(h, b"pile", n, ctypes.c_uint(seed), st.ctypes.data_as(ctypes.c_void_p), len(st))
    if rc != 0:
        raise RuntimeError(f"reference state load failed rc={rc}")
    t0 = time.perf_counter()
    f("world_step")(h, ctypes.c_float(DT), vel, pos, 1)
    first_s = time.perf_counter() - t0
    steps = 0
    t0 = time.perf_counter()
    while steps < k:
        f("world_step")(h, ctypes.c_float(DT), vel, pos, 1)
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    el = time.perf_counter() - t0
    f("world_destroy")(h)
    return {"value": n * steps / el, "unit": "body-steps/s", "cores": 1, "kind": kind,
            "sample": f"steps {first + 2}-{first + 1 + steps} of the same {n}-body pile (seed {seed}): the reference "
                      f"started from the device's state after step {first} (poses and velocities), step "
                      f"{first + 1} untimed to create its contacts ({first_s:.1f} s), then {steps} steps timed on "
                      f"one host thread ({el:.1f} s)"}


def cpu_reference_prefix(n, seed, vel, pos, budget_s):
    """Reference b2World::Step on one core: the first steps of the pile that
    fit the budget (used where the reference build is absent)."""
    f, kind = _ref_lib()
    h = ctypes.c_void_p(f("world_create")(0.0, -10.0))
    f("world_set_flags")(h, 0, 1, 0, 0)
    f("world_load_scene")(h, b"pile", n, ctypes.c_uint(seed))
    steps = 0
    t0 = time.perf_counter()
    while True:
        f("world_step")(h, ctypes.c_float(DT), vel, pos, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 1000:
            break
    f("world_destroy")(h)
    return {"value": n * steps / el, "unit": "body-steps/s", "cores": 1, "kind": kind,
            "sample": f"steps 1-{steps} of 1000 of the same {n}-body pile (seed {seed}), one host thread, "
                      f"{el:.1f} s; early steps are cheaper than the settled pile"}


def reference_stats(n, seed, at):
    """The reference build's post-step measures of the same pile at step
    `at`, from the committed fixture (tests/golden/pile_ref_stats.json,
    written by tests/golden/make_pile_stats.py)."""
    path = os.path.join(ROOT, "tests", "golden", "pile_ref_stats.json")
    if not os.path.exists(path):
        return None
    data = json.load(open(path))
    for v in data.values():
        if v.get("bodies") == n and v.get("seed") == seed and str(at) in v.get("at", {}):
            m = dict(v["at"][str(at)])
            m["step"] = at
            m["source"] = "tests/golden/pile_ref_stats.json (reference build, one core, offline)"
            return m
    return None


def gpu_measures(lib, h):
    """Post-step measures of a device world (box2dpp_amd.measures, FP64)."""
    from box2dpp_amd import measures
    nb = lib.b2pp_world_body_count(h)
    nf = lib.b2pp_world_fixture_count(h)
    bodies = np.zeros((nb, 12), np.float32)
    lib.b2pp_world_get_bodies(h, bodies.ctypes.data_as(ctypes.c_void_p))
    nc = lib.b2pp_world_contact_count(h)
    meta = np.zeros((max(nc, 1), 4), np.int32)
    man = np.zeros(max(nc, 1), MANIFOLD)
    got = lib.b2pp_world_get_contacts(h, meta.ctypes.data_as(ctypes.c_void_p), man.ctypes.data_as(ctypes.c_void_p),
                                      nc)
    nc = min(nc, got)
    fb = np.zeros(max(nf, 1), np.int32)
    fm = np.zeros((max(nf, 1), 3), np.float32)
    lib.b2pp_world_fixture_info(h, fb.ctypes.data_as(ctypes.c_void_p), fm.ctypes.data_as(ctypes.c_void_p))
    mass = np.zeros((max(nb, 1), 2), np.float32)
    lib.b2pp_world_body_mass(h, mass.ctypes.data_as(ctypes.c_void_p))
    m = measures.all_measures(bodies, meta[:nc], man[:nc], fb[:nf], fm[:nf], mass[:nb])
    m["contacts"] = int(nc)
    b2g = box2dpp_amd.b2g()
    counts = np.zeros(65, np.int32)
    b2g.b2g_get_colour_counts(ctypes.c_void_p(lib.b2pp_world_device_context(h)), counts.ctypes.data_as(ctypes.c_void_p),
                              65, None)
    m["colour_sizes"] = [int(v) for v in counts[:int(np.nonzero(counts)[0].max()) + 1]] if counts.any() else []
    return m


def get_bodies(lib, h):
    nb = lib.b2pp_world_body_count(h)
    out = np.zeros((nb, 12), np.float32)
    lib.b2pp_world_get_bodies(h, out.ctypes.data_as(ctypes.c_void_p))
    return out


MLP_PARAM_FLOATS = 64 * 16 + 64 + 2 * (64 * 64 + 64) + 2 * 64 + 2 + 16 + 16 + 2 + 2
MLP_TRAIN_FLOPS = 2.0 * 2.0 * (16 * 64 + 2 * 64 * 64 + 64 * 16) + 2.0 * (16 * 64 + 2 * 64 * 64)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def _check(b2g, ctx, rc):
    if rc != 0:
        raise RuntimeError(f"b2g call failed rc={rc}: {b2g.b2g_last_error(ctx).decode()}")


def learned_warm_start(lib, args, scope=0, stride=1):
    """Second half of the metric: velocity iterations to reach a residual,
    plain (cold, warm) against the learned warm start.

    1. log: fresh contact points and their solved impulses from the plain
       solver on a pile of the same size (seed args.log_seed), every step;
    2. train the warm-start network on the log with its own kernels;
    3. evaluate on an unseen pile (seed args.eval_seed): at every
       args.trace_every-th step, the same velocity constraints are solved
       from the three starts for args.trace_iters iterations, recording the
       RMS complementarity residual (FP64, host) after each;
    4. time the learned step (network inference for fresh points inside the
       step) over the same steps as the plain timed run.

    scope 0 serves fresh points (newly created contact points); scope 1
    serves every solved point, with the carried impulse as an input, and
    logs every `stride`-th step.
    """
    b2g = box2dpp_amd.b2g()
    b2g.b2g_last_error.restype = ctypes.c_char_p
    out = {}
    # 1. log
    progress("learned warm start: logging fresh points")
    h = make_world(lib, args.bodies, args.log_seed)
    ctx = ctypes.c_void_p(lib.b2pp_world_device_context(h))
    _check(b2g, ctx, b2g.b2g_set_learned_scope(ctx, scope, stride))
    _check(b2g, ctx, b2g.b2g_log_fresh(ctx, args.log_capacity))
    t0 = time.perf_counter()
    step(lib, h, args.learned_steps, args.vel, args.pos)
    lib.b2pp_world_sync(h)
    n = ctypes.c_int()
    b2g.b2g_log_size(ctx, ctypes.byref(n))
    out["log"] = {"seed": args.log_seed, "steps": args.learned_steps, "step_stride": stride, "points": n.value,
                  "seconds": round(time.perf_counter() - t0, 2)}
    # 2. train (kernel time of the gradient kernel from dispatch-packet events)
    progress(f"learned warm start: training on {n.value} points")
    _check(b2g, ctx, b2g.b2g_mlp_init(ctx, 5))
    losses = np.zeros(args.epochs, np.float64)
    b2g.b2g_set_profiling(ctx, 1)
    t0 = time.perf_counter()
    _check(b2g, ctx, b2g.b2g_mlp_train(ctx, args.epochs, args.batch, ctypes.c_float(args.lr), 9,
                                       losses.ctypes.data_as(ctypes.c_void_p)))
    train_s = time.perf_counter() - t0
    prof = (PhaseProfile * 32)()
    cnt = ctypes.c_int(0)
    b2g.b2g_get_profile(ctx, prof, 32, ctypes.byref(cnt))
    b2g.b2g_set_profiling(ctx, 0)
    entries = {prof[i].name.decode(): (prof[i].launches, prof[i].ms) for i in range(min(cnt.value, 32))}
    g_launches, g_ms = entries.get("k_mlp_grad", (0, 0.0))
    flops = MLP_TRAIN_FLOPS * n.value * args.epochs
    tflops = flops / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
    params = (ctypes.c_float * MLP_PARAM_FLOATS)()
    _check(b2g, ctx, b2g.b2g_mlp_get_params(ctx, params))
    lib.b2pp_world_destroy(h)
    out["train"] = {"epochs": args.epochs, "batch": args.batch, "lr": args.lr,
                    "loss_first_epoch": round(float(losses[0]), 4), "loss_last_epoch": round(float(losses[-1]), 4),
                    "seconds": round(train_s, 2)}
    out["mfma"] = {"kernel": "k_mlp_grad", "dtype": "bf16", "launches": g_launches,
                   "avg_launch_us": round(1e3 * g_ms / max(g_launches, 1), 2),
                   "achieved_tflops": round(tflops, 2), "peak_tflops": MFMA_BF16_PEAK_TFLOPS,
                   "frac": tflops / MFMA_BF16_PEAK_TFLOPS,
                   "flops_per_sample": MLP_TRAIN_FLOPS}
    # 3. residual traces on an unseen pile
    progress("learned warm start: residual traces")
    h = make_world(lib, args.bodies, args.eval_seed)
    ctx = ctypes.c_void_p(lib.b2pp_world_device_context(h))
    _check(b2g, ctx, b2g.b2g_set_learned_scope(ctx, scope, stride))
    _check(b2g, ctx, b2g.b2g_mlp_set_params(ctx, params))
    T = args.trace_iters
    curves = []
    buf = np.zeros(3 * (T + 1), np.float64)
    it = ctypes.c_int()
    for s in range(1, args.learned_steps + 1):
        traced = s >= args.trace_from and s % args.trace_every == 0
        if traced:
            b2g.b2g_set_trace(ctx, T)
        step(lib, h, 1, args.vel, args.pos)
        if traced:
            b2g.b2g_set_trace(ctx, 0)
            _check(b2g, ctx, b2g.b2g_get_trace(ctx, buf.ctypes.data_as(ctypes.c_void_p), len(buf), ctypes.byref(it)))
            if it.value == T:
                curves.append(buf.reshape(3, T + 1).copy())
    lib.b2pp_world_destroy(h)
    if not curves:
        raise RuntimeError(f"no residual trace in steps {args.trace_from}-{args.learned_steps}")
    cur = np.array(curves).mean(axis=0)
    names = ("cold", "warm", "learned")

    def iters_to(row, target):
        hit = np.nonzero(row <= target)[0]
        return int(hit[0]) if len(hit) else None

    plain8 = float(cur[1, min(8, T)])

    out["residual"] = {
        "definition": "RMS over solved points of r_n (vn - bias if normal impulse > 0 else min(vn - bias, 0)) "
                      "and r_t (vt if |tangent impulse| < friction * normal impulse else 0), m/s, FP64 on host",
        "traced_steps": len(curves), "trace_every": args.trace_every, "trace_from": args.trace_from, "eval_seed": args.eval_seed,
        "mean_at_iteration": {nm: {str(k): round(float(cur[i, k]), 5) for k in (0, 1, 2, 4, 8, 16, 32, 64, T) if k <= T}
                              for i, nm in enumerate(names)},
        "iterations_to_1e-4": {nm: iters_to(cur[i], 1e-4) for i, nm in enumerate(names)},
        "plain_warm_residual_at_8": round(plain8, 5),
        "iterations_to_plain_warm_residual_at_8": {nm: iters_to(cur[i], plain8) for i, nm in enumerate(names)},
    }
    # 4. learned step throughput and penetration on the evaluation pile
    progress("learned warm start: learned step")
    h = make_world(lib, args.bodies, args.eval_seed)
    ctx = ctypes.c_void_p(lib.b2pp_world_device_context(h))
    _check(b2g, ctx, b2g.b2g_set_learned_scope(ctx, scope, stride))
    _check(b2g, ctx, b2g.b2g_mlp_set_params(ctx, params))
    lib.b2pp_world_set_learned(h, 1, ctypes.c_float(0.0))
    lib.b2pp_world_sync(h)
    t0 = time.perf_counter()
    step(lib, h, args.learned_steps, args.vel, args.pos)
    lib.b2pp_world_sync(h)
    wall = time.perf_counter() - t0
    pen = np.zeros(3, np.float64)
    lib.b2pp_world_penetration(h, pen.ctypes.data_as(ctypes.c_void_p))
    lib.b2pp_world_destroy(h)
    out["learned_step"] = {"value": args.bodies * args.learned_steps / wall, "unit": "body-steps/s",
                           "ms_per_step": 1e3 * wall / args.learned_steps, "mean_penetration_m": float(pen[0]),
                           "max_penetration_m": float(pen[1])}
    return out, params


def stress(lib, args, params):
    """Stress: args.stress_worlds worlds, each a random pile of
    args.stress_bodies bodies plus args.stress_bodies / 50 free chains of 32
    boxes 0.5 x 0.1 m joined by revolute joints (the links are bodies in
    addition to the pile's), in sub-worlds of one context. The plain solver
    at 8/3 iterations against the learned warm start at 4/2, each from a
    fresh load over args.stress_steps steps: throughput, penetration, joint
    drift and energy at the end."""
    b2g = box2dpp_amd.b2g()
    b2g.b2g_last_error.restype = ctypes.c_char_p
    variants = [("plain", args.vel, args.pos, False)]
    if params is not None:
        variants.append(("learned", 4, 2, True))
    global _schedule
    pile_schedule = _schedule

    def load():
        h = ctypes.c_void_p(lib.b2pp_world_create(0.0, -10.0))
        lib.b2pp_world_set_flags(h, 0, 1, 0, 0)
        rc = lib.b2pp_world_load_worlds(h, b"stress", args.stress_worlds, args.stress_bodies,
                                        ctypes.c_uint(args.stress_seed))
        if rc != 0:
            raise RuntimeError(f"stress load failed rc={rc}: {lib.b2pp_world_last_error(h).decode()}")
        return h

    stress_schedule = None
    if args.tune_rounds > 0:
        progress("stress: loading to autotune")
        h = load()
        stress_schedule = autotune(lib, h, args, 150)
        lib.b2pp_world_destroy(h)
    out = {"workload": f"{args.stress_worlds} worlds x (random pile of {args.stress_bodies} bodies + "
                       f"{args.stress_bodies // 50} chains of 32 links), seed {args.stress_seed}, "
                       f"steps 1-{args.stress_steps}, sleep and CCD off, one context"}
    for name, vel, pos, learned in variants:
        progress(f"stress: loading ({name})")
        t0 = time.perf_counter()
        h = load()
        _schedule = stress_schedule
        apply_schedule(lib, h)
        _schedule = pile_schedule
        ctx = ctypes.c_void_p(lib.b2pp_world_device_context(h))
        if learned:
            _check(b2g, ctx, b2g.b2g_mlp_set_params(ctx, params))
            lib.b2pp_world_set_learned(h, 1, ctypes.c_float(0.0))
        lib.b2pp_world_sync(h)
        load_s = time.perf_counter() - t0
        bodies = lib.b2pp_world_body_count(h)
        joints = lib.b2pp_world_joint_count(h)
        progress(f"stress: {bodies} bodies, {joints} joints loaded in {load_s:.1f} s; {vel}/{pos} iterations")
        t0 = time.perf_counter()
        for k in range(0, args.stress_steps, 100):
            step(lib, h, min(100, args.stress_steps - k), vel, pos)
        lib.b2pp_world_sync(h)
        wall = time.perf_counter() - t0
        pen = np.zeros(3, np.float64)
        lib.b2pp_world_penetration(h, pen.ctypes.data_as(ctypes.c_void_p))
        drift = np.zeros(3, np.float64)
        _check(b2g, ctx, b2g.b2g_joint_drift(ctx, drift.ctypes.data_as(ctypes.c_void_p)))
        energy = np.zeros(2, np.float64)
        _check(b2g, ctx, b2g.b2g_energy(ctx, energy.ctypes.data_as(ctypes.c_void_p)))
        st = stats_of(lib, h)
        lib.b2pp_world_destroy(h)
        out[name] = {"value": bodies * args.stress_steps / wall, "unit": "body-steps/s", "bodies": bodies,
                     "joints": joints, "vel_iters": vel, "pos_iters": pos, "ms_per_step": 1e3 * wall / args.stress_steps,
                     "load_seconds": round(load_s, 2), "contacts": st.contacts, "colours": st.colours,
                     "mean_penetration_m": float(pen[0]), "max_penetration_m": float(pen[1]),
                     "joint_drift_mean_m": float(drift[0]), "joint_drift_max_m": float(drift[1]),
                     "kinetic_energy_J": float(energy[0]), "potential_energy_J": float(energy[1])}
    out["schedule"] = sched_dict(stress_schedule)
    return out


def batched_worlds(lib, args):
    """Batched worlds: args.batch_worlds independent worlds (dynamic body
    counts log-uniform in [16, 1024], world w built from seed + w, each in its
    own container) in sub-worlds of one b2World, advanced together by one
    Step; body-steps/s over args.batch_steps steps, state resident on the
    device before timing. CPU leg: the reference b2World::Step, one world per
    thread over the host's cores, on a bounded sample of the worlds."""
    global _schedule
    pile_schedule = _schedule

    def load():
        h = ctypes.c_void_p(lib.b2pp_world_create(0.0, -10.0))
        lib.b2pp_world_set_flags(h, 0, 1, 0, 0)
        sizes = np.zeros(args.batch_worlds, np.int32)
        rc = lib.b2pp_world_load_batch(h, args.batch_worlds, ctypes.c_uint(args.batch_seed), 16, 1024,
                                       sizes.ctypes.data_as(ctypes.c_void_p))
        if rc != 0:
            raise RuntimeError(f"batch load failed rc={rc}: {lib.b2pp_world_last_error(h).decode()}")
        lib.b2pp_world_sync(h)
        return h, sizes

    batch_schedule = None
    if args.tune_rounds > 0:
        progress(f"batched worlds: loading {args.batch_worlds} worlds to autotune")
        h, _ = load()
        batch_schedule = autotune(lib, h, args, 100)
        lib.b2pp_world_destroy(h)
    progress(f"batched worlds: loading {args.batch_worlds} worlds")
    t0 = time.perf_counter()
    h, sizes = load()
    _schedule = batch_schedule
    apply_schedule(lib, h)
    _schedule = pile_schedule
    load_s = time.perf_counter() - t0
    bodies = int(sizes.sum())
    progress(f"batched worlds: {bodies} bodies loaded in {load_s:.1f} s; stepping")
    t0 = time.perf_counter()
    step(lib, h, args.batch_steps, args.vel, args.pos)
    lib.b2pp_world_sync(h)
    wall = time.perf_counter() - t0
    st = stats_of(lib, h)
    pen = np.zeros(3, np.float64)
    lib.b2pp_world_penetration(h, pen.ctypes.data_as(ctypes.c_void_p))
    lib.b2pp_world_destroy(h)
    out = {"value": bodies * args.batch_steps / wall, "unit": "body-steps/s", "worlds": args.batch_worlds,
           "bodies": bodies, "steps": args.batch_steps, "ms_per_step": 1e3 * wall / args.batch_steps,
           "load_seconds": round(load_s, 2), "contacts": st.contacts, "colours": st.colours,
           "mean_penetration_m": float(pen[0]), "schedule": sched_dict(batch_schedule),
           "workload": f"{args.batch_worlds} worlds, bodies log-uniform in [16, 1024] (seed {args.batch_seed}), "
                       f"{args.vel}/{args.pos} iterations, sleep and CCD off, one context"}
    if not args.no_cpu:
        progress("batched worlds: CPU reference sample")
        out["cpu_baseline"] = cpu_batched(sizes, args)
    return out


def cpu_batched(sizes, args):
    """Reference b2World::Step on host threads, one world per task."""
    import concurrent.futures as cf
    ref = os.path.join(ROOT, "oracle", "_ref", "libb2ref.so")
    kind = "reference"
    if os.path.exists(ref):
        lib, pre = ctypes.CDLL(ref), "ref_"
    else:
        lib, pre = ctypes.CDLL(os.path.join(ROOT, "oracle", "_build", "liborc.so")), "orc_"
        kind = "port"
    f = lambda name: getattr(lib, pre + name)
    f("world_create").restype = ctypes.c_void_p
    f("world_create").argtypes = [ctypes.c_float, ctypes.c_float]
    threads = max(1, min(16, os.cpu_count() or 1))
    rng = np.random.default_rng(0)
    sample = np.sort(rng.choice(len(sizes), size=min(args.batch_cpu_worlds, len(sizes)), replace=False))

    def run(w):
        hw = ctypes.c_void_p(f("world_create")(0.0, -10.0))
        f("world_set_flags")(hw, 0, 1, 0, 0)
        f("world_load_scene")(hw, b"world", int(sizes[w]), ctypes.c_uint(args.batch_seed + int(w)))
        for _ in range(args.batch_cpu_steps):
            f("world_step")(hw, ctypes.c_float(1.0 / 60.0), args.vel, args.pos, 1)
        f("world_destroy")(hw)
        return int(sizes[w]) * args.batch_cpu_steps

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        total = sum(ex.map(run, sample))
    el = time.perf_counter() - t0
    return {"value": total / el, "unit": "body-steps/s", "cores": threads, "kind": kind,
            "sample": f"{len(sample)} of the {len(sizes)} worlds (random, seed 0), steps 1-{args.batch_cpu_steps} "
                      f"of {args.batch_steps}, one world per thread, {threads} threads, {el:.1f} s"}


def continuous(lib, args):
    """Continuous collision at the reference's default (SetContinuousPhysics
    on): the random pile of args.ccd_bodies bodies, sleep off, steps
    1..args.ccd_steps, coloured solver plus the TOI event loop; body-steps/s
    with the state resident on the device. CPU leg: the reference
    b2World::Step with its SolveTOI on the same steps of the same pile, one
    core."""
    progress(f"continuous collision: {args.ccd_bodies}-body pile, {args.ccd_steps} steps")
    h = ctypes.c_void_p(lib.b2pp_world_create(0.0, -10.0))
    lib.b2pp_world_set_flags(h, 0, 1, 1, 0)
    rc = lib.b2pp_world_load_scene(h, b"pile", args.ccd_bodies, ctypes.c_uint(args.seed))
    if rc != 0:
        raise RuntimeError(f"pile load failed rc={rc}")
    lib.b2pp_world_sync(h)
    events = 0
    t0 = time.perf_counter()
    for _ in range(args.ccd_steps):
        lib.b2pp_world_step(h, ctypes.c_float(1.0 / 60.0), args.vel, args.pos, 1)
        events += stats_of(lib, h).toiEvents
    lib.b2pp_world_sync(h)
    wall = time.perf_counter() - t0
    pen = np.zeros(3, np.float64)
    lib.b2pp_world_penetration(h, pen.ctypes.data_as(ctypes.c_void_p))
    lib.b2pp_world_destroy(h)
    out = {"value": args.ccd_bodies * args.ccd_steps / wall, "unit": "body-steps/s",
           "ms_per_step": 1e3 * wall / args.ccd_steps, "toi_events_per_step": events / args.ccd_steps,
           "us_per_toi_event": 1e6 * wall / max(events, 1), "mean_penetration_m": float(pen[0]),
           "workload": f"random pile, {args.ccd_bodies} bodies, seed {args.seed}, steps 1-{args.ccd_steps}, "
                       f"{args.vel}/{args.pos} iterations, coloured solver, sleep off, CCD on"}
    if not args.no_cpu:
        progress("continuous collision: CPU reference on the same steps")
        ref = os.path.join(ROOT, "oracle", "_ref", "libb2ref.so")
        kind = "reference"
        if os.path.exists(ref):
            rl, pre = ctypes.CDLL(ref), "ref_"
        else:
            rl, pre = ctypes.CDLL(os.path.join(ROOT, "oracle", "_build", "liborc.so")), "orc_"
            kind = "port"
        f = lambda name: getattr(rl, pre + name)
        f("world_create").restype = ctypes.c_void_p
        f("world_create").argtypes = [ctypes.c_float, ctypes.c_float]
        hc = ctypes.c_void_p(f("world_create")(0.0, -10.0))
        f("world_set_flags")(hc, 0, 1, 1, 0)
        f("world_load_scene")(hc, b"pile", args.ccd_bodies, ctypes.c_uint(args.seed))
        t0 = time.perf_counter()
        done = 0
        for k in range(args.ccd_steps):
            f("world_step")(hc, ctypes.c_float(1.0 / 60.0), args.vel, args.pos, 1)
            done += 1
            if time.perf_counter() - t0 > args.ccd_cpu_seconds:
                break
        el = time.perf_counter() - t0
        f("world_destroy")(hc)
        out["cpu_baseline"] = {"value": args.ccd_bodies * done / el, "unit": "body-steps/s", "cores": 1,
                               "kind": kind,
                               "sample": f"steps 1-{done} of the same pile with SolveTOI, one host thread, "
                                         f"{el:.1f} s"}
    return out


def headline(lib, args):
    """The metric's workload: the pile over `window` steps from a fresh load;
    the last K (--steps) are timed, the steps before them run untimed."""
    total = max(args.window, args.steps)
    prefix = total - args.steps
    if args.warmup > 0 and prefix < args.warmup:
        progress(f"warm-up: {args.warmup} steps on a separate copy")
        hw = make_world(lib, args.bodies, args.seed)
        step(lib, hw, args.warmup, args.vel, args.pos)
        lib.b2pp_world_sync(hw)
        lib.b2pp_world_destroy(hw)
    h = make_world(lib, args.bodies, args.seed)
    lib.b2pp_world_sync(h)
    progress(f"headline: {args.bodies} bodies, steps 1-{prefix} untimed, {prefix + 1}-{total} timed")
    t0 = time.perf_counter()
    for k in range(0, prefix, 100):
        step(lib, h, min(100, prefix - k), args.vel, args.pos)
        if k % 500 == 400:
            progress(f"headline: step {k + 100} of the untimed prefix")
    lib.b2pp_world_sync(h)
    prefix_s = time.perf_counter() - t0
    state = get_bodies(lib, h) if prefix > 0 else None
    lib.b2pp_world_sync(h)
    t0 = time.perf_counter()
    step(lib, h, args.steps, args.vel, args.pos)
    lib.b2pp_world_sync(h)
    wall = time.perf_counter() - t0
    st = stats_of(lib, h)
    progress(f"headline: {1e3 * wall / args.steps:.3f} ms/step over the timed steps; measuring the state")
    meas = gpu_measures(lib, h)
    lib.b2pp_world_destroy(h)
    return {"total": total, "prefix": prefix, "prefix_s": prefix_s, "wall": wall, "state": state, "stats": st,
            "measures": meas}


# Algorithmic bytes of the step's main kernels, per unit (the loads and
# stores each unit's code names, every one counted once per unit):
#   k_solve_velocity  per constraint and iteration 132 B (body indices 8,
#                     info 4, masses 16, normal and friction 16, point masses
#                     16, impulses read 16 and written 16, velocity biases 8,
#                     the two bodies' (v, w) rows read 32 and written 32) + per
#                     manifold point 48 B (anchor offsets 16, and for the
#                     second point the block matrix and its inverse 32, taken
#                     as 16 per point): 180 B for one point, 228 B for two;
#   k_solve_position  per constraint and iteration 156 B (island and its
#                     flag 8, body indices 8, info 4, masses 16, local
#                     centres 16, local points 16, normal and plane point
#                     16, radii 8, body positions read 32 and written 32);
#   k_collide         per contact 360 B (order 4, flags read and written 8,
#                     fixture ids 8, body ids 8, body type and flags 4, fat
#                     boxes 32, transforms 32, manifold read 64 and written
#                     64, shapes 2 x 68: a 24 B header plus 16 B per polygon
#                     vertex, 5.5 vertices on half the pile's fixtures);
#   k_pairs_query     36 B per query (proxy id 4, fat box 16, cell range 16)
#                     + 28 B per grid entry scanned (cell key 8, id 4, fat box
#                     16) + 8 B per new pair written;
#   k_solve_persistent (the one-launch solve: every pass of the step's solve)
#                     the warm start, 124 B per constraint (body indices 8,
#                     info 4, masses 16, normal 16, impulses 16, the two
#                     bodies' (v, w) rows read and written 64) + 16 B per point
#                     (anchor offsets); the velocity passes as above; the
#                     impulse store, 24 B per constraint (slot 4, info 4,
#                     impulses 16) + 8 B per point (manifold impulses); the
#                     position integration, 56 B per body (flags and island 8,
#                     (v, w) 16, (c, a) read and written 32); the position
#                     passes as above; the transform sync, 48 B per body (flags
#                     and island 8, (c, a) 16, local centre 8, transform 16).
BYTES_PER_CONSTRAINT_V, BYTES_PER_POINT_V = 132.0, 48.0
BYTES_PER_CONSTRAINT_WARM, BYTES_PER_POINT_WARM = 124.0, 16.0
BYTES_PER_CONSTRAINT_STORE, BYTES_PER_POINT_STORE = 24.0, 8.0
BYTES_PER_BODY_INTEGRATE_X, BYTES_PER_BODY_SYNC = 56.0, 48.0
BYTES_PER_CONSTRAINT_X = 156.0
BYTES_PER_CONTACT_COLLIDE = 360.0
BYTES_PER_QUERY, BYTES_PER_SCANNED, BYTES_PER_NEW_PAIR = 36.0, 28.0, 8.0


def kernel_bytes(s2, contacts_before, vel):
    """Algorithmic bytes of one step's launches of each main kernel, from the
    step's counters (the per-colour solver kernels see the constraints of the
    colours launched one by one; their points are taken in the step's mean
    ratio of points to constraints)."""
    wide = s2.wideConstraints
    pts = s2.points * wide / s2.constraints if s2.constraints else 0.0
    cons, allpts = s2.constraints, s2.points
    solve = (BYTES_PER_CONSTRAINT_WARM * cons + BYTES_PER_POINT_WARM * allpts
             + vel * (BYTES_PER_CONSTRAINT_V * cons + BYTES_PER_POINT_V * allpts)
             + BYTES_PER_CONSTRAINT_STORE * cons + BYTES_PER_POINT_STORE * allpts
             + (BYTES_PER_BODY_INTEGRATE_X + BYTES_PER_BODY_SYNC) * s2.bodies
             + s2.positionIterations * BYTES_PER_CONSTRAINT_X * cons)
    return {"k_solve_persistent": solve,
            "k_solve_velocity": vel * (BYTES_PER_CONSTRAINT_V * wide + BYTES_PER_POINT_V * pts),
            "k_solve_position": s2.positionIterations * BYTES_PER_CONSTRAINT_X * wide,
            "k_collide": BYTES_PER_CONTACT_COLLIDE * contacts_before,
            "k_pairs_query": BYTES_PER_QUERY * s2.queries + BYTES_PER_SCANNED * s2.scanned
            + BYTES_PER_NEW_PAIR * s2.newPairs}


def roofline_pass(lib, args, total):
    """The same workload from a fresh pile with per-phase events on the step
    stream and the main kernels' execution times from events carried by
    their own dispatch packets, over all `total` steps of the workload."""
    b2g = box2dpp_amd.b2g()
    progress(f"roofline pass: {total} steps with kernel events")
    h = make_world(lib, args.bodies, args.seed)
    ctx = ctypes.c_void_p(lib.b2pp_world_device_context(h))
    b2g.b2g_set_profiling(ctx, 1)
    alg = {}
    contacts_before = 0
    for _ in range(total):
        step(lib, h, 1, args.vel, args.pos)
        s2 = stats_of(lib, h)
        for k, v in kernel_bytes(s2, contacts_before, args.vel).items():
            alg[k] = alg.get(k, 0.0) + v
        contacts_before = s2.contacts
    prof = (PhaseProfile * 32)()
    cnt = ctypes.c_int(0)
    b2g.b2g_get_profile(ctx, prof, 32, ctypes.byref(cnt))
    b2g.b2g_set_profiling(ctx, 0)
    lib.b2pp_world_destroy(h)
    entries = {prof[i].name.decode(): (prof[i].launches, prof[i].ms) for i in range(min(cnt.value, 32))}
    phases = {k: v for k, v in entries.items() if not k.startswith("k_")}
    kernels = {k: v for k, v in entries.items() if k.startswith("k_")}
    traffic = args.traffic_bytes or {}
    # the dominant kernel: the one-launch solve when the schedule runs it,
    # else the per-colour velocity solve
    dominant = "k_solve_persistent" if kernels.get("k_solve_persistent", (0, 0.0))[0] > 0 else "k_solve_velocity"

    def line(name):
        launches, ms = kernels.get(name, (0, 0.0))
        avg_ms = ms / max(launches, 1)
        per_launch = alg.get(name, 0.0) / max(launches, 1)
        achieved = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        return {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic.get(name, traffic.get("dominant") if name == dominant else None),
                "bytes_per_launch": per_launch,
                "avg_launch_us": avg_ms * 1e3, "launches_per_step": launches / max(total, 1),
                "kernel_ms_per_step": ms / max(total, 1)}

    roof = line(dominant)
    ph_launches, ph_ms = phases.get("solve_velocity", (0, 0.0))
    roof["stream_us_per_launch"] = 1e3 * ph_ms / max(ph_launches, 1)
    roof["window"] = f"steps 1-{total} of a fresh pile (the whole workload; the timed steps are its last ones)"
    total_phase_ms = sum(ms for _, ms in phases.values())
    others = [k for k in ("k_solve_velocity", "k_solve_position", "k_pairs_query", "k_collide")
              if k != dominant and kernels.get(k, (0, 0.0))[0] > 0]
    extra = {"kernel_rooflines": {k: line(k) for k in others},
             "phase_ms_per_step": {k: round(v[1] / max(total, 1), 4) for k, v in phases.items()},
             "phase_share": {k: round(v[1] / total_phase_ms, 4) for k, v in phases.items()} if total_phase_ms else {},
             "kernel_us_per_launch": {k: round(1e3 * v[1] / max(v[0], 1), 3) for k, v in kernels.items()},
             "kernel_launches_per_step": {k: round(v[0] / max(total, 1), 2) for k, v in kernels.items()}}
    return roof, extra


def pyramid_leg(lib, args):
    """The 20-row pyramid (210 boxes) over 600 steps at the reference's
    defaults (sleeping and continuous collision on), 8/3 iterations: the
    device step (coloured solver, and the reference-order mode) against the
    reference b2World::Step on one core, same flags, same steps."""
    out = {"workload": "20-row box pyramid, 210 dynamic boxes, 600 steps, 8/3 iterations, sleep and CCD on"}
    for name, mode in (("gpu_coloured", 0), ("gpu_reference_order", 1)):
        h = ctypes.c_void_p(lib.b2pp_world_create(0.0, -10.0))
        lib.b2pp_world_set_flags(h, 1, 1, 1, 0)
        lib.b2pp_world_set_mode(h, mode)
        rc = lib.b2pp_world_load_scene(h, b"pyramid", 20, ctypes.c_uint(0))
        if rc != 0:
            raise RuntimeError(f"pyramid load failed rc={rc}")
        lib.b2pp_world_sync(h)
        t0 = time.perf_counter()
        step(lib, h, 600, args.vel, args.pos)
        lib.b2pp_world_sync(h)
        wall = time.perf_counter() - t0
        m = gpu_measures(lib, h)
        awake = int(get_bodies(lib, h)[:, 10].sum())
        lib.b2pp_world_destroy(h)
        out[name] = {"value": 210 * 600 / wall, "unit": "body-steps/s", "ms_per_step": 1e3 * wall / 600,
                     "mean_penetration_m": m["mean_penetration_m"], "max_penetration_m": m["max_penetration_m"],
                     "awake_bodies_at_end": awake}
    f, kind = _ref_lib()
    h = ctypes.c_void_p(f("world_create")(0.0, -10.0))
    f("world_set_flags")(h, 1, 1, 1, 0)
    f("world_load_scene")(h, b"pyramid", 20, ctypes.c_uint(0))
    t0 = time.perf_counter()
    f("world_step")(h, ctypes.c_float(DT), args.vel, args.pos, 600)
    wall = time.perf_counter() - t0
    pen = np.zeros(3, np.float64)
    f("world_penetration")(h, pen.ctypes.data_as(ctypes.c_void_p))
    f("world_destroy")(h)
    out["cpu_baseline"] = {"value": 210 * 600 / wall, "unit": "body-steps/s", "cores": 1, "kind": kind,
                           "ms_per_step": 1e3 * wall / 600, "mean_penetration_m": float(pen[0]),
                           "max_penetration_m": float(pen[1]), "sample": "the whole workload, one host thread"}
    return out


def _kernel_short(name):
    """A kernel's own name out of a rocprofv3 row (namespaces, template
    arguments beyond the first and the parameter list dropped)."""
    import re
    m = re.search(r"(k_[A-Za-z0-9_]+(?:<[^>]*>)?)", name)
    if m:
        return m.group(1)
    return name.split("(")[0][-48:]


def rocprof_summary(paths, steps, top):
    """Summarises rocprofv3 CSV output (--output-format csv): from a kernel
    trace, per kernel the dispatches, mean duration and mean idle time of the
    queue before each dispatch (the gap since the previous kernel ended), and
    the traced span per step; from counter collections, per kernel the mean
    of each counter per dispatch (rows of one dispatch summed). Prints a
    markdown table."""
    import csv
    from collections import defaultdict
    out = []
    for path in paths:
        with open(path, newline="") as f:
            rows = list(csv.DictReader(f))
        if not rows:
            continue
        if "Counter_Name" in rows[0]:
            per = defaultdict(float)
            names = {}
            for r in rows:
                key = (r["Dispatch_Id"], r["Counter_Name"])
                per[key] += float(r["Counter_Value"])
                names[r["Dispatch_Id"]] = _kernel_short(r["Kernel_Name"])
            agg = defaultdict(lambda: [0, 0.0])
            for (d, c), v in per.items():
                a = agg[(names[d], c)]
                a[0] += 1
                a[1] += v
            out.append(f"counters from {os.path.basename(path)} (mean per dispatch):")
            out.append("| kernel | counter | dispatches | mean |")
            out.append("|---|---|---:|---:|")
            for (k, c), (n, v) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
                out.append(f"| {k} | {c} | {n} | {v / n:.1f} |")
        else:
            ev = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), _kernel_short(r["Kernel_Name"]))
                        for r in rows)
            agg = defaultdict(lambda: [0, 0, 0])
            prev_end = None
            for t0, t1, k in ev:
                a = agg[k]
                a[0] += 1
                a[1] += t1 - t0
                if prev_end is not None:
                    a[2] += max(0, t0 - prev_end)
                prev_end = t1 if prev_end is None else max(prev_end, t1)
            span = ev[-1][1] - ev[0][0]
            busy = sum(a[1] for a in agg.values())
            out.append(f"kernel trace {os.path.basename(path)}: {len(ev)} dispatches, span {span / 1e6:.1f} ms "
                       f"({span / 1e6 / max(steps, 1):.3f} ms per step over {steps} steps), kernels busy "
                       f"{100.0 * busy / span:.1f}% of it")
            out.append("| kernel | dispatches | per step | mean us | idle before, mean us | share of kernel time % |")
            out.append("|---|---:|---:|---:|---:|---:|")
            for k, (n, d, g) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
                out.append(f"| {k} | {n} | {n / max(steps, 1):.2f} | {d / n / 1e3:.2f} | {g / n / 1e3:.2f} | "
                           f"{100.0 * d / busy:.2f} |")
    print("\n".join(out))


def check_abi():
    """The ctypes layouts this script passes to the libraries must be the
    libraries' own sizes: a smaller ctypes struct handed to a call that fills
    the C struct overruns the Python heap (b2g_abi_sizes, include/b2g.h)."""
    b2g = box2dpp_amd.b2g()
    n = b2g.b2g_abi_sizes(None, 0)
    sizes = np.zeros(n, np.int32)
    b2g.b2g_abi_sizes(sizes.ctypes.data_as(ctypes.c_void_p), n)
    want = {"b2g_manifold": (sizes[1], MANIFOLD.itemsize), "b2g_step_stats": (sizes[6], ctypes.sizeof(StepStats)),
            "b2g_phase_profile": (sizes[7], ctypes.sizeof(PhaseProfile)),
            "b2g_mlp_params": (sizes[8], 4 * MLP_PARAM_FLOATS), "b2g_schedule": (sizes[12], SCHEDULE.itemsize)}
    bad = {k: v for k, v in want.items() if int(v[0]) != int(v[1])}
    if bad:
        raise RuntimeError(f"ctypes layouts differ from the library's (C size, ctypes size): {bad}")


def section(name, fn, *a):
    """Runs one side section; a failure is recorded, not raised, so the
    headline line always prints."""
    try:
        return fn(*a)
    except Exception as e:  # noqa: BLE001 - reported in the JSON line
        progress(f"{name}: failed: {e!r}")
        return {"error": repr(e)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000,
                    help="timed steps: the last K steps of the workload's window (all of it when K >= window)")
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed warm-up steps (on a separate copy when the window has no untimed prefix)")
    ap.add_argument("--window", type=int, default=1000, help="steps of the workload (1000 on the random pile)")
    ap.add_argument("--bodies", type=int, default=100000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--vel", type=int, default=8)
    ap.add_argument("--pos", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--demotion", type=int, default=None,
                    help="colour demotion rounds per step of the pile's worlds (default: the library's)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline, its roofline pass and CPU leg (the profiled command)")
    ap.add_argument("--no-roofline", action="store_true", help="skip the roofline pass")
    ap.add_argument("--no-pyramid", action="store_true", help="skip the pyramid leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start evaluation")
    ap.add_argument("--batched", action="store_true",
                    help="run the batched-worlds section (off by default: a host-side crash of this section at "
                         "512 and 4096 worlds has no identified cause, so it is not run unasked)")
    ap.add_argument("--batch-worlds", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--batch-seed", type=int, default=1)
    ap.add_argument("--batch-cpu-worlds", type=int, default=512)
    ap.add_argument("--batch-cpu-steps", type=int, default=500)
    ap.add_argument("--no-stress", action="store_true", help="skip the stress run (piles with jointed chains)")
    ap.add_argument("--stress-worlds", type=int, default=4)
    ap.add_argument("--stress-bodies", type=int, default=125000)
    ap.add_argument("--stress-steps", type=int, default=600)
    ap.add_argument("--stress-seed", type=int, default=1)
    ap.add_argument("--no-ccd", action="store_true", help="skip the continuous-collision run")
    ap.add_argument("--schedule", type=str, default=PILE_SCHEDULE,
                    help="pile step schedule (islands,collideOrder,solverBlock,collideBlock,queryBlock,tailColours,"
                         "clusterBodies), "
                         "or 'auto' to autotune it on a separate copy")
    ap.add_argument("--ccd-bodies", type=int, default=10000)
    ap.add_argument("--ccd-steps", type=int, default=500)
    ap.add_argument("--ccd-cpu-seconds", type=float, default=20.0)
    ap.add_argument("--tune-rounds", type=int, default=3,
                    help="autotuner rounds on the warm-up copy (0: default schedule)")
    ap.add_argument("--tune-at", type=int, default=500, help="step of the warm-up copy at which to autotune")
    ap.add_argument("--all-points-stride", type=int, default=30,
                    help="step stride of the every-point log (0 skips the every-point evaluation)")
    ap.add_argument("--learned-steps", type=int, default=1000, help="steps of the learned warm-start log and eval")
    ap.add_argument("--log-seed", type=int, default=2)
    ap.add_argument("--eval-seed", type=int, default=1001)
    ap.add_argument("--log-capacity", type=int, default=10_000_000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--trace-every", type=int, default=100)
    ap.add_argument("--trace-from", type=int, default=500, help="first step traced (the settled pile)")
    ap.add_argument("--trace-iters", type=int, default=128)
    ap.add_argument("--traffic-bytes", type=float, default=None,
                    help="HBM bytes per launch of the dominant kernel (the one-launch solve, or the per-colour "
                         "velocity solve) measured with rocprofv3 FETCH_SIZE + WRITE_SIZE (replaces the committed "
                         "figures)")
    ap.add_argument("--rocprof-summary", nargs="+", default=None, metavar="CSV",
                    help="summarise rocprofv3 CSV files (kernel trace, counter collection) and exit")
    ap.add_argument("--summary-steps", type=int, default=1, help="steps the summarised trace covers")
    args = ap.parse_args()
    if args.rocprof_summary:
        rocprof_summary(args.rocprof_summary, args.summary_steps, 40)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(batched worlds shard across GPUs by world, a single world by slab; "
                                   "the sharding is a design, not a run)"}))
        return 2
    if args.headline_only:
        args.no_pyramid = args.no_learned = args.no_stress = args.no_ccd = True
        args.batched = False
    default_workload = (args.bodies, args.seed, args.window, args.vel, args.pos) == (100000, 1, 1000, 8, 3) and \
        args.demotion in (None, 1)

    lib = box2dpp_amd.host()
    check_abi()
    global _schedule, _demotion
    _demotion = args.demotion
    pile_schedule = None
    if args.schedule and args.schedule != "auto":
        vals = [int(v) for v in args.schedule.split(",")]
        pile_schedule = np.zeros(1, SCHEDULE)
        for name, v in zip(SCHEDULE.names, vals):
            pile_schedule[0][name] = v
    elif args.tune_rounds > 0:
        hw = make_world(lib, args.bodies, args.seed)
        pile_schedule = autotune(lib, hw, args, args.tune_at)
        lib.b2pp_world_destroy(hw)
    _schedule = pile_schedule
    if args.traffic_bytes is not None:
        args.traffic_bytes = {"dominant": args.traffic_bytes}
    elif default_workload and pile_schedule is not None:
        args.traffic_bytes = DEFAULT_TRAFFIC_BYTES.get(tuple(int(pile_schedule[0][k]) for k in SCHEDULE.names))

    hl = headline(lib, args)
    total, prefix, wall, st, meas = hl["total"], hl["prefix"], hl["wall"], hl["stats"], hl["measures"]
    value = args.bodies * args.steps / wall
    ref_at_end = reference_stats(args.bodies, args.seed, total)
    roof, extra = None, {}
    if not args.no_roofline:
        res = section("roofline", roofline_pass, lib, args, total)
        if isinstance(res, dict):  # the section failed
            extra = {"roofline_error": res["error"]}
        else:
            roof, extra = res

    cpu = None
    if not args.no_cpu:
        progress("CPU reference over the timed window")
        if hl["state"] is not None:
            cpu = section("cpu_baseline", cpu_reference_window, hl["state"], args.bodies, args.seed, prefix,
                          args.steps, args.vel, args.pos, args.cpu_seconds)
        else:
            cpu = section("cpu_baseline", cpu_reference_prefix, args.bodies, args.seed, args.vel, args.pos,
                          args.cpu_seconds)

    pyramid = None if args.no_pyramid else section("pyramid", pyramid_leg, lib, args)
    learned, params = None, None
    if not args.no_learned:
        res = section("learned", learned_warm_start, lib, args)
        learned, params = (res, None) if isinstance(res, dict) else res
    learned_all = None
    if not args.no_learned and args.all_points_stride > 0:
        progress("learned warm start, every point")
        res = section("learned_every_point", learned_warm_start, lib, args, 1, args.all_points_stride)
        learned_all = res if isinstance(res, dict) else res[0]
    batched = section("batched", batched_worlds, lib, args) if args.batched else None
    stressed = None if args.no_stress else section("stress", stress, lib, args, params)
    ccd = None if args.no_ccd else section("ccd", continuous, lib, args)

    matched = None
    if ref_at_end is not None:
        matched = {"gpu_mean_penetration_m": meas["mean_penetration_m"],
                   "reference_mean_penetration_m": ref_at_end["mean_penetration_m"],
                   "bound": "gpu mean <= reference mean + 0.005 m at the same step",
                   "ok": meas["mean_penetration_m"] <= ref_at_end["mean_penetration_m"] + 0.005}
    out = {
        "metric": "body-steps/sec + vel. iters to 1e-4 residual, plain vs learned warm-start",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"random pile, {args.bodies} bodies, seed {args.seed}, {total}-step run from a fresh "
                               f"load: steps {prefix + 1}-{total} timed, steps 1-{prefix} untimed before them; "
                               f"{args.vel} vel / {args.pos} pos iterations, coloured solver, sleep and CCD off",
                   "bodies": args.bodies, "vel_iters": args.vel, "pos_iters": args.pos, "solver": "coloured",
                   "window_steps": total, "timed_steps": f"{prefix + 1}-{total}"},
        "full_window": {"value": args.bodies * total / (hl["prefix_s"] + wall), "unit": "body-steps/s",
                        "ms_per_step": 1e3 * (hl["prefix_s"] + wall) / total, "steps": f"1-{total}"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "state_at_end": dict(meas, step=total, constraints=st.constraints, points=st.points, colours=st.colours),
        "reference_at_end": ref_at_end,
        "matched_penetration": matched,
        "schedule": sched_dict(_schedule),
        "pyramid": pyramid,
        "learned_warm_start": learned,
        "learned_warm_start_every_point": learned_all,
        "batched_worlds": batched,
        "stress": stressed,
        "continuous_collision": ccd,
    }
    out.update(extra)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

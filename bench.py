#!/usr/bin/env python3
"""Benchmark: body-steps/s of b2World::Step on the random pile, one MI355X.

Workload: the random pile of 100k dynamic bodies (50% circles r~U[0.25,0.75],
50% convex polygons of 3-8 points on a circle of R~U[0.3,0.8]), density 1,
friction 0.6, restitution 0, dropped from a 300x600 m grid into a static U
container; dt 1/60, 8 velocity / 3 position iterations, sleeping and
continuous collision off on both sides. A step is one b2World::Step over the
whole pile; the timed region is steps 1..K of a freshly loaded pile (all
state already in HBM), after W warm-up steps on a separate copy.

Also reported on the same JSON line:
  roofline      the dominant kernel (velocity solve) against HBM bandwidth:
                algorithmic bytes per launch from the byte model
                (48 B per constraint + 72 B per manifold point per velocity
                iteration) over its mean execution time, measured in a second
                pass over the same workload (fresh pile, same steps) with HIP
                events that its dispatch packets record on the step stream
                (hipExtLaunchKernelGGL); traffic is the HBM bytes per launch
                measured by rocprofv3 counters (given with --traffic-bytes);
                stream_us_per_launch is the time per launch seen on the
                stream, dispatch gaps included;
  cpu_baseline  the reference b2World::Step (upstream Box2D 2.3.2 built from
                the reference's sources, oracle/_ref) on the same pile on one
                host core, over the first steps that fit the time budget.
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
# HBM bytes per k_solve_velocity launch on the default workload (100k pile,
# seed 1, steps 1-1000, 8/3 iterations), from rocprofv3 FETCH_SIZE and
# WRITE_SIZE averaged over all launches and scaled by the counters' readings
# for a plain 256 MiB copy (x2.000 and x1.000); the summary with the commands
# is under profiles/.
DEFAULT_TRAFFIC_BYTES = 4.929e6


class PhaseProfile(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 24), ("launches", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("ms", ctypes.c_double)]


class StepStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "bodies", "fixtures", "contacts", "constraints", "points", "colours", "colouringRounds", "newPairs",
        "positionIterations", "beginEvents", "endEvents", "overflow")] + [("minSeparation", ctypes.c_float),
                                                                           ("cellSize", ctypes.c_float)]


def make_world(lib, n, seed):
    h = ctypes.c_void_p(lib.b2pp_world_create(0.0, -10.0))
    lib.b2pp_world_set_flags(h, 0, 1, 0, 0)
    rc = lib.b2pp_world_load_scene(h, b"pile", n, ctypes.c_uint(seed))
    if rc != 0:
        raise RuntimeError(f"scene load failed: {lib.b2pp_world_last_error(h)}")
    lib.b2pp_world_sync(h)
    return h


def step(lib, h, k, vel, pos):
    rc = lib.b2pp_world_step(h, ctypes.c_float(1.0 / 60.0), vel, pos, k)
    if rc != 0:
        raise RuntimeError(f"step failed rc={rc}: {lib.b2pp_world_last_error(h).decode()}")


def stats_of(lib, h):
    st = StepStats()
    lib.b2pp_world_stats(h, ctypes.byref(st))
    return st


def cpu_reference(n, seed, vel, pos, budget_s):
    """Reference b2World::Step on one core: steps 1..k of the same pile."""
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
    h = ctypes.c_void_p(f("world_create")(0.0, -10.0))
    f("world_set_flags")(h, 0, 1, 0, 0)
    f("world_load_scene")(h, b"pile", n, ctypes.c_uint(seed))
    steps = 0
    t0 = time.perf_counter()
    while True:
        f("world_step")(h, ctypes.c_float(1.0 / 60.0), vel, pos, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 1000:
            break
    f("world_destroy")(h)
    return {"value": n * steps / el, "unit": "body-steps/s", "cores": 1, "kind": kind,
            "sample": f"steps 1-{steps} of 1000 of the same {n}-body pile (seed {seed}), one host thread, "
                      f"{el:.1f} s; early steps are cheaper than the settled pile"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--bodies", type=int, default=100000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--vel", type=int, default=8)
    ap.add_argument("--pos", type=int, default=3)
    ap.add_argument("--profile-steps", type=int, default=None,
                    help="steps of the timing pass (default: the same as --steps)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-bytes", type=float, default=None,
                    help="HBM bytes per velocity-solve launch measured with rocprofv3 FETCH_SIZE + WRITE_SIZE")
    args = ap.parse_args()
    if args.profile_steps is None:
        args.profile_steps = args.steps
    default_workload = (args.bodies, args.seed, args.steps, args.vel, args.pos) == (100000, 1, 1000, 8, 3)
    if args.traffic_bytes is None and default_workload:
        args.traffic_bytes = DEFAULT_TRAFFIC_BYTES
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(batched worlds shard across GPUs by world; a single world by slab, "
                                   "see DESIGN.md)"}))
        return 2

    lib = box2dpp_amd.host()
    # warm-up on a separate copy of the scene
    if args.warmup > 0:
        hw = make_world(lib, args.bodies, args.seed)
        step(lib, hw, args.warmup, args.vel, args.pos)
        lib.b2pp_world_sync(hw)
        lib.b2pp_world_destroy(hw)

    h = make_world(lib, args.bodies, args.seed)
    lib.b2pp_world_sync(h)
    t0 = time.perf_counter()
    step(lib, h, args.steps, args.vel, args.pos)
    lib.b2pp_world_sync(h)
    t1 = time.perf_counter()
    wall = t1 - t0
    value = args.bodies * args.steps / wall
    st = stats_of(lib, h)
    pen = np.zeros(3, np.float64)
    lib.b2pp_world_penetration(h, pen.ctypes.data_as(ctypes.c_void_p))

    lib.b2pp_world_destroy(h)

    # Timing pass: the same workload again from a fresh pile, with per-phase
    # events on the step stream and the velocity kernel's own execution time
    # from events carried by its dispatch packets.
    b2g = box2dpp_amd.b2g()
    h = make_world(lib, args.bodies, args.seed)
    ctx = ctypes.c_void_p(lib.b2pp_world_device_context(h))
    b2g.b2g_set_profiling(ctx, 1)
    prof_bytes = 0.0
    for _ in range(args.profile_steps):
        step(lib, h, 1, args.vel, args.pos)
        s2 = stats_of(lib, h)
        prof_bytes += args.vel * (48.0 * s2.constraints + 72.0 * s2.points)
    prof = (PhaseProfile * 32)()
    cnt = ctypes.c_int(0)
    b2g.b2g_get_profile(ctx, prof, 32, ctypes.byref(cnt))
    b2g.b2g_set_profiling(ctx, 0)
    entries = {prof[i].name.decode(): (prof[i].launches, prof[i].ms) for i in range(min(cnt.value, 32))}
    phases = {k: v for k, v in entries.items() if not k.startswith("k_")}
    kernels = {k: v for k, v in entries.items() if k.startswith("k_")}
    # kernel execution time from events carried by the dispatch packets
    sv_launches, sv_ms = kernels.get("k_solve_velocity", (0, 0.0))
    avg_launch_ms = sv_ms / max(sv_launches, 1)
    ph_launches, ph_ms = phases.get("solve_velocity", (0, 0.0))
    bytes_per_launch = prof_bytes / max(sv_launches, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0
    lib.b2pp_world_destroy(h)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_reference(args.bodies, args.seed, args.vel, args.pos, args.cpu_seconds)

    total_phase_ms = sum(ms for _, ms in phases.values())
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
        "config": {"workload": f"random pile, {args.bodies} bodies, seed {args.seed}, steps 1-{args.steps}, "
                               f"{args.vel} vel / {args.pos} pos iterations, coloured solver, sleep and CCD off",
                   "bodies": args.bodies, "vel_iters": args.vel, "pos_iters": args.pos, "solver": "coloured"},
        "roofline": {"bound": "hbm", "kernel": "k_solve_velocity", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": args.traffic_bytes,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_launch_ms * 1e3,
                     "launches_per_step": sv_launches / max(args.profile_steps, 1),
                     "stream_us_per_launch": 1e3 * ph_ms / max(ph_launches, 1)},
        "cpu_baseline": cpu,
        "final_state": {"contacts": st.contacts, "constraints": st.constraints, "points": st.points,
                        "colours": st.colours, "mean_penetration_m": float(pen[0]), "max_penetration_m": float(pen[1])},
        "phase_ms_per_step": {k: round(v[1] / max(args.profile_steps, 1), 4) for k, v in phases.items()},
        "phase_share": {k: round(v[1] / total_phase_ms, 4) for k, v in phases.items()} if total_phase_ms else {},
        "kernel_us_per_launch": {k: round(1e3 * v[1] / max(v[0], 1), 3) for k, v in kernels.items()},
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

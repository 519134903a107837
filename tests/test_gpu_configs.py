"""The benchmarked workloads at their own size, with the coloured solver (the
path every throughput figure is measured on), held to the reference.

The coloured solver visits a pass's constraints by colour instead of in
Box2D's island order, so its trajectories leave the reference's after the
first contacts; it is held to bounds on what the order does not change. All
state measures are box2dpp_amd.measures (FP64), applied alike to every
engine's post-step state:

* random pile, 100 000 bodies (seed 1, 8/3 iterations, sleeping and
  continuous collision off), 1000 steps, against the reference build's
  statistics of the same pile (tests/golden/pile_ref_stats.json, written by
  tests/golden/make_pile_stats.py) at steps 100, 250, 500 and 1000:
  - mean penetration at most the reference's + 0.5 cm (the matched
    penetration of the benchmark);
  - kinetic energy within 10% of the reference's while the pile falls
    (steps 100-500: free fall of the bodies still in the air dominates it and
    does not depend on the solver's order) and within [0.25, 2] x the
    reference's at step 1000 (the settling pile: sliding and collapses
    differ between two valid visit orders; a broken solver leaves
    metres-per-second motion, orders of magnitude above the band);
  - RMS complementarity residual at most 1.5 x the reference's while the
    pile falls and 2 x at step 1000 (the residual after 8 Gauss-Seidel
    passes depends on the visit order; a solver that stopped converging
    would leave the approach speeds, ~1 m/s, at step 1000);
  - contact count at step 1000 within 15% of the reference's.
* stress world at one world's share reduced (a pile of 3000 or 6000 bodies
  dropped on jointed chains of 32 links), against the oracle (CPU
  restatement, the reference's order) over 300 steps: penetration at most
  the oracle's + 0.5 cm, kinetic energy within 5%, residual at most 2 x the
  oracle's + 0.05 m/s; and the stress configuration at full size (4 worlds
  of a 125 000-body pile on 2500 chains each: 820 004 bodies, 310 000
  joints), 600 steps: finite state, mean penetration under 20 cm, mean
  hinge gap under 1 cm.
* learned warm start at its configuration: logged from the plain solver on a
  100 000-body pile (seed 2, 1000 steps, 10 M fresh points), trained by the
  project's kernels, evaluated on an unseen 100 000-body pile (seed 1001):
  after 2 and 4 velocity iterations its residual is below the plain warm
  start's, and after 8 within 2% of it, averaged over the traced steps of
  the settled pile; the learned step keeps the mean penetration within
  0.5 cm of the plain step's.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np
import pytest

import engines as E

pytestmark = pytest.mark.gpu

MLP_PARAM_FLOATS = 64 * 16 + 64 + 2 * (64 * 64 + 64) + 2 * 64 + 2 + 16 + 16 + 2 + 2
STATS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pile_ref_stats.json")


def test_pile100k_coloured_against_reference(gpu):
    ref = json.load(open(STATS))["pile100k"]
    assert ref["bodies"] == 100000 and ref["seed"] == 1
    print("case: pile 100000 bodies, seed 1, 1000 steps, coloured solver", flush=True)
    g = gpu.world("pile", 100000, 1, sleep=False)
    done = 0
    for step in (100, 250, 500, 1000):
        g.step(step - done, vel=8, pos=3)
        done = step
        m = g.measures()
        r = ref["at"][str(step)]
        msg = f"step {step}: gpu {m}, reference {r}"
        assert np.isfinite(g.bodies()).all(), msg
        assert m["mean_penetration_m"] <= r["mean_penetration_m"] + 0.005, msg
        if step < 1000:
            assert abs(m["kinetic_energy_J"] / r["kinetic_energy_J"] - 1.0) <= 0.10, msg
            assert m["residual_rms_mps"] <= 1.5 * r["residual_rms_mps"], msg
        else:
            assert 0.25 <= m["kinetic_energy_J"] / r["kinetic_energy_J"] <= 2.0, msg
            assert m["residual_rms_mps"] <= 2.0 * r["residual_rms_mps"], msg
            contacts = gpu.fn("world_contact_count")(g.h)
            assert abs(contacts / r["contacts"] - 1.0) <= 0.15, (contacts, r["contacts"])
    g.close()


@pytest.mark.parametrize("n,seed", [(3000, 11), (6000, 5)])
def test_stress_world_against_oracle(gpu, orc, n, seed):
    g = gpu.world("stress", n, seed, sleep=False)
    o = orc.world("stress", n, seed, sleep=False)
    print(f"case: stress world, pile of {n} bodies, seed {seed}", flush=True)
    for step in (100, 200, 300):
        g.step(100)
        o.step(100)
        mg, mo = g.measures(), o.measures()
        msg = f"step {step}: gpu {mg}, oracle {mo}"
        assert mg["mean_penetration_m"] <= mo["mean_penetration_m"] + 0.005, msg
        assert abs(mg["kinetic_energy_J"] / mo["kinetic_energy_J"] - 1.0) <= 0.05, msg
        assert mg["residual_rms_mps"] <= 2.0 * mo["residual_rms_mps"] + 0.05, msg
    g.close()
    o.close()


def test_stress_config_full_size(gpu):
    lib = E.b2g_lib()
    print("case: stress configuration, 4 worlds x (125000 bodies + 2500 chains)", flush=True)
    w = E.World.worlds(gpu, "stress", 4, 125000, 1, sleep=False)
    nb = gpu.fn("world_body_count")(w.h)
    nj = gpu.fn("world_joint_count")(w.h)
    assert nb == 820004 and nj == 310000
    for _ in range(6):
        w.step(100)
    ctx = w.device_context()
    pen = np.zeros(3, np.float64)
    gpu.fn("world_penetration")(w.h, E.ptr(pen))
    drift = np.zeros(3, np.float64)
    E.b2g_check(lib, ctx, lib.b2g_joint_drift(ctx, E.ptr(drift)))
    assert np.isfinite(w.bodies()).all()
    assert pen[2] > 100000 and pen[0] < 0.20, pen
    assert drift[2] == nj and drift[0] < 0.01, drift
    w.close()


def _trace(lib, ctx, iters):
    cur = np.zeros(3 * (iters + 1), np.float64)
    it = ctypes.c_int(0)
    E.b2g_check(lib, ctx, lib.b2g_get_trace(ctx, E.ptr(cur), len(cur), ctypes.byref(it)))
    assert it.value == iters
    return cur.reshape(3, iters + 1)


def test_learned_warm_start_config(gpu):
    lib = E.b2g_lib()
    print("case: learned warm start, 100000-body piles (log seed 2, evaluation seed 1001)", flush=True)
    lw = gpu.world("pile", 100000, 2, sleep=False)
    c0 = lw.device_context()
    E.b2g_check(lib, c0, lib.b2g_log_fresh(c0, 10_000_000))
    lw.step(1000)
    n = ctypes.c_int(0)
    E.b2g_check(lib, c0, lib.b2g_log_size(c0, ctypes.byref(n)))
    assert n.value >= 5_000_000, n.value
    E.b2g_check(lib, c0, lib.b2g_mlp_init(c0, 5))
    losses = np.zeros(20, np.float64)
    E.b2g_check(lib, c0, lib.b2g_mlp_train(c0, 20, 65536, ctypes.c_float(2e-3), 9, E.ptr(losses)))
    assert losses[-1] < losses[0]
    params = np.zeros(MLP_PARAM_FLOATS, np.float32)  # b2g_mlp_params
    E.b2g_check(lib, c0, lib.b2g_mlp_get_params(c0, E.ptr(params)))
    lw.close()

    ev = gpu.world("pile", 100000, 1001, sleep=False)
    c = ev.device_context()
    E.b2g_check(lib, c, lib.b2g_mlp_set_params(c, E.ptr(params)))
    curves = []
    for s in range(1, 1001):
        traced = s >= 500 and s % 100 == 0
        if traced:
            E.b2g_check(lib, c, lib.b2g_set_trace(c, 8))
        ev.step(1)
        if traced:
            E.b2g_check(lib, c, lib.b2g_set_trace(c, 0))
            curves.append(_trace(lib, c, 8))
    plain_pen = ev.penetration()[0]
    ev.close()
    cur = np.array(curves).mean(axis=0)
    cold, warm, learned = cur
    print("cold", cold, "\nwarm", warm, "\nlearned", learned)
    assert learned[2] < warm[2] and learned[4] < warm[4], (warm, learned)
    assert learned[8] <= 1.02 * warm[8], (warm, learned)

    ls = gpu.world("pile", 100000, 1001, sleep=False)
    c = ls.device_context()
    E.b2g_check(lib, c, lib.b2g_mlp_set_params(c, E.ptr(params)))
    ls.set_learned(True, 0.0)
    ls.step(1000)
    learned_pen = ls.penetration()[0]
    assert np.isfinite(ls.bodies()).all()
    assert learned_pen <= plain_pen + 0.005, (learned_pen, plain_pen)
    ls.close()

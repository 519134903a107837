"""Writes tests/golden/pile_ref_stats.json: post-step measures of the random
pile stepped by the reference build (oracle/_ref/libb2ref.so, upstream
Box2D 2.3.2 compiled from the reference's own sources by oracle/Makefile).

Workloads (scene "pile" of box2dpp_amd/scenes/scenes.h, 8/3 iterations,
dt 1/60, sleep and continuous collision off, warm starting on):
  pile100k  100 000 bodies, seed 1, measured after steps 100, 250, 500, 1000
  pile10k   10 000 bodies, seed 1, measured after steps 100, 250, 500, 1000

Measures (box2dpp_amd.measures, FP64, the same function every engine's
state is measured with): mean and max penetration, touching points, RMS
complementarity residual, kinetic energy; plus contact count and wall time.
The 100k run takes about 13 minutes on one core.

Run from the repository root after building oracle/_ref:
    python tests/golden/make_pile_stats.py [pile10k] [pile100k]
"""
from __future__ import annotations

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import engines as E  # noqa: E402

WORKLOADS = {"pile10k": (10000, 1), "pile100k": (100000, 1)}
STEPS = (100, 250, 500, 1000)


def run(name: str) -> dict:
    n, seed = WORKLOADS[name]
    ref = E.Engine("ref")
    w = ref.world("pile", n, seed, sleep=False, ccd=False)
    out = {"bodies": n, "seed": seed, "vel_iters": 8, "pos_iters": 3, "sleep": False, "ccd": False, "at": {}}
    done, t0 = 0, time.perf_counter()
    for s in STEPS:
        w.step(s - done, vel=8, pos=3)
        done = s
        m = w.measures()
        m["contacts"] = int(ref.fn("world_contact_count")(w.h))
        m["wall_s"] = round(time.perf_counter() - t0, 2)
        out["at"][str(s)] = m
        print(name, s, m, flush=True)
    w.close()
    return out


def main(names) -> None:
    path = os.path.join(HERE, "pile_ref_stats.json")
    data = json.load(open(path)) if os.path.exists(path) else {}
    for name in names:
        data[name] = run(name)
        with open(path, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    print(f"wrote {path}")


if __name__ == "__main__":
    main(sys.argv[1:] or list(WORKLOADS))

"""The C ABI library: it loads without a GPU, exports every function that
include/b2g.h declares, and the struct layouts the tests, the benchmark and
the oracle bind match the library's (b2g_abi_sizes)."""
from __future__ import annotations

import ctypes
import os
import re
import sys

import numpy as np

import engines as E

sys.path.insert(0, os.path.join(E.ROOT, "oracle"))
sys.path.insert(0, E.ROOT)


def _lib():
    return ctypes.CDLL(E.GPU_B2G_LIB)


def test_exports_every_declared_function():
    header = open(os.path.join(E.ROOT, "include", "b2g.h")).read()
    names = sorted(set(re.findall(r"^(?:int|const char\*)\s+(b2g_[a-z0-9_]+)\(", header, re.M)))
    assert len(names) > 30
    lib = _lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_the_library():
    import mlp_ref
    import bench
    lib = _lib()
    n = lib.b2g_abi_sizes(None, 0)
    sizes = np.zeros(n, np.int32)
    lib.b2g_abi_sizes(E.ptr(sizes), n)
    shape, manifold, body, fixture, world, params, stats, profile, mlp, aabb, xf, revolute, schedule, sweep, distance, \
        rope, weld, friction, motor, pulley, mouse, prismatic, wheel, gear = sizes[:24]
    assert E.SHAPE.itemsize == shape
    assert E.MANIFOLD.itemsize == manifold
    assert E.STATS.itemsize == stats
    assert E.AABB.itemsize == aabb
    assert E.XF.itemsize == xf
    assert mlp_ref.PARAMS_DTYPE.itemsize == mlp
    assert ctypes.sizeof(bench.StepStats) == stats
    assert ctypes.sizeof(bench.PhaseProfile) == profile
    assert bench.MLP_PARAM_FLOATS * 4 == mlp
    assert body == 26 * 4  # 2 int32 + 24 float
    assert revolute == 56  # 2 body indices, 9 floats, 3 switches
    assert schedule == E.SCHEDULE.itemsize
    assert sweep == E.SWEEP.itemsize == 36
    assert distance == 40  # 2 body indices, 7 floats, 1 switch
    assert rope == 32      # 2 body indices, 5 floats, 1 switch
    assert weld == 40      # 2 body indices, 7 floats, 1 switch
    assert friction == 36  # 2 body indices, 6 floats, 1 switch
    assert motor == 36     # 2 body indices, 6 floats, 1 switch
    assert pulley == 56    # 2 body indices, 11 floats, 1 switch
    assert mouse == 44     # 2 body indices, 8 floats, 1 switch
    assert prismatic == 64 # 2 body indices, 11 floats, 3 switches
    assert wheel == 56     # 2 body indices, 10 floats, 2 switches
    assert gear == 92      # 4 body indices, 2 kinds, 16 floats, 1 switch
    assert params == E.STEP_PARAMS.itemsize

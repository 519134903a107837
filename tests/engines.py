"""Uniform access to the three Box2D step engines used by the tests.

  ref  oracle/_ref/libb2ref.so   upstream Box2D 2.3.2 built from the reference's sources
  orc  oracle/_build/liborc.so   the CPU restatement (the oracle)
  gpu  box2dpp_amd/lib/libbox2d.so  the product: box2d:: API over the gfx950 step (C ABI b2g.h)

All three export the same world entry points with their own prefix
(ref_/orc_/b2pp_), so a test builds the same scene in each and compares.
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libb2ref.so")
ORC_LIB = os.path.join(ROOT, "oracle", "_build", "liborc.so")
GPU_HOST_LIB = os.path.join(ROOT, "box2dpp_amd", "lib", "libbox2d.so")
GPU_B2G_LIB = os.path.join(ROOT, "box2dpp_amd", "lib", "libb2g.so")

VP = ctypes.c_void_p

# ---------------------------------------------------------------- C layouts
VEC2 = np.dtype([("x", "<f4"), ("y", "<f4")])
SHAPE = np.dtype([
    ("type", "<i4"), ("count", "<i4"), ("radius", "<f4"), ("flags", "<i4"),
    ("centroid", VEC2), ("v0", VEC2), ("v3", VEC2), ("pad", VEC2),
    ("verts", VEC2, (8,)), ("normals", VEC2, (8,)),
])
assert SHAPE.itemsize == 176
XF = np.dtype([("px", "<f4"), ("py", "<f4"), ("s", "<f4"), ("c", "<f4")])
SWEEP = np.dtype([("localCenter", "<f4", 2), ("c0", "<f4", 2), ("c", "<f4", 2), ("a0", "<f4"), ("a", "<f4"),
                  ("alpha0", "<f4")])
MPOINT = np.dtype([("lp", VEC2), ("ni", "<f4"), ("ti", "<f4"), ("id", "<u4")])
MANIFOLD = np.dtype([("points", MPOINT, (2,)), ("ln", VEC2), ("lpt", VEC2), ("type", "<i4"), ("count", "<i4")])
assert MANIFOLD.itemsize == 64
AABB = np.dtype([("lx", "<f4"), ("ly", "<f4"), ("ux", "<f4"), ("uy", "<f4")])
STATS = np.dtype([
    ("bodies", "<i4"), ("fixtures", "<i4"), ("contacts", "<i4"), ("constraints", "<i4"), ("points", "<i4"),
    ("colours", "<i4"), ("colouringRounds", "<i4"), ("newPairs", "<i4"), ("positionIterations", "<i4"),
    ("beginEvents", "<i4"), ("endEvents", "<i4"), ("overflow", "<i4"), ("minSeparation", "<f4"),
    ("cellSize", "<f4"), ("freshPoints", "<i4"), ("fallback", "<i4"), ("residual", "<f4"), ("toiEvents", "<i4"),
    ("queries", "<i4"), ("scanned", "<i4"), ("wideConstraints", "<i4"),
])

SCHEDULE = np.dtype([("islands", "<i4"), ("collideOrder", "<i4"), ("solverBlock", "<i4"), ("collideBlock", "<i4"),
                     ("queryBlock", "<i4"), ("tailColours", "<i4"), ("clusterBodies", "<i4"),
                     ("persistentSolve", "<i4")])
STEP_PARAMS = np.dtype([("dt", "<f4"), ("velocityIterations", "<i4"), ("positionIterations", "<i4"), ("mode", "<i4"),
                        ("learnedWarmStart", "<i4"), ("residualThreshold", "<f4")])

# body row columns of *_world_get_bodies
PX, PY, QS, QC, CX, CY, ANG, VX, VY, W, AWAKE, SLEEP = range(12)


def ptr(a: np.ndarray) -> VP:
    return a.ctypes.data_as(VP)


def have(path: str) -> bool:
    return os.path.exists(path)


class Engine:
    """One step engine behind the shared world entry points."""

    def __init__(self, name: str):
        self.name = name
        if name == "ref":
            self.lib, self.pre = ctypes.CDLL(REF_LIB), "ref_"
        elif name == "orc":
            self.lib, self.pre = ctypes.CDLL(ORC_LIB), "orc_"
        elif name == "gpu":
            ctypes.CDLL(GPU_B2G_LIB, mode=ctypes.RTLD_GLOBAL)
            self.lib, self.pre = ctypes.CDLL(GPU_HOST_LIB), "b2pp_"
        else:
            raise ValueError(name)
        self.fn("world_create").restype = VP
        self.fn("world_create").argtypes = [ctypes.c_float, ctypes.c_float]
        if name == "gpu":
            self.fn("world_last_error").restype = ctypes.c_char_p

    def fn(self, n: str):
        return getattr(self.lib, self.pre + n)

    def world(self, kind: str, n: int = 0, seed: int = 0, sleep: bool = True, warm: bool = True,
              gravity=(0.0, -10.0), ccd: bool = False, substep: bool = False) -> "World":
        return World(self, kind, n, seed, sleep, warm, gravity, ccd, substep)


class World:
    def __init__(self, eng: Engine, kind: str, n: int, seed: int, sleep: bool, warm: bool, gravity,
                 ccd: bool = False, substep: bool = False):
        self.e = eng
        self.h = VP(eng.fn("world_create")(gravity[0], gravity[1]))
        eng.fn("world_set_flags")(self.h, int(sleep), int(warm), int(ccd), int(substep))
        if kind is None:
            return
        rc = eng.fn("world_load_scene")(self.h, kind.encode(), int(n), ctypes.c_uint(seed))
        if rc != 0:
            raise RuntimeError(f"{eng.name}: load {kind} failed rc={rc} {self.error()}")

    @classmethod
    def batch(cls, eng: "Engine", n_worlds: int, seed: int, lo: int, hi: int, sleep: bool = False):
        """Batched worlds in sub-worlds of one GPU world; returns (world,
        dynamic body count per sub-world)."""
        w = cls(eng, None, 0, 0, sleep, True, (0.0, -10.0))
        sizes = np.zeros(n_worlds, np.int32)
        rc = eng.lib.b2pp_world_load_batch(w.h, int(n_worlds), ctypes.c_uint(seed), int(lo), int(hi), ptr(sizes))
        if rc != 0:
            raise RuntimeError(f"batch load failed rc={rc} {w.error()}")
        return w, sizes

    @classmethod
    def worlds(cls, eng: "Engine", kind: str, n_worlds: int, n: int, seed: int, sleep: bool = False):
        """n_worlds copies of a named scene (world w from seed + w) in
        sub-worlds of one GPU world, joints included."""
        w = cls(eng, None, 0, 0, sleep, True, (0.0, -10.0))
        rc = eng.lib.b2pp_world_load_worlds(w.h, kind.encode(), int(n_worlds), int(n), ctypes.c_uint(seed))
        if rc != 0:
            raise RuntimeError(f"worlds load failed rc={rc} {w.error()}")
        return w

    def joints(self) -> np.ndarray:
        """Per joint: impulse x, y, z and motor impulse."""
        nj = self.e.fn("world_joint_count")(self.h)
        out = np.zeros(4 * nj, np.float32)
        self.e.fn("world_get_joints")(self.h, ptr(out))
        return out.reshape(-1, 4)

    def error(self) -> str:
        if self.e.name == "gpu":
            return self.e.fn("world_last_error")(self.h).decode()
        return ""

    def set_mode(self, mode: int) -> None:
        self.e.fn("world_set_mode")(self.h, int(mode))

    def step(self, n: int = 1, dt: float = 1.0 / 60.0, vel: int = 8, pos: int = 3) -> None:
        rc = self.e.fn("world_step")(self.h, ctypes.c_float(dt), int(vel), int(pos), int(n))
        if self.e.name != "ref" and rc not in (0, None):
            raise RuntimeError(f"{self.e.name}: step failed rc={rc} {self.error()}")

    def sync(self) -> None:
        if self.e.name == "gpu":
            self.e.fn("world_sync")(self.h)

    def bodies(self) -> np.ndarray:
        nb = self.e.fn("world_body_count")(self.h)
        out = np.zeros((nb, 12), np.float32)
        self.e.fn("world_get_bodies")(self.h, ptr(out))
        return out

    def contacts(self):
        """(meta[int32 n x 4], manifolds[MANIFOLD n]) in the engine's own order."""
        n = self.e.fn("world_contact_count")(self.h)
        meta = np.zeros((max(n, 1), 4), np.int32)
        man = np.zeros(max(n, 1), MANIFOLD)
        got = self.e.fn("world_get_contacts")(self.h, ptr(meta), ptr(man), n)
        n = min(n, got)
        return meta[:n], man[:n]

    def fat_aabbs(self) -> np.ndarray:
        nf = self.e.fn("world_fixture_count")(self.h)
        out = np.zeros(nf, AABB)
        self.e.fn("world_get_fat_aabbs")(self.h, ptr(out))
        return out

    def penetration(self):
        out = np.zeros(3, np.float64)
        self.e.fn("world_penetration")(self.h, ptr(out))
        return float(out[0]), float(out[1]), int(out[2])

    def fixture_info(self):
        """(body index per fixture proxy, (radius, friction, restitution) per proxy)."""
        nf = self.e.fn("world_fixture_count")(self.h)
        body = np.zeros(max(nf, 1), np.int32)
        mat = np.zeros((max(nf, 1), 3), np.float32)
        self.e.fn("world_fixture_info")(self.h, ptr(body), ptr(mat))
        return body[:nf], mat[:nf]

    def body_mass(self) -> np.ndarray:
        """(mass, rotational inertia about the centre) per body."""
        nb = self.e.fn("world_body_count")(self.h)
        out = np.zeros((max(nb, 1), 2), np.float32)
        self.e.fn("world_body_mass")(self.h, ptr(out))
        return out[:nb]

    def measures(self) -> dict:
        """Penetration, complementarity residual and kinetic energy of the
        current state, FP64 (box2dpp_amd.measures), the same function for
        every engine."""
        from box2dpp_amd import measures
        fb, fm = self.fixture_info()
        meta, man = self.contacts()
        return measures.all_measures(self.bodies(), meta, man, fb, fm, self.body_mass())

    def device_context(self) -> VP:
        """The world's b2g_context (C ABI of include/b2g.h)."""
        self.e.lib.b2pp_world_device_context.restype = VP
        return VP(self.e.lib.b2pp_world_device_context(self.h))

    def set_learned(self, on: bool, threshold: float = 0.0) -> None:
        self.e.lib.b2pp_world_set_learned(self.h, int(on), ctypes.c_float(threshold))

    def stats(self):
        if self.e.name != "gpu":
            return None
        out = np.zeros(1, STATS)
        self.e.fn("world_stats")(self.h, ptr(out))
        return out[0]

    def close(self) -> None:
        if self.h:
            self.e.fn("world_destroy")(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def contact_table(meta: np.ndarray, man: np.ndarray):
    """Contacts keyed by (fixtureA, fixtureB), sorted, for order-free comparison."""
    order = np.lexsort((meta[:, 1], meta[:, 0]))
    return meta[order], man[order]


# ---------------------------------------------------------------- shapes
def oracle():
    lib = ctypes.CDLL(ORC_LIB)
    lib.orc_find_max_separation.restype = ctypes.c_float
    return lib


def reference():
    lib = ctypes.CDLL(REF_LIB)
    lib.ref_find_max_separation.restype = ctypes.c_float
    lib.ref_bp_create.restype = VP
    return lib


def polygon_from_points(lib, pre: str, pts: np.ndarray) -> np.ndarray:
    out = np.zeros(1, SHAPE)
    p = np.ascontiguousarray(pts, np.float32).reshape(-1)
    getattr(lib, pre + "polygon_set")(ptr(p), len(p) // 2, ptr(out))
    return out[0]


def box(lib, pre: str, hx: float, hy: float, cx=0.0, cy=0.0, angle=0.0, rotated=False) -> np.ndarray:
    out = np.zeros(1, SHAPE)
    getattr(lib, pre + "polygon_box")(ctypes.c_float(hx), ctypes.c_float(hy), ctypes.c_float(cx),
                                      ctypes.c_float(cy), ctypes.c_float(angle), int(rotated), ptr(out))
    return out[0]


def circle(r: float, cx=0.0, cy=0.0) -> np.ndarray:
    s = np.zeros(1, SHAPE)[0]
    s["type"] = 0
    s["radius"] = r
    s["centroid"] = (cx, cy)
    return s


def edge(v1, v2, v0=None, v3=None) -> np.ndarray:
    s = np.zeros(1, SHAPE)[0]
    s["type"] = 1
    s["count"] = 2
    s["radius"] = np.float32(2.0) * np.float32(0.005)
    s["verts"][0] = tuple(v1)
    s["verts"][1] = tuple(v2)
    flags = 0
    if v0 is not None:
        s["v0"] = tuple(v0)
        flags |= 1
    if v3 is not None:
        s["v3"] = tuple(v3)
        flags |= 2
    s["flags"] = flags
    return s


def transform(x: float, y: float, angle: float) -> np.ndarray:
    t = np.zeros(1, XF)[0]
    t["px"], t["py"] = x, y
    t["s"], t["c"] = np.sin(np.float32(angle)), np.cos(np.float32(angle))
    return t


def manifold_bits(m: np.ndarray) -> np.ndarray:
    """Manifold bytes with unused point slots zeroed (engines leave them stale)."""
    m = m.copy()
    for i in range(len(m)):
        c = int(m[i]["count"])
        for k in range(c, 2):
            m[i]["points"][k] = np.zeros(1, MPOINT)[0]
        if c == 0:
            m[i] = np.zeros(1, MANIFOLD)[0]
    return m.view(np.uint8).reshape(len(m), MANIFOLD.itemsize)


def geometry_bits(m: np.ndarray) -> np.ndarray:
    """Manifold bytes without the impulse fields."""
    m = m.copy()
    m["points"]["ni"] = 0
    m["points"]["ti"] = 0
    return manifold_bits(m)


# ---------------------------------------------------------------- C ABI
def b2g_lib():
    """libb2g.so with the return types of the calls the tests make."""
    lib = ctypes.CDLL(GPU_B2G_LIB, mode=ctypes.RTLD_GLOBAL)
    lib.b2g_last_error.restype = ctypes.c_char_p
    return lib


def b2g_check(lib, ctx, rc) -> None:
    if rc != 0:
        raise RuntimeError(f"b2g call failed rc={rc}: {lib.b2g_last_error(ctx).decode()}")

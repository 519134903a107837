"""The drop-in boundary from C++: the reference's HelloWorld
(HelloWorld/HelloWorld.cpp:30-101), with only its include line changed,
compiles against include/box2d and links libbox2d.so; on the GPU its output
is the reference build's trajectory (committed fixture hello_ccd_bodies:
6/2 iterations, sleeping and continuous collision on, the reference's
defaults) within 1e-5.

The program is tests/cpp/hello_world.cpp, and INTEGRATION.md shows the same
text, which the first test checks. tests/cpp/body_api.cpp exercises the
b2Body calls beyond HelloWorld's (contact lists, mass and rotation setters).
"""
from __future__ import annotations

import os
import re
import subprocess

import numpy as np
import pytest

import engines as E

SRC = os.path.join(E.ROOT, "tests", "cpp", "hello_world.cpp")
LIBDIR = os.path.join(E.ROOT, "box2dpp_amd", "lib")


BODY_API = os.path.join(E.ROOT, "tests", "cpp", "body_api.cpp")


def _compile(out: str, *defines: str, src: str = SRC) -> None:
    cmd = ["g++", "-std=c++14", "-Wall", "-Werror", f"-I{E.ROOT}/include", *defines, src, f"-L{LIBDIR}",
           "-lbox2d", "-lb2g", f"-Wl,-rpath,{LIBDIR}", "-o", out]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert res.returncode == 0, res.stdout


def test_integration_sample_is_the_compiled_program():
    text = open(os.path.join(E.ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, re.S)
    assert open(SRC).read() in blocks, "INTEGRATION.md's HelloWorld differs from tests/cpp/hello_world.cpp"


def test_hello_world_compiles_and_links(tmp_path):
    _compile(str(tmp_path / "hello_world"))


@pytest.mark.gpu
def test_hello_world_matches_reference(tmp_path, gold):
    exe = str(tmp_path / "hello_world")
    _compile(exe, '-DHELLO_FORMAT="%.9g %.9g %.9g\\n"')
    print("case: tests/cpp/hello_world.cpp on the GPU", flush=True)
    res = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    got = np.array([[float(v) for v in line.split()] for line in res.stdout.strip().splitlines()])
    want = gold["hello_ccd_bodies"][:, [0, 1, 6]].astype(np.float64)
    assert got.shape == want.shape
    assert np.abs(got - want).max() <= 1e-5, np.abs(got - want).max(axis=0)


def test_body_api_compiles_and_links(tmp_path):
    _compile(str(tmp_path / "body_api"), src=BODY_API)


@pytest.mark.gpu
def test_body_api_facts(tmp_path):
    """tests/cpp/body_api.cpp on the GPU path: each body's contact list is the
    world's list restricted to it (order, links, other body), SetMassData's
    mass, centre, inertia and velocity shift, a fixed-rotation body keeping
    its angle, SetBullet, the local-point velocity."""
    exe = str(tmp_path / "body_api")
    _compile(exe, src=BODY_API)
    print("case: tests/cpp/body_api.cpp on the GPU", flush=True)
    res = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.startswith("ok"), res.stdout

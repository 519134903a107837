"""The drop-in boundary from C++: the reference's HelloWorld
(HelloWorld/HelloWorld.cpp:30-101), with only its include line changed,
compiles against include/box2d and links libbox2d.so; on the GPU its output
is the reference build's trajectory (committed fixture hello_ccd_bodies:
6/2 iterations, sleeping and continuous collision on, the reference's
defaults) within 1e-5.

The program is tests/cpp/hello_world.cpp, and INTEGRATION.md shows the same
text, which the first test checks.
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


def _compile(out: str, *defines: str) -> None:
    cmd = ["g++", "-std=c++14", "-Wall", "-Werror", f"-I{E.ROOT}/include", *defines, SRC, f"-L{LIBDIR}",
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

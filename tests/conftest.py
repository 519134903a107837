import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import engines  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP step through the C ABI")


@pytest.fixture(scope="session")
def orc():
    if not engines.have(engines.ORC_LIB):
        pytest.fail("oracle/_build/liborc.so is missing: run `python -m box2dpp_amd.build`")
    return engines.Engine("orc")


@pytest.fixture(scope="session")
def ref():
    if not engines.have(engines.REF_LIB):
        pytest.skip("oracle/_ref/libb2ref.so not built (needs the reference sources)")
    return engines.Engine("ref")


@pytest.fixture(scope="session")
def gpu():
    if not engines.have(engines.GPU_HOST_LIB):
        pytest.fail("box2dpp_amd/lib/libbox2d.so is missing: run `python -m box2dpp_amd.build`")
    return engines.Engine("gpu")


@pytest.fixture(scope="session")
def gold():
    """The committed reference-build fixtures (tests/golden/ref_vectors.npz)."""
    import numpy as np
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def orclib():
    return engines.oracle()


@pytest.fixture(scope="session")
def reflib():
    if not engines.have(engines.REF_LIB):
        pytest.skip("oracle/_ref/libb2ref.so not built")
    return engines.reference()

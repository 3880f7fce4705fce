import os
import sys

import numpy as np
import pytest

# torch brings its own HIP runtime; it must be loaded before libpumas_hip in
# a process that uses both (device buffers of the GPU tests).
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def ref_standard():
    import oracle as orc
    if not orc.available():
        pytest.skip("oracle library not built (make -C oracle)")
    p = orc.RefPhysics("standard_c005")
    yield p
    p.close()


@pytest.fixture(scope="session")
def ref_iron():
    import oracle as orc
    if not orc.available():
        pytest.skip("oracle library not built (make -C oracle)")
    p = orc.RefPhysics("iron_c001")
    yield p
    p.close()


@pytest.fixture(scope="session")
def gpu_standard():
    import pumas_amd as pa
    return pa.Physics(pa.physics_path("standard_c005"))


@pytest.fixture(scope="session")
def gpu_iron():
    import pumas_amd as pa
    return pa.Physics(pa.physics_path("iron_c001"))
import os, sys; sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))

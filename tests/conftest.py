import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "autodock-gpu_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def c1():
    """The C1 workload: 30-atom / 6-torsion seeded ligand on 60^3 maps at 0.375 A."""
    import helpers
    return helpers.synthetic_case(30, 6, lig_seed=7, map_seed=11, npts=60)


@pytest.fixture(scope="session")
def gpu_ctx():
    import adgpu
    ctx = adgpu.Context(0)
    yield ctx
    ctx.close()

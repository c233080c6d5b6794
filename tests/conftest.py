import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "pbrt-renderer-_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the product through the C ABI")
    config.addinivalue_line("markers", "slow: long-running parity case")


@pytest.fixture(scope="session")
def gpu():
    import nori_gpu
    ctx = nori_gpu.NoriGPU(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def cornell():
    import nori_gpu
    return nori_gpu.generate_scene(1)


@pytest.fixture(scope="session")
def disney_small():
    import nori_gpu
    return nori_gpu.generate_scene(2, subdiv=3, grid=4, env_width=256, env_height=128)

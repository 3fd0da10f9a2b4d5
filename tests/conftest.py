import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)

from ve_loader import load_package  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def ve():
    return load_package()


@pytest.fixture(scope="session")
def scenes(ve):
    from voxelengine_amd import scenes as s
    return s


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    return o


@pytest.fixture(scope="session")
def small_scene(ve, scenes):
    depth, ops = scenes.scene_small(1)
    words, info = ve.octree_build(ops, depth)
    return dict(depth=depth, ops=ops, words=words, info=info)


@pytest.fixture(scope="session")
def renderer(ve):
    import torch  # noqa: F401  (device enumeration only; the product path is the HIP library)
    r = ve.Renderer(512, 512, hash_width=16384, max_bounces=8)
    yield r
    r.close()

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libasuna.so")


@pytest.fixture(scope="session")
def scene_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("scenes")


@pytest.fixture(scope="session")
def box_scene(scene_dir):
    import asuna_amd as A
    return A.generate_scene(0, str(scene_dir / "box"))


@pytest.fixture(scope="session")
def spheres_small(scene_dir):
    import asuna_amd as A
    return A.generate_scene(1, str(scene_dir / "spheres_s1"), scale=1)


@pytest.fixture(scope="session")
def multishot_scene(scene_dir):
    import asuna_amd as A
    return A.generate_scene(2, str(scene_dir / "multishot"))

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "blender-lucid-renderer_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def lucid():
    import lucidrenderer

    return lucidrenderer


def make_renderer(lucid, scene, cam, algorithm="mis"):
    import workloads

    r = lucid.Renderer()
    assert r.load_scene_json(workloads.to_json(scene))
    r.set_camera(*cam["pos"], *cam["dir"], *cam["up"], cam["fov"])
    r.set_algorithm(algorithm)
    return r


def make_oracle(scene, cam):
    import oracle

    o = oracle.OracleScene(scene)
    o.set_camera(cam["pos"], cam["dir"], cam["up"], cam["fov"])
    return o

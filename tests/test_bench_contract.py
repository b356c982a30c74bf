"""bench.py's command-line contract, host side: more than one GPU is refused with a message, every
workload builds its scene, and the committed counter summaries give the traffic figure of the
1080p workloads that have one."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("name", ["mandelbulb8", "material", "mandelbox", "menger", "hybrid", "deepzoom"])
def test_workload_scenes_build(name):
    p = bench.make_scene(name, 1920, 1080)
    assert p.imageWidth > 0 and p.imageHeight > 0
    assert name in bench.WORKLOADS and name in bench.F_IT


@pytest.mark.parametrize("name", ["mandelbulb8", "material", "mandelbox", "menger", "hybrid", "image"])
def test_profiles_give_traffic(name):
    traffic, src = bench.measured_traffic(name, 1920, 1080)
    assert traffic and traffic > 40e6 and src.startswith("profiles/")
    assert bench.measured_traffic(name, 640, 480) == (None, None)

"""GPU render vs the FP64 oracle on identical scenes, through the C ABI (libmbx.so).

Bars (FP32 kernel vs FP64 restatement of the reference CPU renderer):
  hit/miss mask equal on >= 99.9% of pixels; iteration count equal on >= 99.9% of pixels
  (Mandelbulb) -- both are integer outputs compared exactly;
  depth within 1e-3 relative on >= 99.9% of pixels both hit; RGB within 1e-2 absolute on
  >= 99.9% of pixels.
"""
import sys

import numpy as np
import pytest

from mandelbulber2_amd import abi, scene
from tests.parity_util import compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from mandelbulber2_amd import runtime
    runtime.load()
    return runtime


def _report(name, st):
    print("\n[%s] " % name + " ".join("%s=%s" % (k, ("%.6g" % v) if isinstance(v, float) else v)
                                      for k, v in st.items()), flush=True)


def test_scene_info_matches_oracle(rt, oracle_mod):
    for p in (scene.mandelbulb8(320, 240), scene.mandelbox_shadow_ao(320, 180, seed=3)):
        ctx = rt.Context(0)
        ctx.upload(p)
        m, fov, light, kind = ctx.scene_info()
        ctx.close()
        m_o, fov_o, light_o = oracle_mod.frame_constants(p)
        np.testing.assert_allclose(m, m_o, rtol=0, atol=1e-14)
        assert abs(fov - fov_o) <= 1e-15
        np.testing.assert_allclose(light, light_o, rtol=0, atol=1e-14)


def test_distance_matches_oracle(rt, oracle_mod):
    rs = np.random.default_rng(7)
    pts = rs.uniform(-1.3, 1.3, size=(4096, 3))
    p = scene.mandelbulb8(640, 480)
    ctx = rt.Context(0)
    ctx.upload(p)
    d, it = ctx.distance(pts, 1e-3)
    ctx.close()
    d_o, it_o = oracle_mod.distance(p, pts, 1e-3)
    outside = d_o > 1e-3
    assert (it == it_o).mean() >= 0.995
    rel = np.abs(d[outside] - d_o[outside]) / d_o[outside]
    assert np.quantile(rel, 0.99) < 1e-4, np.quantile(rel, 0.99)


def test_mandelbulb8_small_parity(rt, oracle_mod):
    p = scene.mandelbulb8(160, 120)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("mandelbulb8 160x120", st)
    assert st["mask_agree"] >= 0.999
    assert st["iters_agree"] >= 0.995
    assert st["rgb_within_1e-2"] >= 0.995


def test_mandelbulb8_640x480_parity(rt, oracle_mod):
    p = scene.mandelbulb8(640, 480)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("mandelbulb8 640x480", st)
    assert st["mask_agree"] >= 0.999
    assert st["iters_agree"] >= 0.999
    assert st["depth_within_1e-3"] >= 0.999
    assert st["rgb_within_1e-2"] >= 0.999
    # the work counters agree to within the divergence the FP32 march allows
    assert abs(g["de_evals"] - r["de_evals"]) / r["de_evals"] < 0.01

"""Fish-eye, equirectangular and fish-eye-cut projections and antialiasing on the device against
the FP64 oracle (projection_3d.cpp:41-89, render_worker.cpp:96-202, 669, 688).

The projections' view vectors are FP64 on every kind (the device's sin and cos may differ from the
host's in the last bit, so the FP64 kinds agree statistically rather than bit for bit); the FP32
kinds march them in FP32 as they march the three-point rays.  Bars: mask and iterations on
>= 99.9% of pixels, depth within 1e-4 (FP32) / 1e-6 (FP64) relative on >= 99.9% of common hits,
RGB within 1e-2 on >= 99.9% of pixels; the fish-eye cut's blank pixels exactly as the oracle's.
"""
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


def _bulb(W, H):
    p = scene.reference_material(scene.plain_material(scene.mandelbulb8(W, H)))
    p.light.castShadows = 1
    p.light.softShadowCone = 0.0
    return p


def _box(W, H):
    p = scene.mandelbox_shadow_ao(W, H, seed=3)
    p.ambientOcclusionEnabled = 0
    return p


def _generic(W, H):
    p = scene.plain_material(scene.default_params(W, H))
    p.fractal.formula = abi.MBX_FORMULA_QUATERNION
    p.fractal.formulaMaxiter = 60
    p.hybridCount = 2
    p.hybridFractals[0].formula = abi.MBX_FORMULA_MANDELBULB_POWER2
    p.light.castShadows = 0
    abi.set3(p.camera, (1.4, -2.8, 1.0))
    return p


KINDS = {"bulb8": (_bulb, False), "mandelbox": (_box, True), "generic": (_generic, True)}
PERSP = {1: 160.0, 2: 300.0, 3: 200.0}


def _check(rt, oracle_mod, p, name, fp64):
    gpu = rt.render_to_host(p)
    ref = oracle_mod.render(p)
    st = compare(gpu, ref)
    _report(name, st)
    assert st["mask_agree"] >= 0.999 and st["iters_agree"] >= 0.999, st
    assert "depth_rel_p999" in st, st
    assert st["depth_rel_p999"] <= (1e-6 if fp64 else 1e-4), st
    assert st["rgb_within_1e-2"] >= 0.999, st
    return gpu, ref


@pytest.mark.parametrize("kind", sorted(KINDS))
@pytest.mark.parametrize("persp", [1, 2, 3])
def test_projection_matches_oracle(rt, oracle_mod, kind, persp):
    make, fp64 = KINDS[kind]
    p = make(200, 150)
    p.perspectiveType = persp
    p.fov = PERSP[persp]
    gpu, ref = _check(rt, oracle_mod, p, "persp%d-%s" % (persp, kind), fp64)
    if persp == 3:
        blank = (ref["depth"] == 1e20) & (ref["rgb"] == 0).all(axis=-1)
        assert blank.sum() > 1000
        assert np.array_equal(gpu["rgb"][blank], ref["rgb"][blank])
        assert np.all(gpu["opacity"][blank] == 65535)


@pytest.mark.parametrize("kind,n,mc", [("bulb8", 2, False), ("bulb8", 3, False), ("bulb8", 2, True),
                                       ("mandelbox", 2, False), ("generic", 2, True)])
def test_antialiasing_matches_oracle(rt, oracle_mod, kind, n, mc):
    make, fp64 = KINDS[kind]
    p = make(160, 120)
    p.antialiasingEnabled = 1
    p.antialiasingSize = n
    if mc:
        p.DOFMonteCarlo = 1
        p.DOFSamples = 2
        p.DOFMinSamples = 1000
    _check(rt, oracle_mod, p, "aa%d%s-%s" % (n, "-mc" if mc else "", kind), fp64)


def test_projection_and_antialiasing_in_one_batch(rt, oracle_mod):
    """Frames of every projection, with and without antialiasing, in one batch launch."""
    frames = []
    for persp in (0, 1, 2, 3):
        p = _bulb(96, 72)
        p.perspectiveType = persp
        p.fov = PERSP.get(persp, 53.13)
        p.antialiasingEnabled = persp % 2
        frames.append(p)
    import torch
    ctx = rt.Context(0)
    try:
        bufs = [ctx.alloc_outputs(p) for p in frames]
        ctx.render_batch(frames, bufs)
        st = ctx.stats()
        torch.cuda.synchronize()
        # antialiased frames march size^2 rays per pixel
        assert st["rays"] == sum(p.imageWidth * p.imageHeight * (4 if p.antialiasingEnabled else 1) for p in frames)
        for i, (p, b) in enumerate(zip(frames, bufs)):
            g = {k: v.cpu().numpy() for k, v in b.items() if hasattr(v, "cpu")}
            single = rt.render_to_host(p)
            for key in ("rgb", "depth", "iters", "mask"):
                np.testing.assert_array_equal(g[key], single[key], err_msg=key)
            s = compare(g, oracle_mod.render(p))
            _report("batch-%d" % i, s)
            assert s["mask_agree"] >= 0.999 and s["rgb_within_1e-2"] >= 0.999, s
    finally:
        ctx.close()

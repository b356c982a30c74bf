"""Example scenes loaded with mbx_load_fract, rendered through the C ABI on the GPU and compared
with the FP64 oracle's render of the same parameter block.

This mirrors the reference's example test (mandelbulber2/src/test.cpp:88-247), which loads every
example scene, overrides its image size and renders it: here the fixtures under
tests/golden/fract are loaded, shrunk to 192 pixels wide (aspect kept), and rendered.  Effects the
path does not render (mbx_fract_info.unsupported) are absent from both renders (iteration fog,
clouds and shadows in fog, which the render refuses, are switched off); screen-space AO is the
post effect of mbx_render_ssao over the rendered planes, checked on the post image; multi-ray AO
takes its directions and colours from the default lightmap texture (mbx_prepare_ao_vectors on the
decoded fixture); basic and distance fog are rendered (opacity within 2 of 65535).  Bars as
tests/test_gpu_parity.py: mask and iterations equal on
>= 99.9% of pixels (>= 99.5% iterations for the FP32 general-power Mandelbulb, whose asin / atan2 /
pow run in FP32), depth within 1e-3 relative on >= 99.9% of common hits (the FP64 kinds, Mandelbox
and the deep-zoom Mandelbulb: 99.9th percentile within 1e-6), RGB within 1e-2 on >= 99.9% of pixels.
"""
import os

import pytest

from mandelbulber2_amd import abi
from tests.parity_util import compare

pytestmark = pytest.mark.gpu

FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fract")

SCENES = [
    "mandelbulb001.fract",
    "mandelbulb002.fract",
    "mandelbulb_power_8.fract",
    "mandelbox001.fract",
    "mandelbox_22.fract",
    "keyframe_anim_mandelbox_boxes.fract",
    "IFS_001.fract",
    "menger_sponge_001.fract",
    # hybrids (the general formula kernel, FP64)
    "hybrid002.fract",
    "hybrid003.fract",
    "hybrid14.fract",
    "hybrid16.fract",
    "hybrid22_foldingIntPow.fract",
    "hybrid001.fract",
    # single formulas on the general kernel
    "aexion01.fract",
    "xenodreambuie2.fract",
    "hypercomplex_03.fract",
    "box_fold_bulb_pow2_001.fract",
    # Mandelbox main rotation: the specialised FP64 kernel and a hybrid
    "mandelbox_rot_001.fract",
    "FoldIntPow2_3.fract",
    # rotated Mandelbox folds (general kernel)
    "mandelbox62.fract",
    # global box and spherical foldings before every iteration (general kernel)
    "folded_mender_sponge.fract",
    # slow-shading normals (iteration counts of an 11^3 grid of orbits)
    "aexion05.fract",
    # iteration-threshold mode: the surface is where orbits reach maxiter (general kernel)
    "mandelbox8_maxiter.fract",
    "glow.fract",
    # multi-ray AO with the lightmap's directions and colours
    "IFS_25.fract",
    "mandelbox33_spiral.fract",
    # basic and volumetric distance fog
    "IFS_23.fract",
    "mandelbox55.fract",
    # a fish-eye projection and an antialiased IFS
    "fish_eye.fract",
    "IFS_19_maxiter.fract",
]


@pytest.fixture(scope="module")
def rt():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from mandelbulber2_amd import runtime
    runtime.load()
    return runtime


def lightmap():
    """The default lightmap texture as cTexture holds an 8-bit image: value / 256 per channel."""
    import numpy as np
    from PIL import Image
    img = Image.open(os.path.join(os.path.dirname(FIXTURES), "textures", "lightmap.jpg")).convert("RGB")
    return np.asarray(img, np.float32) / 256.0


def _load(rt, name, width=192):
    with open(os.path.join(FIXTURES, name), "rb") as f:
        p, info = rt.load_fract(f.read())
    p.imageHeight = max(1, round(width * p.imageHeight / p.imageWidth))
    p.imageWidth = width
    if p.ambientOcclusionEnabled and p.ambientOcclusionMode == abi.MBX_AO_MULTIPLE_RAYS:
        # multi-ray AO: the directions and colours of the default lightmap (the scenes name no other)
        rt.prepare_ao_vectors(p, lightmap())
    # the fog the render refuses (the reader flags it): the scene's image without it
    p.iterFogEnabled = 0
    p.cloudsEnabled = 0
    p.basicFogCastShadows = 0
    p.distanceFogShadows = 0
    if p.basicFogEnabled or p.volFogEnabled or p.glowEnabled:
        p.volumetricLightDEFactor = 1.0
    # the projections and antialiasing render without the lens variant's effects (the reader
    # flags the combination): a fish-eye scene without its lens and fog, an antialiased scene with
    # its lens and fog but without the antialiasing grid
    lens_fx = (p.DOFMonteCarlo and (p.DOFEnabled or p.DOFMonteCarloChromaticAberration)) or \
        p.basicFogEnabled or p.volFogEnabled
    if lens_fx and p.perspectiveType != 0:
        p.DOFEnabled = p.DOFMonteCarloChromaticAberration = p.basicFogEnabled = p.volFogEnabled = 0
    elif lens_fx:
        p.antialiasingEnabled = 0
    return p, info


@pytest.mark.parametrize("name", SCENES)
def test_loaded_example_scene_parity(rt, oracle_mod, name):
    p, info = _load(rt, name)
    flags = abi.fract_flag_names(info.unsupported)
    # (perspective: stereo, or a projection with the lens effects, which _load turns off)
    assert not {"hybrid", "primitives", "formula", "interior", "disabled"} & set(flags)
    assert "perspective" not in flags or p.perspectiveType != 0
    print("\n[%s] %dx%d formula %d, not rendered: %s" % (name, p.imageWidth, p.imageHeight,
                                                        p.fractal.formula, ",".join(flags)), flush=True)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    print("  kernel kind %d; " % g["de_kind"] + " ".join("%s=%s" % (k, ("%.6g" % v) if isinstance(v, float) else v)
                          for k, v in st.items()), flush=True)
    trig = p.fractal.formula == abi.MBX_FORMULA_MANDELBULB and not (
        p.fractal.bulb.power == 8.0 and p.fractal.bulb.alphaAngleOffset == 0.0
        and p.fractal.bulb.betaAngleOffset == 0.0)
    if g["de_kind"] == 6:   # general kernel: FP64 libm-class trigonometry on the device
        trig = any(f.formula in (abi.MBX_FORMULA_MANDELBULB, abi.MBX_FORMULA_XENODREAMBUIE)
                   for f in [p.fractal] + list(p.hybridFractals)[:max(0, p.hybridCount - 1)])
    fp64 = p.fractal.formula == abi.MBX_FORMULA_MANDELBOX or g["de_kind"] in (5, 6)
    assert st["mask_agree"] >= 0.999, st
    assert st["iters_agree"] >= (0.995 if trig and (not fp64 or g["de_kind"] == 6) else 0.999), st
    assert "depth_rel_p999" in st, ("no pixel hit in both renders", st)
    if fp64 and trig:
        assert st["depth_within_1e-3"] >= 0.999, st
    elif fp64:
        assert st["depth_rel_p999"] <= 1e-6, st
    else:
        assert st["depth_within_1e-3"] >= 0.999, st
    # multi-ray AO on an FP32 kind: each direction's walk stops at the first sample below the
    # threshold and subtracts (end - r) / end, so where FP32 sample positions move the crossing to
    # the neighbouring sample, the pixel's AO steps by a fraction of one direction's share
    multiray_fp32 = (p.ambientOcclusionEnabled and p.ambientOcclusionMode == abi.MBX_AO_MULTIPLE_RAYS
                     and not fp64)
    assert st["rgb_within_1e-2"] >= (0.998 if multiray_fp32 else 0.999), st
    if p.basicFogEnabled or p.volFogEnabled:
        assert st["opacity_within_2"] >= 0.999, st
    if "post_within_1e-2" in st:
        assert st["post_within_1e-2"] >= (0.998 if multiray_fp32 else 0.999), st


def test_mandelbulb_precision_choice(rt):
    """The deep-zoom Mandelbulb scene marches in FP64 (kind 5), the others in FP32."""
    from mandelbulber2_amd import scene
    want = {"mandelbulb_power_8.fract": 5, "mandelbulb001.fract": 1, "mandelbulb002.fract": 1}
    for name, kind in want.items():
        p, _ = _load(rt, name)
        ctx = rt.Context(0)
        ctx.upload(p)
        assert ctx.scene_info()[3] == kind, name
        ctx.tune(rt.MBX_TUNE_PRECISION, 1)
        ctx.upload(p)
        assert ctx.scene_info()[3] in (0, 1), name
        ctx.close()
    ctx = rt.Context(0)
    ctx.upload(scene.mandelbulb8(1920, 1080))
    assert ctx.scene_info()[3] == 0          # the 1080p benchmark scene stays on the FP32 kernel
    ctx.close()


def test_ifs_precision_choice(rt):
    """IFS 23 looks at the sponge from 2e-5 away (thresholds near 1e-8 at coordinates near 0.6):
    it takes the general FP64 kernel (kind 6); the other IFS scenes stay on the FP32 IFS kernel
    (kind 4) unless FP64 is forced."""
    want = {"IFS_23.fract": 6, "IFS_001.fract": 4, "menger_sponge_001.fract": 4, "IFS_25.fract": 4}
    for name, kind in want.items():
        p, _ = _load(rt, name)
        ctx = rt.Context(0)
        try:
            ctx.upload(p)
            assert ctx.scene_info()[3] == kind, name
            ctx.tune(rt.MBX_TUNE_PRECISION, 2)
            ctx.upload(p)
            assert ctx.scene_info()[3] == 6, name
            ctx.tune(rt.MBX_TUNE_PRECISION, 1)
            ctx.upload(p)
            assert ctx.scene_info()[3] == 4, name
        finally:
            ctx.close()


@pytest.mark.parametrize("power", [8.0, 3.0])
def test_forced_fp64_mandelbulb_parity(rt, oracle_mod, power):
    """MBX_TUNE_PRECISION = 2 on ordinary views: the FP64 march follows the oracle's operation
    order, so hits and iterations agree and depths to FP64 rounding of the trigonometry."""
    from mandelbulber2_amd import scene
    p = scene.mandelbulb8(160, 120)
    p.fractal.bulb.power = power
    g = rt.render_to_host(p, tune={rt.MBX_TUNE_PRECISION: 2})
    assert g["de_kind"] == 5
    r = oracle_mod.render(p)
    st = compare(g, r)
    print("\n[fp64 mandelbulb power %g] " % power + " ".join(
        "%s=%s" % (k, ("%.6g" % v) if isinstance(v, float) else v) for k, v in st.items()), flush=True)
    assert st["mask_agree"] >= 0.999 and st["iters_agree"] >= 0.999, st
    assert st["depth_rel_p999"] <= 1e-6, st
    assert st["rgb_within_1e-3"] >= 0.999, st
    assert g["de_evals"] == pytest.approx(r["de_evals"], rel=1e-3)

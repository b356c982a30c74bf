"""Palette colouring, specular highlights and glow on the GPU vs the FP64 oracle.

Scenes carry the reference's material defaults (reference_material: palette colouring from the
default surface gradient, plastic and metallic highlights with roughness 0.1, glow on).  Bars as in
test_gpu_parity: mask and iteration counts equal on >= 99.9% of pixels, depth within 1e-3
(FP32 march) or 1e-6 (FP64 Mandelbox march), RGB within 1e-2 on >= 99.9% of pixels.
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


def _check(st, depth_tol=1e-3, iters=0.999):
    assert st["mask_agree"] >= 0.999, st
    assert st["iters_agree"] >= iters, st
    if depth_tol == 1e-3:
        assert st["depth_within_1e-3"] >= 0.999, st
    else:
        assert st["depth_rel_p999"] <= depth_tol, st
    assert st["rgb_within_1e-2"] >= 0.999, st


def _scene(name):
    if name == "mandelbulb8":
        return scene.reference_material(scene.mandelbulb8(128, 96)), 1e-3, 0.999
    if name == "mandelbulb_trig":
        p = scene.reference_material(scene.mandelbulb8(128, 96))
        p.fractal.bulb.power = 5.5
        return p, 1e-3, 0.995
    if name == "menger_ifs":
        return scene.reference_material(scene.menger_ifs(160, 120)), 1e-3, 0.999
    if name == "mandelbox":
        return scene.reference_material(scene.mandelbox_shadow_ao(192, 108, seed=1)), 1e-6, 0.999
    if name == "rotated_mandelbox":
        p = scene.reference_material(scene.mandelbox_shadow_ao(128, 72, seed=1))
        abi.set3(p.formulaRotation, (17.0, -33.0, 71.0))
        return p, 1e-6, 0.999
    raise KeyError(name)


@pytest.mark.parametrize("name", ["mandelbulb8", "mandelbulb_trig", "menger_ifs", "mandelbox",
                                  "rotated_mandelbox"])
def test_reference_material_parity(rt, oracle_mod, name):
    p, depth_tol, iters = _scene(name)
    print("[case] %s %dx%d" % (name, p.imageWidth, p.imageHeight), flush=True)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("%s reference material" % name, st)
    _check(st, depth_tol, iters)
    # the shading is not the plain material's
    q = scene.plain_material(scene.copy_params(p))
    plain = oracle_mod.render(q)
    assert np.abs(plain["rgb"] - r["rgb"]).max() > 0.05


@pytest.mark.parametrize("algorithm", [1, 2, 3, 4])
@pytest.mark.parametrize("name", ["mandelbulb8", "mandelbox"])
def test_colouring_algorithms_parity(rt, oracle_mod, name, algorithm):
    """Colouring measures z.p, sphere, cross and line (standard is covered above), before and
    after the v2.15 bailout update."""
    p, depth_tol, iters = _scene(name)
    p.imageWidth, p.imageHeight = 96, 72 if name == "mandelbulb8" else 54
    p.coloringAlgorithm = algorithm
    p.coloringPreV215 = 1 if algorithm == 2 else 0
    p.coloringSphereRadius = 0.8
    abi.set3(p.coloringLineDirection, (0.3, -0.5, 0.81))
    p.coloringSpeed = 2.3
    p.coloringPaletteOffset = 0.4
    print("[case] %s algorithm %d" % (name, algorithm), flush=True)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("%s colouring %d" % (name, algorithm), st)
    _check(st, depth_tol, iters)


def test_monte_carlo_gi_reference_material_parity(rt, oracle_mod):
    """GI bounces tinted by the palette colour of each bounce surface, with specular highlights
    and emissive luminosity seen by the bounces."""
    p = scene.reference_material(scene.mandelbulb8(96, 72))
    p.DOFMonteCarlo = 1
    p.DOFSamples = 4
    p.DOFMonteCarloGlobalIllumination = 1
    p.reflectionsMax = 2
    p.light.castShadows = 1
    p.luminosity = 0.3
    p.luminosityEmissive = 2.0
    print("[case] MC GI reference material 96x72", flush=True)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("mandelbulb8 MC GI reference material", st)
    _check(st)


def test_batch_with_reference_material_equals_single_frames(rt):
    """Frames with palette colouring, highlights and glow in one grid launch equal their
    single-frame renders bit for bit."""
    import torch
    frames = []
    for k, power in enumerate((6.0, 7.0, 9.5)):
        p = scene.reference_material(scene.mandelbulb8(64, 48))
        p.fractal.bulb.power = power
        p.coloringPaletteOffset = 0.1 * k
        frames.append(p)
    ctx = rt.Context(0)
    bufs = [ctx.alloc_outputs(p) for p in frames]
    ctx.render_batch(frames, bufs)
    torch.cuda.synchronize()
    for p, b in zip(frames, bufs):
        single = rt.render_to_host(p)
        for key in ("rgb", "depth", "iters", "mask", "steps"):
            np.testing.assert_array_equal(b[key].cpu().numpy(), single[key], err_msg=key)
    ctx.close()


def test_shading_settings_are_validated(rt):
    import torch
    p = scene.reference_material(scene.mandelbulb8(32, 24))
    bad = scene.copy_params(p)
    bad.surfaceGradientCount = 1
    with pytest.raises(rt.MbxError, match="surfaceGradientCount"):
        rt.render_to_host(bad)
    bad = scene.copy_params(p)
    bad.surfaceGradientPosition[3], bad.surfaceGradientPosition[4] = 0.9, 0.1
    with pytest.raises(rt.MbxError, match="sorted"):
        rt.render_to_host(bad)
    bad = scene.copy_params(p)
    bad.DEFactor = 1e-10
    with pytest.raises(rt.MbxError, match="glow"):
        rt.render_to_host(bad)
    bad = scene.copy_params(p)
    bad.specularWidth = 0.0
    with pytest.raises(rt.MbxError, match="specular"):
        rt.render_to_host(bad)
    torch.cuda.synchronize()


def _split_scenes():
    from mandelbulber2_amd import scene as sc
    bulb = sc.reference_material(sc.mandelbulb8(160, 120))
    bulb.light.castShadows = 1
    box = sc.mandelbox_shadow_ao(192, 108, seed=3)
    ifs = sc.reference_material(sc.menger_ifs(128, 96))
    glow = sc.reference_material(sc.mandelbulb8(96, 72))
    glow.glowEnabled = 1
    fast = sc.mandelbox_shadow_ao(160, 90, seed=5)
    fast.ambientOcclusionMode = abi.MBX_AO_FAST
    return {"bulb8-material-shadow": bulb, "mandelbox-shadow-ao": box, "menger-material": ifs,
            "bulb8-glow": glow, "mandelbox-fast-ao": fast}


@pytest.mark.parametrize("name", list(_split_scenes().keys()))
def test_split_render_matches_the_one_kernel_render(rt, name):
    """MBX_TUNE_SPLIT = 2: the march launch plus the shading launch over the hit lists give the
    geometry planes of the one-kernel render bit for bit, and the same work counts; the shading
    planes too on the FP64 kinds (their shading arithmetic has contraction off), and to the last
    bit of FP32 rounding on the FP32 kinds (the compiler fuses their multiply-adds per kernel:
    the shadow factor of 22 of 19,200 pixels differed by 1 ulp when measured)."""
    p = _split_scenes()[name]
    a = rt.render_to_host(p, tune={rt.MBX_TUNE_SPLIT: 2}, optional=True)
    b = rt.render_to_host(p, tune={rt.MBX_TUNE_SPLIT: 0}, optional=True)
    fp64 = a["de_kind"] in (2, 3, 5, 6)
    for k in ("depth", "iters", "mask", "normal", "steps", "colour", "opacity", "world", "gi"):
        assert np.array_equal(a[k], b[k]), (name, k)
    for k in ("rgb", "specular", "shadow"):
        if fp64:
            assert np.array_equal(a[k], b[k]), (name, k)
        else:
            np.testing.assert_allclose(a[k], b[k], rtol=1e-6, atol=1e-6, err_msg=name + " " + k)
    assert (a["de_evals"], a["iterations"]) == (b["de_evals"], b["iterations"])


def test_split_render_batch_of_mixed_frames(rt):
    """A batch of frames of different sizes and kinds of shading in one split render equals the
    frames rendered one at a time without it."""
    import torch
    from mandelbulber2_amd import scene as sc
    frames = [sc.mandelbox_shadow_ao(96, 54, seed=s) for s in (1, 2)]
    frames.append(sc.mandelbox_shadow_ao(64, 64, seed=4))
    frames[1].light.castShadows = 0
    ctx = rt.Context(0)
    try:
        ctx.tune(rt.MBX_TUNE_SPLIT, 2)
        bufs = [ctx.alloc_outputs(p, optional=True) for p in frames]
        ctx.render_batch(frames, bufs)
        torch.cuda.synchronize(0)
        got = [{k: v.cpu().numpy() for k, v in b.items() if not k.startswith("_")} for b in bufs]
    finally:
        ctx.close()
    for p, g in zip(frames, got):
        r = rt.render_to_host(p, tune={rt.MBX_TUNE_SPLIT: 0}, optional=True)
        for k in ("rgb", "depth", "iters", "mask", "normal", "colour", "specular", "shadow"):
            assert np.array_equal(g[k], r[k]), k

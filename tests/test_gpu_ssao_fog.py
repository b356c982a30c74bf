"""Screen-space AO (mbx_render_ssao) and fog on the device against the FP64 oracle.

* mbx_render_ssao on the same input planes as the oracle's restatement of cSSAOWorker::doWork
  (synthetic relief, every perspective, random mode, qualities 1 to 10): the same FP64
  arithmetic, so the images agree to 1e-5 (the device's atan, sin and cos may differ from the
  host's in the last bit).
* The render's colour and opacity planes (the reference's colour and opacity buffers: colour
  within 1 of 255 on >= 99.9% of pixels, equal on >= 99%), and render + SSAO end to end against
  the oracle's render + SSAO (post image within 1e-2 on >= 99.9%).
* Basic fog and the volumetric distance fog, alone, with glow, on the FP32 Mandelbulb, the FP64
  Mandelbox and the general kernel, and with Monte Carlo samples: the device accumulates the
  blends front to back in float while the oracle walks the step buffer back to front, so the bars
  are RGB within 1e-2 on >= 99.9% of pixels and the opacity within 2 of 65535 on >= 99.9%.
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


def _ssao_params(W, H, quality, persp, random_mode):
    p = scene.default_params(W, H)
    p.ambientOcclusionEnabled = 1
    p.ambientOcclusionMode = abi.MBX_AO_SCREEN_SPACE
    p.ambientOcclusionQuality = quality
    p.ambientOcclusion = 1.3
    abi.set3(p.ambientOcclusionColor, (1.0, 0.6, 0.3))
    p.perspectiveType = persp
    p.fov = 70.0 if persp else 53.0
    p.SSAORandomMode = random_mode
    p.randomSeed = 11
    return p


def _device_ssao(rt, p, depth, colour, opacity, rgb):
    import torch
    ctx = rt.Context(0)
    try:
        dev = torch.device("cuda", 0)
        bufs = {"depth": torch.from_numpy(depth).to(dev), "colour": torch.from_numpy(colour).to(dev),
                "rgb": torch.from_numpy(rgb).to(dev),
                "opacity": torch.from_numpy(opacity.view(np.int16)).to(dev).view(torch.uint16)
                if opacity is not None else None}
        out = ctx.render_ssao(bufs, params=p)
        torch.cuda.synchronize(0)
        return out.cpu().numpy()
    finally:
        ctx.close()


@pytest.mark.parametrize("quality,persp,random_mode", [(4, 0, 0), (1, 0, 0), (10, 0, 0), (4, 0, 1), (3, 1, 0),
                                                       (4, 2, 0), (4, 3, 1)])
def test_ssao_kernel_matches_oracle_on_relief(rt, oracle_mod, quality, persp, random_mode):
    W, H = 301, 203
    p = _ssao_params(W, H, quality, persp, random_mode)
    rng = np.random.default_rng(quality * 10 + persp)
    yy, xx = np.mgrid[0:H, 0:W]
    depth = (3.0 + 0.4 * np.sin(xx / 9.0) * np.cos(yy / 7.0) + 0.3 * (((xx // 40) + (yy // 30)) % 2)
             + 0.01 * rng.random((H, W))).astype(np.float32)
    depth[rng.random((H, W)) < 0.05] = 1e20
    colour = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    opacity = rng.integers(0, 65536, (H, W), dtype=np.uint16)
    rgb = rng.random((H, W, 3), dtype=np.float32)
    print("case quality=%d perspective=%d random=%d" % (quality, persp, random_mode), flush=True)
    got = _device_ssao(rt, p, depth, colour, opacity, rgb)
    want = oracle_mod.ssao(p, depth, colour, opacity, rgb)
    err = np.abs(got.astype(np.float64) - want.astype(np.float64))
    print("max |diff| %.3g" % err.max(), flush=True)
    assert err.max() <= 1e-5
    assert np.abs(want - rgb).max() > 1e-3  # the effect is there


def test_ssao_in_place_and_without_opacity(rt, oracle_mod):
    import torch
    W, H = 128, 96
    p = _ssao_params(W, H, 4, 0, 0)
    yy, xx = np.mgrid[0:H, 0:W]
    depth = (2.0 + 0.5 * np.sin(xx / 5.0) + 0.5 * (yy > 40)).astype(np.float32)
    colour = np.full((H, W, 3), 200, np.uint8)
    rgb = np.random.default_rng(2).random((H, W, 3), dtype=np.float32)
    want = oracle_mod.ssao(p, depth, colour, None, rgb)
    ctx = rt.Context(0)
    try:
        dev = torch.device("cuda", 0)
        bufs = {"depth": torch.from_numpy(depth).to(dev), "colour": torch.from_numpy(colour).to(dev),
                "rgb": torch.from_numpy(rgb).to(dev)}
        ctx.render_ssao(bufs, params=p, rgb_out=bufs["rgb"])
        torch.cuda.synchronize(0)
        got = bufs["rgb"].cpu().numpy()
    finally:
        ctx.close()
    assert np.abs(got - want).max() <= 1e-5


def _material_bulb(W, H):
    p = scene.reference_material(scene.mandelbulb8(W, H))
    p.glowEnabled = 0
    return p


def test_render_colour_plane_and_ssao_end_to_end(rt, oracle_mod):
    p = _material_bulb(320, 240)
    p.ambientOcclusionEnabled = 1
    p.ambientOcclusionMode = abi.MBX_AO_SCREEN_SPACE
    p.ambientOcclusionQuality = 4
    gpu = rt.render_to_host(p)
    ref = oracle_mod.render(p)
    st = compare(gpu, ref)
    _report("ssao-e2e", st)
    assert st["mask_agree"] >= 0.999 and st["rgb_within_1e-2"] >= 0.999, st
    # the colour buffer truncates c * 255: an FP32 surface colour one rounding away from the
    # oracle's FP64 one changes a byte by 1 where c * 255 is near an integer, and a palette
    # position on the far side of a gradient stop moves it further, as it moves the RGB
    assert st["colour_equal"] >= 0.99 and st["colour_within_1"] >= 0.999, st
    assert st["opacity_max_diff"] == 0, st
    assert st["post_within_1e-2"] >= 0.999, st
    # the hits carry their surface colour, the misses none
    assert not gpu["colour"][gpu["mask"] == 0].any()
    assert gpu["colour"][gpu["mask"] > 0].any()


def _fog(p, basic=True, vol=False, glow=False):
    p.basicFogEnabled = int(basic)
    p.basicFogVisibility = 6.0
    p.volFogEnabled = int(vol)
    p.volFogDensity = 2.0
    p.volFogDistanceFactor = 0.3
    p.volFogColour1Distance = 0.2
    p.volFogColour2Distance = 0.6
    p.glowEnabled = int(glow)
    return p


@pytest.mark.parametrize("name,basic,vol,glow", [("basic", True, False, False), ("distance", False, True, False),
                                                 ("both+glow", True, True, True)])
def test_fog_bulb8_matches_oracle(rt, oracle_mod, name, basic, vol, glow):
    p = _fog(_material_bulb(320, 240), basic, vol, glow)
    gpu = rt.render_to_host(p)
    ref = oracle_mod.render(p)
    st = compare(gpu, ref)
    _report("fog-bulb8-" + name, st)
    assert st["mask_agree"] >= 0.999 and st["iters_agree"] >= 0.999, st
    assert st["depth_rel_p999"] <= 1e-4, st
    assert st["rgb_within_1e-2"] >= 0.999, st
    assert st["opacity_within_2"] >= 0.999, st
    assert ref["opacity"].max() > 1000  # the fog is there


def test_fog_mandelbox_fp64_matches_oracle(rt, oracle_mod):
    p = scene.mandelbox_shadow_ao(192, 108, seed=3)
    p.ambientOcclusionEnabled = 0
    p = _fog(p, basic=True, vol=True)
    p.basicFogVisibility = 20.0
    gpu = rt.render_to_host(p)
    ref = oracle_mod.render(p)
    st = compare(gpu, ref)
    _report("fog-mandelbox", st)
    assert st["mask_agree"] >= 0.999 and st["iters_agree"] >= 0.999, st
    assert st["depth_rel_p999"] <= 1e-6, st
    assert st["rgb_within_1e-2"] >= 0.999, st
    assert st["opacity_within_2"] >= 0.999, st


def test_fog_general_kernel_matches_oracle(rt, oracle_mod):
    """A hybrid of the quaternion and the power-2 Mandelbulb on the general FP64 kernel."""
    p = scene.plain_material(scene.default_params(160, 120))
    p.fractal.formula = abi.MBX_FORMULA_QUATERNION
    p.fractal.formulaMaxiter = 60
    p.hybridCount = 2
    p.hybridFractals[0].formula = abi.MBX_FORMULA_MANDELBULB_POWER2
    p.light.castShadows = 0
    abi.set3(p.camera, (1.4, -2.8, 1.0))
    p = _fog(p, basic=True, vol=True, glow=True)
    gpu = rt.render_to_host(p)
    ref = oracle_mod.render(p)
    st = compare(gpu, ref)
    _report("fog-generic", st)
    assert st["mask_agree"] >= 0.999 and st["iters_agree"] >= 0.999, st
    assert st["depth_rel_p999"] <= 1e-6, st
    assert st["rgb_within_1e-2"] >= 0.999, st
    assert st["opacity_within_2"] >= 0.999, st


def test_fog_monte_carlo_and_ssao(rt, oracle_mod):
    """Monte Carlo samples average the fog colour and opacity (integer mean of the opacity
    buffer); SSAO then scales its term by 1 - opacity."""
    p = _fog(_material_bulb(160, 120), basic=True, vol=True)
    p.DOFMonteCarlo = 1
    p.DOFSamples = 4
    p.DOFMinSamples = 10
    p.ambientOcclusionEnabled = 1
    p.ambientOcclusionMode = abi.MBX_AO_SCREEN_SPACE
    gpu = rt.render_to_host(p)
    ref = oracle_mod.render(p)
    st = compare(gpu, ref)
    _report("fog-mc-ssao", st)
    assert st["mask_agree"] >= 0.999, st
    assert st["rgb_within_1e-2"] >= 0.999, st
    assert st["opacity_within_2"] >= 0.999, st
    assert st["post_within_1e-2"] >= 0.999, st


def test_unsupported_fog_kinds_are_refused(rt):
    for field, value in (("iterFogEnabled", 1), ("cloudsEnabled", 1), ("volumetricLightDEFactor", 0.5)):
        p = _fog(_material_bulb(64, 48))
        setattr(p, field, value)
        ctx = rt.Context(0)
        try:
            with pytest.raises(rt.MbxError):
                ctx.upload(p)
        finally:
            ctx.close()
    p = _fog(_material_bulb(64, 48))
    p.basicFogCastShadows = 1
    ctx = rt.Context(0)
    try:
        with pytest.raises(rt.MbxError):
            ctx.upload(p)
    finally:
        ctx.close()


def _channel_stats(g, r, name, mask):
    e = np.abs(g[name].astype(np.float64) - r[name].astype(np.float64)).max(axis=-1)
    return float((e[mask] <= 1e-2).mean()) if mask.any() else 1.0


def test_optional_channels_match_oracle(rt, oracle_mod):
    """The world position, specular, shadow and GI channels (mbx_outputs.world / .specular /
    .shadow / .gi) against the oracle's: world within 1e-4 of the point's size on >= 99.9% of the
    pixels both renders hit, the others within 1e-2 on >= 99.9% of pixels."""
    p = _material_bulb(200, 150)
    p.light.castShadows = 1
    abi.set3(p.light.rotation, (30.0, 20.0, 0.0))
    g = rt.render_to_host(p, optional=True)
    r = oracle_mod.render(p, optional=True)
    both = (g["mask"] > 0) & (r["mask"] > 0)
    dw = np.abs(g["world"] - r["world"]).max(axis=-1) / np.maximum(np.abs(r["world"]).max(axis=-1), 1e-30)
    st = {"world_1e-4": float((dw[both] <= 1e-4).mean())}
    for name in ("specular", "shadow", "gi"):
        st[name] = _channel_stats(g, r, name, np.ones_like(both))
    _report("channels", st)
    assert st["world_1e-4"] >= 0.999 and all(st[k] >= 0.999 for k in ("specular", "shadow", "gi")), st
    assert np.abs(r["specular"]).max() > 1e-3 and (r["shadow"][both] < 1.0).any()
    # Monte Carlo GI over 4 samples: the channel is the mean of the samples' GI
    q = scene.plain_material(scene.mandelbulb8(96, 72))
    q.DOFMonteCarlo = 1
    q.DOFSamples = 4
    q.DOFMinSamples = 1000
    q.DOFMonteCarloGlobalIllumination = 1
    q.reflectionsMax = 1
    g = rt.render_to_host(q, optional=True)
    r = oracle_mod.render(q, optional=True)
    gi = _channel_stats(g, r, "gi", np.ones(g["mask"].shape, bool))
    print("[channels] gi (MC 4 samples) within 1e-2: %.6f" % gi, flush=True)
    assert gi >= 0.999 and np.abs(r["gi"]).max() > 1e-3

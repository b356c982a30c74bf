"""Pins of the oracle's optional image channels (render_worker.cpp:355-366, 405-420, 474-476) by
facts outside their code:

* world position: the march's end point is camera + t * direction for the ray of the pixel, with
  t the returned depth (the ray geometry of the frame constants), and a hit lies within one
  distance threshold of the surface;
* shadow: 1 on a hit without cast shadows, 0 with the light off, 1 on a miss; with cast shadows
  the Lambert colour closed form surface * light * intensity * shade * shadow holds with the
  channel's value;
* specular: the image with highlights minus the image without them (no shadow or AO rays, so
  the highlight's random draw moves nothing else) is the channel;
* GI: with one Monte Carlo sample the image with GI minus the image without it is the channel
  (the GI rays draw last), and with several samples the channel is the mean.
"""
import numpy as np
import pytest

from mandelbulber2_amd import abi, scene
from oracle import oracle


def _bulb(W=96, H=72):
    p = scene.mandelbulb8(W, H)
    abi.set3(p.fillLightColor, (0.0, 0.0, 0.0))
    return p


def test_world_position_is_the_ray_end_point():
    p = _bulb()
    r = oracle.render(p, optional=True)
    cam = np.array(p.camera)
    checked = 0
    for (x, y) in [(10, 10), (48, 36), (60, 30), (30, 50), (5, 70), (90, 2)]:
        m = oracle.march_pixel(p, x, y)
        d = np.asarray(m["dir"]) / np.linalg.norm(m["dir"])
        want = cam + d * m["depth"] * np.linalg.norm(m["dir"])
        np.testing.assert_allclose(r["world"][y, x], want, rtol=1e-6, atol=1e-6)
        if m["found"]:
            dist, _ = oracle.distance(p, r["world"][y, x].astype(np.float64)[None, :], m["thr"])
            assert dist[0] < m["thr"] * 1.01
            checked += 1
    assert checked >= 2


def test_shadow_channel_values_and_the_lambert_closed_form():
    p = _bulb()
    r = oracle.render(p, optional=True)
    hit = r["mask"] > 0
    assert np.all(r["shadow"][hit] == 1.0) and np.all(r["shadow"][~hit] == 1.0)
    p.light.enabled = 0
    r0 = oracle.render(p, optional=True)
    assert np.all(r0["shadow"][r0["mask"] > 0] == 0.0)
    # cast shadows: a hit's colour is surface * light * intensity * (1 - s + s max(n.l, 0)) * shadow
    p = _bulb(120, 90)
    p.light.castShadows = 1
    p.light.softShadowCone = 0.0
    p.light.intensity = 1.3
    abi.set3(p.surfaceColor, (0.6, 0.3, 0.2))
    abi.set3(p.light.rotation, (30.0, 20.0, 0.0))
    r = oracle.render(p, optional=True)
    _, _, light = oracle.frame_constants(p)
    hit = r["mask"].ravel() > 0
    sh = r["shadow"].reshape(-1, 3)[hit]
    assert np.all(sh[:, 0] == sh[:, 1]) and np.all(sh[:, 1] == sh[:, 2])
    assert (sh[:, 0] < 1.0).any() and (sh[:, 0] > 0.0).any()
    n = r["normal"].reshape(-1, 3)[hit].astype(np.float64)
    shade = (np.maximum(n @ np.asarray(light), 0.0)) * 1.3
    expect = (shade * sh[:, 0])[:, None] * np.array([0.6, 0.3, 0.2])[None, :]
    # (the normal plane is float: n.l near grazing carries its rounding)
    np.testing.assert_allclose(r["rgb"].reshape(-1, 3)[hit], expect, rtol=1e-4, atol=2e-6)
    # the channel is the factor the hard shadow applies: the same scene without cast shadows
    # (no random draws either way) differs by it alone
    p.light.castShadows = 0
    r1 = oracle.render(p, optional=True)
    np.testing.assert_allclose(r["rgb"].reshape(-1, 3)[hit], r1["rgb"].reshape(-1, 3)[hit] * sh, rtol=1e-6,
                               atol=1e-7)


def test_specular_channel_is_the_highlight_difference():
    p = scene.reference_material(_bulb())
    p.glowEnabled = 0
    p.useColorsFromPalette = 0
    on = oracle.render(p, optional=True)
    q = scene.reference_material(_bulb())
    q.glowEnabled = 0
    q.useColorsFromPalette = 0
    q.specularPlasticEnable = 0
    q.metallic = 0
    off = oracle.render(q, optional=True)
    hit = on["mask"] > 0
    assert np.abs(on["specular"][hit]).max() > 1e-3
    np.testing.assert_allclose(on["rgb"][hit] - off["rgb"][hit], on["specular"][hit], rtol=1e-5, atol=2e-6)
    assert not off["specular"].any() and not on["specular"][~hit].any()


def _gi_scene(samples):
    p = scene.plain_material(_bulb(48, 36))
    p.DOFMonteCarlo = 1
    p.DOFSamples = samples
    p.DOFMinSamples = 1000
    p.reflectionsMax = 1
    return p


def test_gi_channel_is_the_gi_difference_and_mean():
    p = _gi_scene(1)
    p.DOFMonteCarloGlobalIllumination = 1
    on = oracle.render(p, optional=True)
    p.DOFMonteCarloGlobalIllumination = 0
    off = oracle.render(p, optional=True)
    assert np.abs(on["gi"]).max() > 1e-3 and not off["gi"].any()
    np.testing.assert_allclose(on["rgb"] - off["rgb"], on["gi"], rtol=1e-5, atol=2e-6)
    # several samples: the channel is the mean of the per-sample GI terms, each bounded by the
    # radiance limit times the surface colour
    p = _gi_scene(4)
    p.DOFMonteCarloGlobalIllumination = 1
    r = oracle.render(p, optional=True)
    assert np.all(r["gi"] >= 0.0) and np.all(r["gi"] <= p.monteCarloGIRadianceLimit + 1e-6)
    # a pixel all of whose samples miss has none (the mask is the first sample's: jittered later
    # samples may still hit, so the test takes the pixels far from any hit)
    far = r["mask"] == 0
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            far &= np.roll(np.roll(r["mask"] == 0, dy, 0), dx, 1)
    assert far.any() and not r["gi"][far].any()

"""Pins of the oracle's fog walk (VolumetricShader, shader_volumetric.cpp:70-700) by closed forms.

The scene is the unit ball (the power-2 Mandelbulb in Julia mode with c = 0) and advanced quality
with abs_min = abs_max = s, so every march step is s.  A ray that misses runs until the scan
passes view_distance_max = M: its step buffer holds entries 0 .. floor(M / s), and the walk blends
entries floor(M / s) .. 1, each with step s (entry 0 is the start and carries no step).  Every
fog blend is then the same affine map, and n of them have a closed form:

* basic fog: opacity o = s / visibility per step: colour -> a^n x + (1 - a^n) C with a = 1 - o,
  fog opacity 1 - a^n;
* glow before the fog within a step: x -> (1 - o)((1 - g) x + g G) + o C, g = glow / stepCount,
  G from stepCount = n + 1 (the order and the glow colour's step count are pinned);
* distance fog at a density so large that DistanceFogOpacity saturates at 0.3 - 0.001: colour
  distances near 0 give fog colour 3, huge ones fog colour 1, and at a shifted distance of
  ~half the first colour distance and ~1/8 of the second the mix (C1 + C2) / 2 * 7/8 + C3 / 8;
* density 0 or both fogs off: the image without fog, bit for bit.
A hit pixel blends the steps before its hit: n = ceil(depth / s) - 1 entries over the colour of
the same render without fog.
"""
import math

import numpy as np
import pytest

from mandelbulber2_amd import abi, scene
from oracle import oracle

S = 0.25
M = 10.1
N_MISS = int(math.floor(M / S))  # entries 1 .. 40


def _scene(W=32, H=24):
    p = scene.mandelbulb8(W, H)
    p.fractal.bulb.power = 2.0
    p.fractal.juliaMode = 1
    abi.set3(p.fractal.juliaC, (0.0, 0.0, 0.0))
    abi.set3(p.camera, (0.0, -3.0, 0.0))
    abi.set3(p.target, (0.0, 0.0, 0.0))
    p.advancedQuality = 1
    p.absMinMarchingStep = S
    p.absMaxMarchingStep = S
    p.relMinMarchingStep = 1e-9
    p.relMaxMarchingStep = 1e6
    p.detailSizeMin = 1e-12
    p.detailSizeMax = 1.0
    p.viewDistanceMax = M
    p.glowEnabled = 0
    return p


def _blend(x, a, b, n):
    """n applications of x -> a x + b (float64)."""
    an = a ** n
    return an * x + (1.0 - an) / (1.0 - a) * b if a != 1.0 else x + n * b


def _misses(r):
    return r["mask"] == 0


def test_basic_fog_on_misses_closed_form():
    p = _scene()
    base = oracle.render(p)
    p.basicFogEnabled = 1
    p.basicFogVisibility = 20.0
    abi.set3(p.basicFogColor, (0.2, 0.6, 0.9))
    r = oracle.render(p)
    miss = _misses(r)
    assert miss.sum() > 50
    o = S / 20.0
    C = np.array([0.2, 0.6, 0.9])
    want = _blend(base["rgb"][miss].astype(np.float64), 1.0 - o, o * C, N_MISS)
    assert np.allclose(r["rgb"][miss], want, rtol=2e-5, atol=1e-6)
    op = 1.0 - (1.0 - o) ** N_MISS
    assert np.all(np.abs(r["opacity"][miss].astype(np.float64) - op * 65535.0) <= 1.5)
    # an entry more or less (the start entry, or the last step's overshoot) is a different power
    wrong = _blend(base["rgb"][miss].astype(np.float64), 1.0 - o, o * C, N_MISS + 1)
    assert not np.allclose(r["rgb"][miss], wrong, rtol=2e-5, atol=1e-6)


def test_glow_then_fog_within_a_step():
    p = _scene()
    p.glowEnabled = 1
    p.glowIntensity = 3.0
    abi.set3(p.glowColor1, (0.9, 0.1, 0.1))
    abi.set3(p.glowColor2, (0.1, 0.1, 0.9))
    p.DEFactor = 1.0
    base = oracle.render(p)  # glow only
    p.basicFogEnabled = 1
    p.basicFogVisibility = 30.0
    abi.set3(p.basicFogColor, (0.3, 0.7, 0.2))
    r = oracle.render(p)
    miss = _misses(r)
    p0 = _scene()
    plain = oracle.render(p0)  # neither
    steps = N_MISS + 1  # stepCount = buffCount
    glow = steps * 3.0 / 512.0
    G = np.array([0.9, 0.1, 0.1]) * max(1.0 - glow, 0.0) + np.array([0.1, 0.1, 0.9]) * glow
    g = glow / steps
    o = S / 30.0
    C = np.array([0.3, 0.7, 0.2])
    x = plain["rgb"][miss].astype(np.float64)
    want = _blend(x, (1 - o) * (1 - g), (1 - o) * g * G + o * C, N_MISS)
    assert np.allclose(r["rgb"][miss], want, rtol=2e-5, atol=1e-6)
    swapped = _blend(x, (1 - o) * (1 - g), (1 - g) * o * C + g * G, N_MISS)
    assert not np.allclose(r["rgb"][miss], swapped, rtol=2e-5, atol=1e-6)
    assert np.allclose(base["rgb"][miss], _blend(x, 1 - g, g * G, N_MISS), rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("c1,c2,colour", [(1e-12, 1e-12, 2), (1e30, 1e30, 0)])
def test_saturated_distance_fog_takes_one_colour(c1, c2, colour):
    p = _scene()
    base = oracle.render(p)
    p.volFogEnabled = 1
    p.volFogDensity = 1e30
    p.volFogColour1Distance = c1
    p.volFogColour2Distance = c2
    cols = [(0.9, 0.2, 0.1), (0.1, 0.8, 0.2), (0.2, 0.3, 0.95)]
    for k in range(3):
        abi.set3(p.volFogColour[k], cols[k])
    r = oracle.render(p)
    miss = _misses(r)
    o = float(np.float32(0.3)) - 0.001  # 0.3f * X / (1 + X) at X ~ 1e30 is 0.3f
    C = np.array(cols[colour])
    want = _blend(base["rgb"][miss].astype(np.float64), 1.0 - o, o * C, N_MISS)
    assert np.allclose(r["rgb"][miss], want, rtol=2e-5, atol=1e-6)
    op = 1.0 - (1.0 - o) ** N_MISS
    assert np.all(np.abs(r["opacity"][miss].astype(np.float64) - op * 65535.0) <= 1.5)


def test_distance_fog_colour_mix():
    """distance_from_surface F = 1e6: the shifted distance |d - F| + 0.1 F is 1.1e6 to within
    1e-5 for the DE values d <= 10, so colour distances 2.2e6 and 8.8e6 give k = 1/2 and
    k2 = 1/8."""
    p = _scene()
    base = oracle.render(p)
    p.volFogEnabled = 1
    p.volFogDensity = 1e30
    p.volFogDistanceFromSurface = 1e6
    p.volFogColour1Distance = 2.2e6
    p.volFogColour2Distance = 4.4e6
    cols = [np.array((0.9, 0.2, 0.1)), np.array((0.1, 0.8, 0.2)), np.array((0.2, 0.3, 0.95))]
    for k in range(3):
        abi.set3(p.volFogColour[k], cols[k])
    r = oracle.render(p)
    miss = _misses(r)
    o = float(np.float32(0.3)) - 0.001
    C = (cols[0] + cols[1]) / 2 * (7 / 8) + cols[2] / 8
    want = _blend(base["rgb"][miss].astype(np.float64), 1.0 - o, o * C, N_MISS)
    assert np.allclose(r["rgb"][miss], want, rtol=1e-4, atol=1e-5)


def test_fog_off_or_empty_is_the_plain_image():
    p = _scene()
    base = oracle.render(p)
    p.volFogEnabled = 1
    p.volFogDensity = 0.0
    r = oracle.render(p)
    assert np.array_equal(r["rgb"], base["rgb"])
    assert not r["opacity"].any()


def test_hit_pixels_blend_the_steps_before_the_hit():
    p = _scene(40, 30)
    base = oracle.render(p)
    p.basicFogEnabled = 1
    p.basicFogVisibility = 25.0
    abi.set3(p.basicFogColor, (0.5, 0.25, 0.75))
    r = oracle.render(p)
    hit = (r["mask"] > 0) & (base["mask"] > 0)
    assert hit.sum() > 100
    # the march's step k reaches scan k s; the hit iteration is the first with k s past the
    # surface, so ceil(depth / s) - 1 of its entries are blended (the binary search ends within
    # the last step)
    n = np.ceil(base["depth"][hit].astype(np.float64) / S - 1e-9) - 1
    o = S / 25.0
    C = np.array([0.5, 0.25, 0.75])
    x = base["rgb"][hit].astype(np.float64)
    an = (1.0 - o) ** n[:, None]
    want = an * x + (1.0 - an) * C
    ok = np.all(np.isclose(r["rgb"][hit], want, rtol=2e-5, atol=1e-6), axis=1)
    assert ok.mean() > 0.99, ok.mean()

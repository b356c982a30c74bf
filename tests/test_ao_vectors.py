"""Multi-ray AO directions from the lightmap (PrepareAOVectors, render_worker.cpp:599-639):
mbx_prepare_ao_vectors against the oracle's restatement (bit for bit) and the oracle against
geometry.  Host only."""
import math
import os

import numpy as np
import pytest

from mandelbulber2_amd import runtime, scene

TEXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "textures", "lightmap.jpg")


def _lightmap():
    from PIL import Image
    return np.asarray(Image.open(TEXTURE).convert("RGB"), np.float32) / 256.0


def _vectors(p):
    n = p.aoVectorsCount
    return (np.array([list(p.aoVectors[i]) for i in range(n)]),
            np.array([list(p.aoColors[i]) for i in range(n)]))


@pytest.mark.parametrize("quality", [1, 3, 4, 7, 10])
@pytest.mark.parametrize("rotation", [0.0, 90.0, -37.5, 180.0])
@pytest.mark.parametrize("texture", ["white", "lightmap"])
def test_library_matches_oracle(oracle_mod, quality, rotation, texture):
    tex = None if texture == "white" else _lightmap()
    a, b = scene.default_params(), scene.default_params()
    for p in (a, b):
        p.ambientOcclusionQuality = quality
        p.aoLightMapRotation = rotation
    n1 = runtime.prepare_ao_vectors(a, tex)
    n2 = oracle_mod.prepare_ao_vectors(b, tex)
    assert n1 == n2 > 0
    va, ca = _vectors(a)
    vb, cb = _vectors(b)
    np.testing.assert_array_equal(va, vb)
    np.testing.assert_array_equal(ca, cb)


def test_directions_are_unit_and_turn_about_z(oracle_mod):
    """Unit directions; a lightmap rotation of 90 degrees turns every white-texture direction a
    quarter turn counter-clockwise about z: (x, y, z) -> (-y, x, z)."""
    p0, p90 = scene.default_params(), scene.default_params()
    for p, r in ((p0, 0.0), (p90, 90.0)):
        p.ambientOcclusionQuality = 6
        p.aoLightMapRotation = r
        oracle_mod.prepare_ao_vectors(p)
    v0, c0 = _vectors(p0)
    v90, _ = _vectors(p90)
    np.testing.assert_allclose(np.linalg.norm(v0, axis=1), 1.0, atol=1e-15)
    np.testing.assert_allclose(v90, np.stack([-v0[:, 1], v0[:, 0], v0[:, 2]], 1), atol=1e-15)
    assert (c0 == 1.0).all()
    # rows of constant latitude: sin(b) over b = -0.49 pi + k / quality
    rows = np.unique(np.round(v0[:, 2], 12))
    assert len(rows) == math.ceil(0.98 * math.pi * 6)


def test_colour_comes_from_the_pixel_at_the_direction(oracle_mod):
    """A texture whose pixels encode their own column and row: each kept direction's colour names
    the pixel at its longitude and latitude, shifted by half the texture (the + 8.5 W, + 8.5 H of
    the lookup): X = floor(phi / 2pi W + W / 2) mod W, Y = floor(asin(z) / pi H + H / 2) mod H."""
    W, H = 37, 23
    tex = np.zeros((H, W, 3), np.float32)
    tex[:, :, 0] = (np.arange(W)[None, :] + 1) / 64.0
    tex[:, :, 1] = (np.arange(H)[:, None] + 1) / 64.0
    p = scene.default_params()
    p.ambientOcclusionQuality = 8
    oracle_mod.prepare_ao_vectors(p, tex)
    v, c = _vectors(p)
    X = np.rint(c[:, 0] * 64).astype(int) - 1
    Y = np.rint(c[:, 1] * 64).astype(int) - 1
    phi = np.arctan2(v[:, 1], v[:, 0]) % (2 * math.pi)
    Xg = np.floor(phi / (2 * math.pi) * W + W / 2).astype(int) % W
    Yg = np.floor(np.arcsin(v[:, 2]) / math.pi * H + H / 2).astype(int) % H
    assert (Y == Yg).all()
    # the longitude is rebuilt from the rounded direction: a pixel edge may fall either side
    assert (X == Xg).mean() > 0.99 and (np.abs((X - Xg + W // 2) % W - W // 2) <= 1).all()


def test_black_lightmap_leaves_one_null_direction(oracle_mod):
    p = scene.default_params()
    p.ambientOcclusionQuality = 4
    oracle_mod.prepare_ao_vectors(p, np.zeros((8, 8, 3), np.float32))
    assert p.aoVectorsCount == 1
    assert list(p.aoVectors[0]) == [0.0, 0.0, 0.0] and list(p.aoColors[0]) == [0.0, 0.0, 0.0]
    q = scene.default_params()
    q.ambientOcclusionQuality = 4
    assert runtime.prepare_ao_vectors(q, np.zeros((8, 8, 3), np.float32)) == 1

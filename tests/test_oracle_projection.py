"""Pins of the oracle's projections and antialiasing (projection_3d.cpp:41-89, 140-159,
render_worker.cpp:96-202, 669, 688) by facts outside their code:

* fish-eye: the ray of image point q makes the angle |q| * fov with the camera's forward axis, in
  the direction of q on the image (the camera's right / up frame), fov in radians;
* equirectangular: the image is two units wide and one high whatever its pixel size, and the ray
  of image point (x, y) has longitude fov * x / 2 and latitude fov * y / 2;
* three-point: an image point at distance d from the centre lies atan(2 d tan(fov / 2)) off the
  axis, so the image edge of a square image is fov / 2 off it;
* the fish-eye cut: outside the circle of radius pi / 2 / fov the pixels are black misses of
  depth 1e20 and fog opacity 1; inside, the fish-eye image unchanged;
* the equirectangular projection halves the distance threshold: with a constant threshold the
  march stops at half of it;
* antialiasing at size n: sample r of pixel (x, y) is the ray of the n W x n H image's pixel
  (n x + r // n, n y - r % n), and with Monte Carlo samples the x step runs on past the pixel;
  the image is the mean of the samples, so it agrees with the finer image box-filtered except
  where the march's per-pixel random jitter moves a silhouette.
"""
import math

import numpy as np
import pytest

from mandelbulber2_amd import abi, scene
from oracle import oracle


def _scene(W, H, persp, fov):
    p = scene.plain_material(scene.mandelbulb8(W, H))
    p.light.castShadows = 0
    p.perspectiveType = persp
    p.fov = fov
    return p


def _local(p, x, y, repeat=0):
    m, _, _ = oracle.frame_constants(p)
    _, vv, _, _ = oracle.sample_ray(p, x, y, 1, repeat)
    v = m.T @ (vv / np.linalg.norm(vv))      # camera frame: x right, y forward, z up
    return v


def _image_point(p, x, y, aspect):
    return (x / p.imageWidth - 0.5) * aspect, 0.5 - y / p.imageHeight


@pytest.mark.parametrize("fov_deg", [180.0, 120.0, 250.0])
def test_fisheye_angle_is_linear_in_the_image_radius(fov_deg):
    p = _scene(64, 48, 1, fov_deg)
    fov = math.radians(fov_deg)
    checked = 0
    for x in range(0, 64, 7):
        for y in range(0, 48, 5):
            ix, iy = _image_point(p, x, y, 64 / 48)
            r = math.hypot(ix, iy)
            if r * fov >= math.pi:
                continue
            v = _local(p, x, y)
            assert math.acos(max(-1.0, min(1.0, v[1]))) == pytest.approx(r * fov, abs=1e-9)
            if r > 1e-9 and r * fov < math.pi - 1e-6:
                assert math.atan2(v[2], v[0]) == pytest.approx(math.atan2(iy, ix), abs=1e-9)
            checked += 1
    assert checked > 50
    # the image centre looks along the camera's forward axis
    v = _local(p, 32, 24)
    assert np.allclose(v, [0.0, 1.0, 0.0], atol=1e-15)


@pytest.mark.parametrize("W,H", [(64, 32), (50, 40)])
def test_equirectangular_longitude_and_latitude(W, H):
    p = _scene(W, H, 2, 360.0)
    for x in range(0, W, 5):
        for y in range(0, H, 4):
            ix, iy = _image_point(p, x, y, 2.0)
            v = _local(p, x, y)
            assert math.asin(v[2]) == pytest.approx(math.pi * iy, abs=1e-9)
            if abs(iy) < 0.5:
                assert math.atan2(v[0], v[1]) == pytest.approx(math.pi * ix, abs=1e-9)


def test_three_point_angle_closed_form():
    p = _scene(40, 40, 0, 60.0)
    for (x, y) in [(0, 20), (20, 0), (5, 33), (39, 39), (20, 20)]:
        ix, iy = _image_point(p, x, y, 1.0)
        v = _local(p, x, y)
        want = math.atan(math.hypot(ix, iy) * 2.0 * math.tan(math.radians(30.0)))
        assert math.acos(min(1.0, v[1])) == pytest.approx(want, abs=1e-12)
    # the left edge of the square image (ix = -0.5) is half the field of view off the axis
    v = _local(p, 0, 20)
    assert math.acos(v[1]) == pytest.approx(math.radians(30.0), abs=1e-12)


def test_fisheye_cut_blanks_outside_the_hemisphere():
    W, H = 48, 40
    cut = oracle.render(_scene(W, H, 3, 200.0), optional=True)
    fish = oracle.render(_scene(W, H, 1, 200.0), optional=True)
    yy, xx = np.mgrid[0:H, 0:W]
    r = np.hypot((xx / W - 0.5) * (W / H), 0.5 - yy / H)
    edge = math.pi * 0.5 / math.radians(200.0)
    out = r > edge
    near = np.abs(r - edge) < 1e-9          # a pixel on the circle may fall either way by rounding
    out &= ~near
    assert out.any() and (~out).any()
    assert not cut["rgb"][out].any() and np.all(cut["depth"][out] == 1e20) and not cut["mask"][out].any()
    assert np.all(cut["opacity"][out] == 65535) and not cut["world"][out].any()
    assert not cut["shadow"][out].any() and not cut["colour"][out].any()
    inside = ~out & ~near
    for k in ("rgb", "depth", "mask", "iters", "normal", "opacity", "colour"):
        assert np.array_equal(cut[k][inside], fish[k][inside]), k
    # the fish-eye without the cut shades every pixel (background at the least)
    assert fish["rgb"][out].any()


def test_equirectangular_halves_the_threshold():
    for const in (1, 0):
        thr = {}
        for persp in (0, 2):
            p = _scene(40, 40, persp, 360.0 if persp else 60.0)
            p.constantDEThreshold = const
            p.DEThresh = 0.004
            m = oracle.march_pixel(p, 20, 20)
            assert m["found"]
            thr[persp] = (m["thr"], np.linalg.norm(m["point"] - np.array(p.camera)))
        if const:
            assert thr[0][0] == 0.004 and thr[2][0] == 0.002
        else:
            # |camera - point| * resolution * fov / detail, halved
            d, fov = thr[2][1], math.radians(360.0)
            assert thr[2][0] == pytest.approx(d * (1.0 / 40) * fov / 1.0 * 0.5, rel=1e-12)


@pytest.mark.parametrize("n,mc", [(2, False), (3, False), (2, True)])
def test_antialiasing_sample_rays_are_the_finer_image_rays(n, mc):
    W, H = 20, 16
    p = _scene(W, H, 0, 53.13)
    p.antialiasingEnabled = 1
    p.antialiasingSize = n
    if mc:
        p.DOFMonteCarlo = 1
        p.DOFSamples = 2
    fine = _scene(n * W, n * H, 0, 53.13)
    repeats = n * n * (2 if mc else 1)
    for (x, y) in [(3, 4), (10, 8), (19, 15), (0, 1)]:
        for r in range(repeats):
            xs, ys = n * x + r // n, n * y - r % n
            _, vv, _, _ = oracle.sample_ray(p, x, y, 1, r)
            if 0 <= xs < n * W and 0 <= ys < n * H:
                _, wv, _, _ = oracle.sample_ray(fine, xs, ys, 1)
                np.testing.assert_allclose(vv / np.linalg.norm(vv), wv / np.linalg.norm(wv), rtol=0, atol=1e-14)
    # no random sub-pixel jitter is drawn when the grid is on
    _, _, _, s_after = oracle.sample_ray(p, 3, 4, 12345)
    assert s_after == 12345


def _const_thr(p):
    # a threshold that does not depend on the pixel size, so that both images march alike
    p.constantDEThreshold = 1
    p.DEThresh = 0.002
    return p


def test_antialiasing_is_the_mean_of_the_finer_image():
    W, H, n = 48, 36, 2
    p = _const_thr(_scene(W, H, 0, 53.13))
    p.antialiasingEnabled = 1
    p.antialiasingSize = n
    aa = oracle.render(p)
    fine = oracle.render(_const_thr(_scene(n * W, n * H, 0, 53.13)))["rgb"].astype(np.float64)
    # sample (i, j) of pixel (x, y) is the fine pixel (2x + i, 2y - j): rows 2y - 1 .. 2y
    box = np.zeros((H, W, 3))
    for i in range(n):
        for j in range(n):
            rows = np.arange(H) * n - j
            box += fine[np.clip(rows, 0, None)][:, np.arange(W) * n + i]
    box /= n * n
    err = np.abs(aa["rgb"][1:] - box[1:]).max(axis=-1)
    assert (err <= 1e-3).mean() >= 0.99, (err <= 1e-3).mean()
    # the grid read downwards instead would miss the silhouette rows
    wrong = np.zeros((H, W, 3))
    for i in range(n):
        for j in range(n):
            rows = np.clip(np.arange(H) * n + j, None, n * H - 1)
            wrong += fine[rows][:, np.arange(W) * n + i]
    wrong /= n * n
    assert (np.abs(aa["rgb"][1:] - wrong[1:]).max(axis=-1) <= 1e-3).mean() < 0.5
    # with size 1 and no Monte Carlo samples the image is the plain one
    p.antialiasingSize = 1
    one = oracle.render(p)
    plain = oracle.render(_const_thr(_scene(W, H, 0, 53.13)))
    assert np.array_equal(one["rgb"], plain["rgb"]) and np.array_equal(one["depth"], plain["depth"])

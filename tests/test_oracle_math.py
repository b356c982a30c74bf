"""The oracle against what the mathematics fixes: closed forms, invariants, brute force.

Each check is chosen so that a plausible restatement mistake (a swapped axis, a wrong sign, an
off-by-one iteration count, a dropped term) fails it.
"""
import math

import numpy as np
import pytest

from mandelbulber2_amd import abi, scene


def multibrot_escape(c, power, bailout, maxiter):
    """Escape-time count of w <- w^power + c (w0 = c) with |w| > bailout, numbered as the
    reference's Compute does (iteration index + 1; maxiter + 1 when the orbit never escapes)."""
    w = c.copy()
    iters = np.full(c.shape, maxiter + 1, np.int32)
    alive = np.ones(c.shape, bool)
    for i in range(maxiter):
        w[alive] = w[alive] ** power + c[alive]
        esc = alive & (np.abs(w) > bailout)
        iters[esc] = i + 1
        alive &= ~esc
        if not alive.any():
            break
    return iters


def test_mandelbulb_plane_is_multibrot(oracle_mod):
    # On z = 0, theta0 = asin(0) = 0, so FormulaCode + c is w -> w^p + c on w = x + iy
    # (fractal_mandelbulb.cpp:32-42).  Escape counts equal the complex escape-time loop.
    for power in (2.0, 8.0):
        p = scene.default_params()
        p.fractal.formula = abi.MBX_FORMULA_MANDELBULB
        p.fractal.bulb.power = power
        p.fractal.formulaMaxiter = 60
        rs = np.random.default_rng(int(power))
        xy = rs.uniform(-1.5, 1.5, size=(3000, 2))
        pts = np.column_stack([xy, np.zeros(len(xy))])
        _, iters, maxiter, _ = oracle_mod.compute(p, pts)
        ref = multibrot_escape(xy[:, 0] + 1j * xy[:, 1], int(power), 10.0, 60)
        agree = (iters == ref).mean()
        assert agree > 0.995, agree
        assert (np.abs(iters - ref) <= 1).mean() > 0.999
        # non-escaping orbits are reported as maxN + 1 with the maxiter flag set
        assert np.all((iters == 61) == (maxiter == 1))


def test_compute_invariants(oracle_mod):
    p = scene.default_params()
    p.fractal.formula = abi.MBX_FORMULA_MANDELBULB
    p.fractal.bulb.power = 8.0
    far = np.array([[20.0, 3.0, -4.0], [0.0, 0.0, 11.0]])
    d, iters, mx, z = oracle_mod.compute(p, far)
    # |p| > bailout escapes after the first formula call: iters = 1, DE = 8 r^7 + 1
    assert list(iters) == [1, 1]
    r0 = np.linalg.norm(far, axis=1)
    r1 = np.linalg.norm(z, axis=1)
    np.testing.assert_allclose(d, 0.5 * r1 * np.log(r1) / (8 * r0 ** 7 + 1), rtol=1e-12)
    # points inside the set never escape: iteration count maxN + 1, log-DE 0 for r <= 1
    d0, it0, mx0, _ = oracle_mod.compute(p, np.array([[0.1, -0.05, 0.02], [0.0, 0.3, 0.1]]))
    assert list(it0) == [251, 251] and list(mx0) == [1, 1]
    assert np.all(d0 == 0.0)
    # the origin makes asin(0/0) NaN: the orbit reverts to the last finite z and stops with the
    # maxiter flag after the first call (compute_fractal.cpp:303-309)
    d1, it1, mx1, z1 = oracle_mod.compute(p, np.zeros((1, 3)))
    assert it1[0] == 1 and mx1[0] == 1 and d1[0] == 0.0 and np.all(z1 == 0.0)


def test_mandelbox_linear_regime(oracle_mod):
    # Inside the folding box and the minimum sphere, one step is z -> (fR2/mR2) scale z + c and
    # DE -> (fR2/mR2) |scale| DE + 1 (fractal_mandelbox.cpp:87-106): with scale 2, factor 4 the
    # orbit of a tiny p is a_n p with a_n = (8^(n+1) - 1) / 7 until |a_n p|^2 >= 0.25.
    p = scene.default_params()
    p.fractal.formula = abi.MBX_FORMULA_MANDELBOX
    p.fractal.formulaMaxiter = 50
    pt = np.array([[1e-6, -2e-6, 0.5e-6]])
    n_lin = 0
    a = 1
    while (a * 8 + 1) ** 2 * np.sum(pt ** 2) < 0.25:
        a = a * 8 + 1
        n_lin += 1
    z4 = np.array([[pt[0, 0], pt[0, 1], pt[0, 2], 0.0]])
    aux = np.array([[np.linalg.norm(pt), 1.0, 1.0]])
    for k in range(n_lin):
        z4, aux = oracle_mod.formula_step(p, z4, aux)
        z4[0, :3] += pt[0]
        aux[0, 0] = np.linalg.norm(z4[0])
    a_n = (8 ** (n_lin + 1) - 1) // 7
    np.testing.assert_allclose(z4[0, :3], a_n * pt[0], rtol=1e-12)
    assert aux[0, 1] == a_n


def test_camera_frame_geometry(oracle_mod):
    rs = np.random.default_rng(11)
    for _ in range(20):
        p = scene.default_params()
        cam = rs.normal(size=3) * rs.uniform(2, 8)
        tgt = rs.normal(size=3) * 0.3
        abi.set3(p.camera, cam)
        abi.set3(p.target, tgt)
        abi.set3(p.topVector, (0.0, 0.0, 1.0))
        m, fov, light = oracle_mod.frame_constants(p)
        fwd = (tgt - cam) / np.linalg.norm(tgt - cam)
        np.testing.assert_allclose(m @ m.T, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(m) - 1.0) < 1e-12
        # image-plane axes: +y of the camera space is forward, +z is up, +x is right
        np.testing.assert_allclose(m @ [0, 1, 0], fwd, atol=1e-12)
        up = np.array([0, 0, 1.0]) - fwd[2] * fwd
        up /= np.linalg.norm(up)
        np.testing.assert_allclose(m @ [0, 0, 1], up, atol=1e-12)
        np.testing.assert_allclose(m @ [1, 0, 0], np.cross(fwd, up), atol=1e-12)
        assert abs(fov - 2 * math.tan(53.13 / 360 * math.pi)) < 1e-15
        # light #1 with the default relative rotation is a unit vector
        assert abs(np.linalg.norm(light) - 1.0) < 1e-12


def test_center_ray_points_at_target(oracle_mod):
    p = scene.mandelbulb8(64, 48)
    r = oracle_mod.march_pixel(p, 32, 24)
    np.testing.assert_allclose(r["dir"], [0.0, 1.0, 0.0], atol=1e-15)
    assert r["found"]
    # the hit lies on the +y ray from the camera at (0, -3.5, 0)
    np.testing.assert_allclose(r["point"], [0.0, -3.5 + r["depth"], 0.0], atol=1e-12)


def test_march_depth_brackets_dense_sampling(oracle_mod):
    """Brute force: a dense fixed-step FP64 walk along the same ray finds the first point with
    DE < threshold; the march's depth lies in that bracket (up to the refinement window)."""
    p = scene.mandelbulb8(96, 72)
    m, fov, _ = oracle_mod.frame_constants(p)
    cam = np.array(p.camera[:])
    checked = 0
    in_bracket = 0
    rs = np.random.default_rng(5)
    for _ in range(40):
        x, y = int(rs.integers(20, 76)), int(rs.integers(15, 57))
        r = oracle_mod.march_pixel(p, x, y)
        if not r["found"]:
            continue
        d = r["dir"]
        t = np.arange(0.0, r["depth"] + 0.02, 2e-4)
        pts = cam[None, :] + t[:, None] * d[None, :]
        thr = np.linalg.norm(pts - cam, axis=1) * (1.0 / 72) * fov
        de, _ = oracle_mod.distance(p, pts, 1e-3)
        # distance() takes one detail size; thresholds vary per point but the DE of an analytic
        # Mandelbulb does not depend on the detail size
        inside = np.nonzero(de < thr)[0]
        assert len(inside) > 0
        k = inside[0]
        lo, hi = t[max(k - 1, 0)], t[k]
        # the refinement stops where DE is within 0.1% below the threshold; allow 1% of the
        # threshold for the DE's slope along the ray
        tol = 0.01 * thr[k]
        assert r["depth"] >= lo - tol, (x, y, r["depth"], lo, hi)  # never stops early
        if r["depth"] <= hi + tol:
            in_bracket += 1
        checked += 1
    assert checked >= 10
    assert in_bracket >= 0.9 * checked, (in_bracket, checked)


def test_normal_matches_gradient(oracle_mod):
    """CalculateNormals is a central difference of the DE with delta = threshold; compare with a
    4th-order stencil of the same field at a quarter of that step."""
    p = scene.mandelbulb8(96, 72)
    rs = np.random.default_rng(9)
    good = 0
    for _ in range(30):
        x, y = int(rs.integers(20, 76)), int(rs.integers(15, 57))
        r = oracle_mod.march_pixel(p, x, y)
        if not r["found"]:
            continue
        n = oracle_mod.normal(p, r["point"], r["thr"])
        h = r["thr"] * 0.25
        g = np.zeros(3)
        for a in range(3):
            e = np.zeros(3)
            e[a] = h
            pts = np.stack([r["point"] + 2 * e, r["point"] + e, r["point"] - e, r["point"] - 2 * e])
            dd, _ = oracle_mod.distance(p, pts, r["thr"], normal_calc=True)
            g[a] = (-dd[0] + 8 * dd[1] - 8 * dd[2] + dd[3]) / (12 * h)
        g /= np.linalg.norm(g)
        assert abs(np.linalg.norm(n) - 1) < 1e-12
        if np.dot(n, g) > 0.9:
            good += 1
        # the normal faces the camera side of the surface
        assert np.dot(n, r["dir"]) < 0.5
    assert good >= 8


def test_background_gradient_and_sun(oracle_mod):
    p = scene.mandelbulb8(64, 48)
    p.light.enabled = 0
    up = oracle_mod.background(p, [0, 0, 1])
    down = oracle_mod.background(p, [0, 0, -1])
    side = oracle_mod.background(p, [1, 0, 0])
    np.testing.assert_allclose(up, p.backgroundColor1[:], atol=1e-6)
    np.testing.assert_allclose(down, p.backgroundColor3[:], atol=1e-6)
    np.testing.assert_allclose(side, p.backgroundColor2[:], atol=1e-6)
    p.light.enabled = 1
    _, _, light = oracle_mod.frame_constants(p)
    sun = oracle_mod.background(p, light)
    p.light.enabled = 0
    sky = oracle_mod.background(p, light)
    # looking straight into the directional light adds intensity * visibility * colour
    want = np.asarray(p.light.color[:]) * p.light.intensity * p.light.visibility
    np.testing.assert_allclose(sun - sky, want, atol=1e-6)


def test_pixel_seed_range(oracle_mod):
    for pix in (0, 1, 12345, 2 ** 31, 2 ** 32 - 1):
        for fs in (0, 7, 2 ** 32 - 1):
            s = oracle_mod.pixel_seed(pix, fs)
            assert 1 <= s <= 2147483646


@pytest.mark.parametrize("rotation,perm", [
    ((0.0, 0.0, 90.0), lambda q: (-q[1], q[0], q[2])),   # Rz(90): (x, y, z) -> (-y, x, z)
    ((90.0, 0.0, 0.0), lambda q: (q[0], -q[2], q[1])),   # Rx(90): (x, y, z) -> (x, -z, y)
    ((0.0, 90.0, 0.0), lambda q: (q[2], q[1], -q[0])),   # Ry(90): (x, y, z) -> (z, y, -x)
    ((90.0, 0.0, 90.0), lambda q: (q[2], q[0], q[1])),   # Rz(90) Rx(90): x first, then z
])
def test_node_rotation_convention(oracle_mod, rotation, perm):
    """A rotated fractal node evaluates the formula at R p with R = Rz(z) Ry(y) Rx(x): quarter
    turns permute coordinates exactly, which pins each axis' sign and the composition order."""
    rs = np.random.default_rng(11)
    pts = rs.uniform(-1.2, 1.2, size=(64, 3))
    p0 = scene.mandelbulb8(64, 48)
    p0.fractal.bulb.power = 5.0  # no symmetry that would hide a wrong permutation
    pr = scene.copy_params(p0)
    abi.set3(pr.formulaRotation, rotation)
    d_rot, it_rot = oracle_mod.distance(pr, pts, 1e-4)
    d_ref, it_ref = oracle_mod.distance(p0, np.array([perm(q) for q in pts]), 1e-4)
    np.testing.assert_allclose(d_rot, d_ref, rtol=1e-9, atol=1e-9)  # cos(pi/2) = 6e-17, not 0
    assert (it_rot == it_ref).mean() >= 0.98


def test_rotated_node_with_counter_rotated_camera_is_invariant(oracle_mod):
    """Rotating the fractal by R and the camera, target and top vector by R^T leaves every ray's
    DE samples unchanged, so mask, iteration counts and depth agree with the unrotated scene."""
    p0 = scene.mandelbulb8(64, 48)
    rot = (23.0, -41.0, 67.0)
    a, b, c = (math.radians(v) for v in rot)
    Rx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    Ry = np.array([[math.cos(b), 0, math.sin(b)], [0, 1, 0], [-math.sin(b), 0, math.cos(b)]])
    Rz = np.array([[math.cos(c), -math.sin(c), 0], [math.sin(c), math.cos(c), 0], [0, 0, 1]])
    R = Rz @ Ry @ Rx
    pr = scene.copy_params(p0)
    abi.set3(pr.formulaRotation, rot)
    for name in ("camera", "target", "topVector"):
        abi.set3(getattr(pr, name), R.T @ np.array(list(getattr(p0, name))))
    r0 = oracle_mod.render(p0)
    r1 = oracle_mod.render(pr)
    assert (r0["mask"] == r1["mask"]).mean() >= 0.99
    assert (r0["iters"] == r1["iters"]).mean() >= 0.97
    both = (r0["mask"] > 0) & (r1["mask"] > 0)
    rel = np.abs(r0["depth"][both] - r1["depth"][both]) / r0["depth"][both]
    assert np.median(rel) < 1e-6

"""Pins of the oracle's shading, delta-DE and rotation restatements by facts independent of their
code: dense brute-force ray walks, closed forms, an independent forward-mode Jacobian, and exact
quarter-turn permutations derived by hand from the reference's rotation definitions.

* AuxShadow (shader_aux_shadow.cpp:94-308): the penetrating hard shadow of a hit point is
  max(0, 1 - (distance - i)/distance) at the first i >= distThresh where DE < distThresh along the
  light vector, with distance = delta / resolution = |camera - point| * fov (CalcDelta,
  render_worker.cpp:684-690); found by a dense walk at distThresh / 4 steps.
* Multi-ray AO (shader_ambient_occlusion.cpp:59-105): the same for each AO direction over
  end = delta / resolution, averaged with the directions' colours.
* Fast AO (shader_fast_ambient_occlusion.cpp:40-66): with ambient_occlusion_fast_tune = 0 the
  DE drops out: max(0, 1 - 0.2 sum_{i=1}^{q^2-1} i^2 2^-i); the result is linear in the tune.
* GI (shader_global_illumination.cpp:58-206): reflections_max = 0 changes no pixel; a bounce ray
  that misses the fractal adds clamp(background(dir), 0, limit) * colour with
  dir = normalize(n + 1.2 u), u from three Random(20000) draws of the pixel stream.
* Delta DE (calculate_distance.cpp:609-711): on orbits that stay in the Mandelbox's affine
  regimes (box folds, inner sphere fold or none) under the finite-difference offsets, the
  4-orbit difference equals 0.5 r / |J^T z / r| (linear) or 0.5 r ln r / |J^T z / r| (log), with J
  the orbit's Jacobian propagated independently here in numpy.
* Rotations (algebra.cpp:198-311, fractal.cpp:655-672, fractal_mandelbox.cpp:33-63,103,
  fractal_kaleidoscopic_ifs.cpp:44-53): quarter turns of the Mandelbox main rotation and of a
  rotated fold frame, and of the IFS main rotation, give hand-derived permutations.
* Light vector (light.cpp:90-190): quarter turns (90.4 degrees, the reference divides by 180.8)
  give hand-derived axis vectors, relative to the camera and absolute.
"""
import math

import numpy as np
import pytest

from mandelbulber2_amd import abi, scene
from oracle import oracle


def _hits(p, step=11):
    """Oracle hit points of a pixel grid: (point, distThresh)."""
    out = []
    for y in range(5, p.imageHeight, step):
        for x in range(5, p.imageWidth, step):
            m = oracle.march_pixel(p, x, y)
            if m["found"]:
                out.append((np.array(m["point"]), m["thr"]))
    return out


def _first_crossing(p, point, direction, thr, start, end, h):
    """First r in [start, end) on a grid of step h with DE(point + r dir) < thr, or None."""
    r = np.arange(start, end, h)
    d, _ = oracle.distance(p, point[None, :] + r[:, None] * np.asarray(direction)[None, :], thr)
    k = np.nonzero(d < np.float32(thr))[0]
    return (r[k[0]] if len(k) else None), (float(d.min()) if len(d) else np.inf)


def _fov(p):
    return 2.0 * math.tan(math.radians(p.fov) / 2.0)


# ------------------------------------------------------------------------------------------
# AuxShadow and multi-ray AO against dense walks
# ------------------------------------------------------------------------------------------
def test_aux_shadow_against_dense_walk():
    p = scene.mandelbulb8(200, 150)
    p.light.castShadows = 1
    p.light.penetrating = 1
    p.light.softShadowCone = 0.0
    _, _, light = oracle.frame_constants(p)
    cam = np.array(list(p.camera))
    errs, shadowed = [], 0
    for pt, thr in _hits(p, step=9):
        distance = np.linalg.norm(cam - pt) * _fov(p)
        sh = float(oracle.shader_probe(p, oracle.SHADOW, pt, (0.0, 0.0, 1.0), thr)[0])
        istar, _ = _first_crossing(p, pt, light, thr, thr, distance, thr / 4.0)
        expect = 1.0 if istar is None else max(0.0, 1.0 - (distance - istar) / distance)
        shadowed += istar is not None
        # the tracer stops at its first sample with DE < thr: never before the dense crossing
        # (one dense step of slack), and at most a fraction of a threshold after it
        errs.append((sh - expect) * distance / thr)
    errs = np.array(errs)
    assert len(errs) >= 60 and shadowed >= 15, (len(errs), shadowed)
    assert errs.min() >= -0.26, errs.min()
    assert np.quantile(np.abs(errs), 0.9) <= 0.5 and np.abs(errs).max() <= 1.5, np.sort(np.abs(errs))[-5:]


def test_multi_ray_ao_against_dense_walks():
    p = scene.mandelbox_shadow_ao(96, 54, seed=1)
    dirs = np.array(scene.fibonacci_hemisphere(16))
    cam = np.array(list(p.camera))
    hits = _hits(p, step=13)[:40]
    assert len(hits) >= 20
    errs, occluded = [], 0
    for pt, thr in hits:
        end = np.linalg.norm(cam - pt) * _fov(p)
        for v in dirs[::3]:
            p.aoVectorsCount = 1
            abi.set3(p.aoVectors[0], tuple(v))
            abi.set3(p.aoColors[0], (1.0, 1.0, 1.0))
            ao = float(oracle.shader_probe(p, oracle.AO_RAYS, pt, (0.0, 0.0, 1.0), thr)[0])
            rstar, _ = _first_crossing(p, pt, v, thr, thr, end, thr / 4.0)
            expect = 1.0 if rstar is None else max(0.0, 1.0 - (end - rstar) / end)
            occluded += rstar is not None
            errs.append((ao - expect) * end / thr)
    errs = np.array(errs)
    assert occluded >= 20, occluded
    # steps of twice the DE may overshoot a crossing by up to one DE: slack of a few thresholds
    assert errs.min() >= -0.26, errs.min()
    assert np.quantile(np.abs(errs), 0.9) <= 2.0, np.quantile(np.abs(errs), 0.9)


def test_multi_ray_ao_averages_direction_colours():
    """Two directions with different colours: the AO colour is the colour-weighted mean of the
    two single-direction results."""
    p = scene.mandelbox_shadow_ao(96, 54, seed=1)
    dirs = np.array(scene.fibonacci_hemisphere(16))
    pt, thr = _hits(p, step=13)[5]
    single = []
    for v in (dirs[2], dirs[9]):
        p.aoVectorsCount = 1
        abi.set3(p.aoVectors[0], tuple(v))
        abi.set3(p.aoColors[0], (1.0, 1.0, 1.0))
        single.append(float(oracle.shader_probe(p, oracle.AO_RAYS, pt, (0.0, 0.0, 1.0), thr)[0]))
    p.aoVectorsCount = 2
    abi.set3(p.aoVectors[0], tuple(dirs[2]))
    abi.set3(p.aoVectors[1], tuple(dirs[9]))
    abi.set3(p.aoColors[0], (1.0, 0.5, 0.0))
    abi.set3(p.aoColors[1], (0.0, 0.25, 1.0))
    ao = oracle.shader_probe(p, oracle.AO_RAYS, pt, (0.0, 0.0, 1.0), thr)
    expect = np.array([single[0] * 1.0, single[0] * 0.5 + single[1] * 0.25, single[1] * 1.0]) / 2.0
    np.testing.assert_allclose(ao, expect, rtol=1e-6, atol=1e-7)


# ------------------------------------------------------------------------------------------
# fast AO closed form and linearity
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("quality,expect", [(1, 1.0), (2, 1.0 - 0.2 * (1 / 2 + 4 / 4 + 9 / 8)), (3, 0.0)])
def test_fast_ao_closed_form_without_tune(quality, expect):
    p = scene.mandelbulb8(120, 90)
    p.ambientOcclusionQuality = quality
    p.ambientOcclusionFastTune = 0.0
    for pt, thr in _hits(p, step=17)[:10]:
        n = oracle.normal(p, pt, thr)
        ao = oracle.shader_probe(p, oracle.AO_FAST, pt, n, thr)
        np.testing.assert_allclose(ao, [expect] * 3, rtol=0, atol=2e-7)


def test_fast_ao_is_linear_in_the_tune():
    p = scene.mandelbulb8(120, 90)
    p.ambientOcclusionQuality = 2
    hits = _hits(p, step=17)[:12]
    vals = {}
    for t in (0.0, 0.5, 1.0):
        p.ambientOcclusionFastTune = t
        vals[t] = np.array([oracle.shader_probe(p, oracle.AO_FAST, pt, oracle.normal(p, pt, thr), thr)[0]
                            for pt, thr in hits])
    inside = (vals[1.0] > 0.0) & (vals[0.0] > 0.0)
    assert inside.sum() >= 5
    np.testing.assert_allclose(vals[0.5][inside], 0.5 * (vals[0.0] + vals[1.0])[inside], atol=2e-6)
    # the tune subtracts a non-negative DE term: a larger tune never darkens
    assert (vals[1.0] >= vals[0.0] - 1e-7).all()


# ------------------------------------------------------------------------------------------
# global illumination
# ------------------------------------------------------------------------------------------
def test_gi_depth_zero_changes_no_pixel():
    p = scene.mandelbulb8(64, 48)
    p.DOFMonteCarlo = 1
    p.DOFSamples = 3
    p.DOFMonteCarloGlobalIllumination = 0
    a = oracle.render(p)
    p.DOFMonteCarloGlobalIllumination = 1
    p.reflectionsMax = 0
    b = oracle.render(p)
    for key in ("rgb", "depth", "iters", "mask", "steps"):
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)


def test_gi_bounce_that_misses_adds_the_background():
    p = scene.mandelbulb8(160, 120)
    p.DOFMonteCarlo = 1
    p.DOFMonteCarloGlobalIllumination = 1
    p.reflectionsMax = 1
    p.monteCarloGIRadianceLimit = 10.0
    colour = np.array([0.7, 0.4, 0.2], np.float32)
    checked = 0
    for k, (pt, thr) in enumerate(_hits(p, step=7)):
        seed = 1000 + 7 * k
        n = np.array(oracle.normal(p, pt, thr))
        u, _ = oracle.random_ints(seed, 3, 20000)
        u = (np.array(u, np.float64) - 10000.0) / 10000.0
        d = n + 1.2 * u
        d /= np.linalg.norm(d)
        # a clear miss: the DE stays above 1.3 thresholds at half-threshold samples from the GI
        # march's start (scan = thr) outwards, so no point of the march comes within one
        # threshold (the DE changes by at most the distance moved)
        _, dmin = _first_crossing(p, pt, d, thr, thr, 6.0, thr / 2.0)
        if dmin < 1.3 * thr:
            continue
        gi = oracle.shader_probe(p, oracle.GI, pt, n, thr, seed=seed, colour=colour)
        bg = np.clip(np.asarray(oracle.background(p, d), np.float32), 0.0, 10.0)
        np.testing.assert_allclose(gi, bg * colour, rtol=2e-6, atol=1e-7)
        checked += 1
    assert checked >= 10, checked


# ------------------------------------------------------------------------------------------
# delta DE against an independent forward-mode Jacobian
# ------------------------------------------------------------------------------------------
def _mandelbox_orbit(pts, scale=2.0, L=1.0, V=2.0, rmin=0.5, rfix=1.0, bailout=100.0, maxiter=50):
    """Mandelbox orbit of many points with c = p, and dz/dp propagated alongside (numpy, FP64).
    Returns final z, J (n x 3 x 3), iteration count (i + 1 at escape, maxiter + 1 otherwise),
    an escape flag and the per-iteration regime pattern (box folds and sphere-fold regime)."""
    n = len(pts)
    z = pts.copy()
    J = np.repeat(np.eye(3)[None], n, axis=0)
    mR2, fR2 = rmin * rmin, rfix * rfix
    alive = np.ones(n, bool)
    iters = np.full(n, maxiter + 1)
    pattern = np.zeros((n, maxiter, 4), np.int8)
    affine = np.ones(n, bool)
    for i in range(maxiter):
        a = alive
        fold = np.abs(z) > L
        zf = np.where(fold, np.sign(z) * V - z, z)
        Jf = np.where(fold[:, :, None], -J, J)
        r2 = (zf * zf).sum(1)
        inner = r2 < mR2
        outer = ~inner & (r2 < fR2)
        k = np.where(inner, fR2 / mR2, np.where(outer, fR2 / np.where(r2 > 0, r2, 1.0), 1.0))
        zn = zf * k[:, None] * scale + pts
        Jn = Jf * (k * scale)[:, None, None] + np.eye(3)[None]
        pattern[:, i, :3] = np.where(a[:, None], fold, 0)
        pattern[:, i, 3] = np.where(a, inner * 1 + outer * 2, 0)
        affine &= ~(a & outer)
        z = np.where(a[:, None], zn, z)
        J = np.where(a[:, None, None], Jn, J)
        esc = a & (np.sqrt((z * z).sum(1)) > bailout)
        iters = np.where(esc, i + 1, iters)
        alive = a & ~esc
    return z, J, iters, ~alive, pattern, affine


@pytest.mark.parametrize("log_de", [False, True])
def test_delta_de_against_forward_mode_jacobian(log_de):
    p = scene.mandelbox_shadow_ao(64, 36, seed=1)
    p.fractal.deltaDEFunction = 2 if log_de else 1
    rs = np.random.default_rng(11)
    pts = rs.uniform(-3.0, 3.0, size=(40000, 3))
    detail = 1e-3
    z, J, iters, escaped, pattern, affine = _mandelbox_orbit(pts)
    # the reference's offset orbits: delta = max(|p| 1e-14, 1e-5 detail), capped at the orbit's
    # iteration count; keep the points whose offset orbits take the same regimes and escape alike
    delta = np.maximum(np.linalg.norm(pts, axis=1) * 1e-14, 1e-5 * detail)
    keep = escaped & affine
    for kdim in range(3):
        q = pts.copy()
        q[:, kdim] += delta
        _, _, it2, esc2, pat2, _ = _mandelbox_orbit(q)
        keep &= esc2 & (it2 == iters) & (pat2 == pattern).all(axis=(1, 2))
    r = np.linalg.norm(z, axis=1)
    grad = np.einsum("nij,ni->nj", J, z) / r[:, None]        # d|z|/dp = J^T z / |z|
    dr = np.linalg.norm(grad, axis=1)
    expect = (0.5 * r * np.log(r) if log_de else 0.5 * r) / dr
    keep &= (expect > 1e-4) & (expect < 9.0)
    assert keep.sum() >= 300, keep.sum()
    d, it = oracle.distance(p, pts[keep], detail)
    np.testing.assert_array_equal(it, iters[keep])
    rel = np.abs(d - expect[keep]) / expect[keep]
    assert np.quantile(rel, 0.99) < 1e-5 and rel.max() < 1e-4, (np.quantile(rel, 0.99), rel.max())


# ------------------------------------------------------------------------------------------
# rotations: quarter turns against hand-derived permutations
# ------------------------------------------------------------------------------------------
# RotateX/Y/Z (algebra.cpp:198-247) right-multiply the counter-clockwise matrices, so a quarter
# turn maps (x, y, z) to: about x (x, -z, y); about y (z, y, -x); about z (-y, x, z).
QX = lambda v: np.array([v[0], -v[2], v[1]])
QY = lambda v: np.array([v[2], v[1], -v[0]])
QZ = lambda v: np.array([-v[1], v[0], v[2]])


def _box_params():
    p = scene.mandelbox_shadow_ao(32, 18, seed=1)
    mb = p.fractal.mandelbox
    abi.set3(mb.offset, (0.0, 0.0, 0.0))
    abi.set3(mb.colorFactor, (0.03, 0.05, 0.07))
    return p


@pytest.mark.parametrize("angles,perm", [((90.0, 0.0, 0.0), QX), ((0.0, 90.0, 0.0), QY), ((0.0, 0.0, 90.0), QZ),
                                         # SetRotation2 = Rz Ry Rx: x first, then y
                                         ((90.0, 90.0, 0.0), lambda v: QY(QX(v)))])
def test_mandelbox_main_rotation_quarter_turns(angles, perm):
    rs = np.random.default_rng(3)
    z4 = np.zeros((64, 4))
    z4[:, :3] = rs.uniform(-2.5, 2.5, size=(64, 3))
    aux = np.tile([0.0, 1.0, 0.0], (64, 1))
    p = _box_params()
    z0, a0 = oracle.formula_step(p, z4, aux)
    p.fractal.mandelbox.mainRotationEnabled = 1
    abi.set3(p.fractal.mandelbox.rotationMain, angles)
    z1, a1 = oracle.formula_step(p, z4, aux)
    expect = np.array([perm(v) for v in z0[:, :3]])
    np.testing.assert_allclose(z1[:, :3], expect, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(a1[:, 1:], a0[:, 1:])   # DE and colour do not see the rotation


def test_mandelbox_rotated_fold_frames():
    """rot[0][x] a quarter turn about z: the x-axis fold of the first frame acts on Rz z = (-y, x, z),
    so a point with y < -limit is folded in y (as the plain fold would) but counted as an x fold
    (colour factor x); with y > limit the frame's fold does not fire and the plain y fold does.
    With every angle zero the rotated folds equal the plain folds bit for bit."""
    p = _box_params()
    mb = p.fractal.mandelbox
    z4 = np.array([[0.3, -1.5, 0.2, 0.0], [0.3, 1.5, 0.2, 0.0], [1.6, 0.4, -0.3, 0.0], [-0.2, 0.7, 1.3, 0.0]])
    aux = np.tile([0.0, 1.0, 0.0], (4, 1))
    z0, a0 = oracle.formula_step(p, z4, aux)
    mb.rotationsEnabled = 1
    zr, ar = oracle.formula_step(p, z4, aux)
    np.testing.assert_array_equal(zr, z0)
    np.testing.assert_array_equal(ar, a0)
    mb.rotation[0][0][2] = 90.0
    z1, a1 = oracle.formula_step(p, z4, aux)
    same = [0, 1, 3]
    np.testing.assert_allclose(z1[same], z0[same], rtol=0, atol=1e-12)
    cf = (0.03, 0.05, 0.07)
    # colour after the step: the fold factors of the folds that fired, then the sphere fold's
    # factor (the same in both runs: the folded point is the same)
    np.testing.assert_allclose(a1[0, 2] - a0[0, 2], cf[0] - cf[1], atol=1e-15)  # x frame instead of y
    np.testing.assert_allclose(a1[1, 2] - a0[1, 2], 0.0, atol=1e-15)             # plain y fold
    # a point folded in x by the plain fold: the frame maps x to y, so the x fold moves to the
    # identity second frame (x < -limit fails for x = 1.6): not folded in x at all
    assert a1[2, 2] != a0[2, 2] or not np.allclose(z1[2], z0[2])


@pytest.mark.parametrize("angles,perm", [((90.0, 0.0, 0.0), QZ), ((0.0, 90.0, 0.0), QY), ((0.0, 0.0, 90.0), QX)])
def test_ifs_main_rotation_quarter_turns(angles, perm):
    """SetRotation3 (algebra.cpp:305-311) = Rz(x) Ry(y) Rx(z); the IFS step rotates about the
    offset: z = s (R (z - o) + o) - o (s - 1) without abs folds, planes and edges."""
    p = scene.menger_ifs(32, 24)
    ifs = p.fractal.IFS
    ifs.absX = ifs.absY = ifs.absZ = 0
    ifs.mengerSpongeMode = 0
    ifs.edgeEnabled = 0
    for i in range(abi.MBX_IFS_VECTOR_COUNT):
        ifs.enabled[i] = 0
    ifs.rotationEnabled = 1
    abi.set3(ifs.rotation, angles)
    s, o = ifs.scale, np.array(list(ifs.offset))
    rs = np.random.default_rng(5)
    z4 = np.zeros((32, 4))
    z4[:, :3] = rs.uniform(-1.5, 1.5, size=(32, 3))
    aux = np.tile([0.0, 1.0, 0.0], (32, 1))
    z1, a1 = oracle.formula_step(p, z4, aux)
    expect = np.array([s * (perm(v - o) + o) - o * (s - 1.0) for v in z4[:, :3]])
    np.testing.assert_allclose(z1[:, :3], expect, rtol=0, atol=1e-12)
    np.testing.assert_allclose(a1[:, 1], abs(s), rtol=0, atol=0)


def test_ifs_plane_rotation_quarter_turn():
    """IFS plane i with rotations (90, 0, 0) (SetRotation3: about z) and direction x: the point is
    turned first, then reflected at distance 0 when its turned x is negative."""
    p = scene.menger_ifs(32, 24)
    ifs = p.fractal.IFS
    ifs.absX = ifs.absY = ifs.absZ = 0
    ifs.mengerSpongeMode = 0
    ifs.edgeEnabled = 0
    ifs.rotationEnabled = 0
    for i in range(abi.MBX_IFS_VECTOR_COUNT):
        ifs.enabled[i] = 0
    ifs.enabled[0] = 1
    abi.set3(ifs.direction[0], (1.0, 0.0, 0.0))
    abi.set3(ifs.rotations[0], (90.0, 0.0, 0.0))
    ifs.distance[0] = 0.0
    ifs.intensity[0] = 1.0
    s, o = ifs.scale, np.array(list(ifs.offset))
    z4 = np.array([[0.4, 0.7, -0.2, 0.0], [0.4, -0.7, -0.2, 0.0]])
    aux = np.tile([0.0, 1.0, 0.0], (2, 1))
    z1, _ = oracle.formula_step(p, z4, aux)
    for k, v in enumerate(z4[:, :3]):
        w = QZ(v)                   # (-y, x, z)
        if w[0] < 0.0:
            w = w - np.array([1.0, 0.0, 0.0]) * (2.0 * w[0])
        np.testing.assert_allclose(z1[k, :3], s * w - o * (s - 1.0), rtol=0, atol=1e-12)


# ------------------------------------------------------------------------------------------
# light vector
# ------------------------------------------------------------------------------------------
Q = 90.4  # a quarter turn in the reference's light angle units (degrees / 180.8 * pi)


@pytest.mark.parametrize("angles,expect", [
    # relative to the camera at (0, -3.5, 0) looking at the origin, top z: forward F = (0, 1, 0),
    # right = F x top = (1, 0, 0); the direction starts at -F, turns about F by z, then about the
    # right vector by -y (directional lights flip y), then about top by x (Rodrigues)
    ((Q, 0.0, 0.0), (1.0, 0.0, 0.0)),
    ((-Q, 0.0, 0.0), (-1.0, 0.0, 0.0)),
    ((0.0, Q, 0.0), (0.0, 0.0, 1.0)),
    ((0.0, -Q, 0.0), (0.0, 0.0, -1.0)),
    ((0.0, 0.0, Q), (0.0, -1.0, 0.0)),   # a turn about F leaves -F
    ((0.0, Q, Q), (0.0, 0.0, 1.0)),      # z first: it does nothing to -F
    ((Q, Q, 0.0), (0.0, 0.0, 1.0)),      # y before x: (0, 0, 1) is on the top axis
])
def test_light_vector_relative_quarter_turns(angles, expect):
    p = scene.mandelbulb8(64, 48)
    p.light.relativePosition = 1
    abi.set3(p.light.rotation, angles)
    _, _, light = oracle.frame_constants(p)
    np.testing.assert_allclose(light, expect, rtol=0, atol=1e-14)


@pytest.mark.parametrize("angles,expect", [
    # absolute: R (0, -1, 0) with R = SetRotation(x, -y, z) = Rz(x) Rx(-y) Ry(z) (algebra.cpp:266-271)
    ((Q, 0.0, 0.0), (1.0, 0.0, 0.0)),
    ((0.0, Q, 0.0), (0.0, 0.0, 1.0)),
    ((0.0, 0.0, Q), (0.0, -1.0, 0.0)),
    ((Q, Q, 0.0), (0.0, 0.0, 1.0)),      # Rx first: (0, 0, 1), then Rz leaves it
    ((Q, 0.0, Q), (1.0, 0.0, 0.0)),      # Ry first leaves (0, -1, 0), then Rz
])
def test_light_vector_absolute_quarter_turns(angles, expect):
    p = scene.mandelbulb8(64, 48)
    p.light.relativePosition = 0
    abi.set3(p.light.rotation, angles)
    _, _, light = oracle.frame_constants(p)
    np.testing.assert_allclose(light, expect, rtol=0, atol=1e-14)


# ------------------------------------------------------------------------------------------
# Lambert light (shader_light_shading.cpp:40-122, shader_object.cpp:40-229)
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("shading", [0.0, 1.0, 0.4])
def test_lambert_light_closed_form(shading):
    """No fill light, AO, luminosity or highlights: a hit's colour is
    surface * light colour * intensity * (1 - shading + shading * max(n . l, 0)), with n the
    render's normal and l the light vector; shading 0 removes the normal altogether."""
    p = scene.mandelbulb8(96, 72)
    p.shading = shading
    p.light.intensity = 1.7
    abi.set3(p.light.color, (0.9, 0.8, 0.5))
    abi.set3(p.surfaceColor, (0.6, 0.3, 0.2))
    abi.set3(p.fillLightColor, (0.0, 0.0, 0.0))
    r = oracle.render(p)
    _, _, light = oracle.frame_constants(p)
    hit = r["mask"].ravel() > 0
    assert hit.sum() > 500
    n = r["normal"].reshape(-1, 3)[hit].astype(np.float64)
    lam = np.maximum(n @ np.asarray(light), 0.0)
    shade = (1.0 - shading + shading * lam) * 1.7
    expect = shade[:, None] * np.array([0.9 * 0.6, 0.8 * 0.3, 0.5 * 0.2])[None, :]
    np.testing.assert_allclose(r["rgb"].reshape(-1, 3)[hit], expect, rtol=2e-6, atol=2e-7)


# ------------------------------------------------------------------------------------------
# slow-shading normals (shader_calculate_normals.cpp:81-109)
# ------------------------------------------------------------------------------------------
def test_slow_shading_normal_follows_the_iteration_gradient():
    """A Mandelbox that never folds (folding limit 1e6, sphere-fold radii 1e-6) and adds no c
    maps z -> 2 z, so the iteration count is the radial step function ceil(log2(100 / |p|)): the
    slow-shading normal, sum q (1 + N - iterations(p + q delta)) over the 11^3 grid, points
    radially outwards (fewer iterations) at points on a step's sphere |p| = 100 / 2^7; without
    slow shading the same points give the DE's central-difference normal, also radial."""
    p = scene.mandelbox_shadow_ao(64, 36, seed=1)
    f = p.fractal
    f.deltaDEMethod = 2                      # analytic DE: r / DE, radial as well
    f.dontAddCConstant = 1
    f.mandelbox.foldingLimit = 1e6
    f.mandelbox.foldingSphericalMin = 1e-6
    f.mandelbox.foldingSphericalFixed = 1e-6
    abi.set3(p.camera, (0.0, -2.5, 0.0))
    rs = np.random.default_rng(9)
    dirs = rs.normal(size=(24, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    R = 100.0 / 2.0 ** 7
    for slow in (1, 0):
        p.slowShading = slow
        cos = [float(np.dot(oracle.normal(p, R * d, 0.01), d)) for d in dirs]
        assert min(cos) > 0.97, (slow, min(cos))
    # iterations step across the sphere (the grid straddles it): inside 7 + 1, outside 7
    _, it = oracle.distance(p, np.array([0.99 * R * dirs[0], 1.01 * R * dirs[0]]), 0.01)
    assert list(it) == [8, 7]


# ------------------------------------------------------------------------------------------
# adaptive Monte Carlo (render_worker.cpp:436-449, 1871-1901)
# ------------------------------------------------------------------------------------------
def _alternating_noise(a, b, n):
    """Noise after each of n samples alternating a, b (grey, 3 equal channels), from the running
    mean: sample 2j sits j (a - b) / (2j + 1) from the mean of the first 2j + 1 samples, sample
    2j + 1 half of (b - a) from the mean (a + b) / 2.  Standard error = sqrt(3 D / k) / sqrt(k)
    for the summed squared deviations D of one channel, divided by the mean brightness
    3 (sum / k) when that exceeds 1."""
    out, D, s = [], 0.0, 0.0
    for k in range(n):
        x = a if k % 2 == 0 else b
        j = k // 2
        dev = j * (a - b) / (2 * j + 1) if k % 2 == 0 else (b - a) / 2
        D += dev * dev
        s += x
        e = math.sqrt(3 * D / (k + 1)) / math.sqrt(k + 1)
        bright = 3 * s / (k + 1)
        out.append(e / bright if bright > 1.0 else e)
    return np.array(out)


def test_mc_noise_follows_the_running_mean_closed_form():
    a, b = 0.25, 0.75                         # dyadic: the float sums are exact
    x = np.array([[a] * 3 if k % 2 == 0 else [b] * 3 for k in range(40)], np.float32)
    np.testing.assert_allclose(oracle.mc_noise(x), _alternating_noise(a, b, 40), rtol=1e-6)
    # brighter than 1: the estimate is relative to the brightness, so scaling by 4 (exact in
    # float) leaves it unchanged once the mean brightness exceeds 1 (a wrong power of the
    # brightness, or the absolute error, would scale it)
    n1, n4 = oracle.mc_noise(x), oracle.mc_noise(4 * x)
    bright1 = np.cumsum(x.sum(1)) / np.arange(1, 41)
    assert (bright1[1:] > 1).all() and n1[0] == 0.0
    np.testing.assert_array_equal(n4, n1)
    np.testing.assert_allclose(n4, _alternating_noise(4 * a, 4 * b, 40), rtol=1e-6)
    # constant samples: no noise at all
    assert (oracle.mc_noise(np.full((12, 3), 0.5, np.float32)) == 0.0).all()


def test_adaptive_stop_takes_min_samples_plus_two():
    """With DOF_max_noise at its maximum (threshold 1.0) every pixel's estimate is below the
    threshold once the loop may stop, at repeat DOF_min_samples + 1: the image equals a fixed
    render of DOF_min_samples + 2 samples (same per-pixel streams), bit for bit.  With threshold 0
    (no estimate is below it; the strict comparison of render_worker.cpp:444) the image equals the
    full fixed render."""
    p = scene.mandelbox_shadow_ao(48, 32, seed=3)
    p.DOFMonteCarlo = 1
    p.DOFSamples = 24
    p.DOFMinSamples = 3
    p.DOFMaxNoise = 100.0
    adaptive = oracle.render(p)
    fixed = scene.mandelbox_shadow_ao(48, 32, seed=3)
    fixed.DOFMonteCarlo = 1
    fixed.DOFSamples = 5
    fixed.DOFMinSamples = 10                  # >= samples: never stops early
    ref5 = oracle.render(fixed)
    np.testing.assert_array_equal(adaptive["rgb"], ref5["rgb"])
    np.testing.assert_array_equal(adaptive["depth"], ref5["depth"])
    p.DOFMaxNoise = 0.0                       # even a pixel of identical samples does not stop
    tight = oracle.render(p)
    fixed.DOFSamples = 24
    fixed.DOFMinSamples = 30
    np.testing.assert_array_equal(tight["rgb"], oracle.render(fixed)["rgb"])


# ------------------------------------------------------------------------------------------
# iteration-threshold and interior modes (calculate_distance.cpp:561-600, render_worker.cpp:
# 653-679, 809-812, 913-933, shader_calculate_normals.cpp:46-47)
# ------------------------------------------------------------------------------------------
def _unit_ball(camera, target):
    """The power-2 Mandelbulb in Julia mode with c = 0 squares |z| at every iteration, so its
    filled set (the points that never bail out) is the closed unit ball."""
    p = scene.mandelbulb8(48, 32)
    p.fractal.bulb.power = 2.0
    p.fractal.juliaMode = 1
    abi.set3(p.fractal.juliaC, (0.0, 0.0, 0.0))
    abi.set3(p.camera, camera)
    abi.set3(p.target, target)
    return p


def _sphere_entry(c, d):
    d = d / np.linalg.norm(d)
    b = float(np.dot(c, d))
    disc = b * b - (float(np.dot(c, c)) - 1.0)
    return None if disc < 0 else -b - math.sqrt(disc)


def test_iteration_threshold_surface_is_the_unit_ball():
    """In iteration-threshold mode the surface is where orbits reach maxiter: the march stops
    on entering the unit ball, then backs off by one threshold, so depth + threshold is the
    ray's entry distance to the ball; the threshold has no detail-level divisor."""
    cam = np.array([0.0, -3.0, 0.0])
    p = _unit_ball(cam, (0.0, 0.0, 0.0))
    p.iterThreshMode = 1
    p.detailLevel = 3.0
    _, fov, _ = oracle.frame_constants(p)
    hits = 0
    for x, y in [(24, 16), (30, 20), (18, 12), (27, 9), (10, 5), (40, 28), (2, 2)]:
        m = oracle.march_pixel(p, x, y)
        t = _sphere_entry(cam, m["dir"])
        assert m["found"] == (t is not None), (x, y)
        if t is None:
            continue
        hits += 1
        dl = float(np.linalg.norm(m["dir"]))
        assert abs(m["depth"] * dl + m["thr"] - t) < 1e-6, (x, y, m, t)
        # CalcDistThresh in this mode: |camera - point| * resolution * fov, detail level unused,
        # taken at the point before the back-off (the surface, at distance t)
        want = (m["depth"] * dl + m["thr"]) * (1.0 / p.imageHeight) * fov
        assert m["thr"] == pytest.approx(want, rel=1e-12)
        n = oracle.normal(p, m["point"], m["thr"])
        assert float(np.dot(n, m["point"])) / float(np.linalg.norm(m["point"])) > 0.95
    assert hits >= 4


def test_interior_mode_keeps_interior_points_open_and_inverts_normals():
    """Interior mode: a point inside the ball (an orbit that reaches maxiter) measures the
    detail size instead of 0, and the normal-vector taps invert the DE near the surface
    (detailSize - d), so the normal at a hit on the ball points into it."""
    cam = np.array([0.0, -3.0, 0.0])
    p = _unit_ball(cam, (0.0, 0.0, 0.0))
    inside = np.array([[0.0, 0.2, 0.1], [0.5, -0.3, 0.2]])
    d0, _ = oracle.distance(p, inside, 0.01)
    assert (d0 == 0.0).all()
    p.interiorMode = 1
    d1, _ = oracle.distance(p, inside, 0.01)
    assert (d1 == 0.01).all()
    # an orbit cut short by maxiter counts as inside whatever its DE: with 2 iterations the point
    # at radius 1.5 reaches |z| = 1.5^4 < the bailout 10, and its positive DE becomes 1e-3
    p.fractal.formulaMaxiter = 2
    _, _, mx, _ = oracle.compute(p, np.array([[0.0, 1.5, 0.0]]))
    assert mx[0] == 1
    p.interiorMode = 0
    d2, _ = oracle.distance(p, np.array([[0.0, 1.5, 0.0]]), 1e-3)
    p.interiorMode = 1
    d3, _ = oracle.distance(p, np.array([[0.0, 1.5, 0.0]]), 1e-3)
    assert d2[0] > 1e-3 and d3[0] == 1e-3
    p.fractal.formulaMaxiter = 250
    for x, y in [(24, 16), (30, 20), (18, 12)]:
        m = oracle.march_pixel(p, x, y)
        t = _sphere_entry(cam, m["dir"])
        # the hit lies in the exterior band where the DE drops below the threshold, a few
        # thresholds before the ball (the Julia set's log DE underestimates |p| - 1 by ~2)
        assert m["found"] and 0.0 < t - m["depth"] * float(np.linalg.norm(m["dir"])) < 5 * m["thr"]
    # 2% outside the ball every tap's DE is well below 0.9 x the threshold 0.05, so all six are
    # inverted and the normal turns inwards; without interior mode it points outwards
    rs = np.random.default_rng(4)
    for _ in range(6):
        u = rs.normal(size=3)
        u /= np.linalg.norm(u)
        p.interiorMode = 1
        assert float(np.dot(oracle.normal(p, 1.02 * u, 0.05), u)) < -0.95
        p.interiorMode = 0
        assert float(np.dot(oracle.normal(p, 1.02 * u, 0.05), u)) > 0.95
    p.interiorMode = 1


# ------------------------------------------------------------------------------------------
# Monte Carlo lens, chromatic aberration and soft shadows (render_worker.cpp:226-264,
# 1832-1844, shader_aux_shadow.cpp:78-87, hsv2rgb.cpp:40-60)
# ------------------------------------------------------------------------------------------
def _lens_scene(dof=True, ca=False):
    p = scene.mandelbulb8(64, 48, seed=2)
    p.DOFMonteCarlo = 1
    p.DOFSamples = 4
    p.DOFEnabled = int(dof)
    p.DOFFocus = 2.5
    p.DOFRadius = 10.0
    p.DOFMonteCarloChromaticAberration = int(ca)
    p.DOFMonteCarloCACameraDispersion = 3.0
    return p


def test_lens_rays_cross_the_undisturbed_ray_at_the_focus():
    """Every lens ray starts on the lens disc (radius 0.0015 DOF_radius DOF_focus, in the plane
    through the camera perpendicular to its forward axis) and meets the ray of the same
    anti-aliasing draws without the lens at distance DOF_focus along the view vector; the lens
    takes exactly two more random draws; the radii fill the disc."""
    p, q = _lens_scene(), _lens_scene(dof=False)
    m, _, _ = oracle.frame_constants(p)
    forward = m[:, 1]
    cam = np.array([p.camera[i] for i in range(3)])
    rmax = 0.0015 * p.DOFRadius * p.DOFFocus
    radii = []
    for seed in range(1, 4000, 13):
        st, vv, tint, s1 = oracle.sample_ray(p, 17, 29, seed)
        st0, v0, _, s0 = oracle.sample_ray(q, 17, 29, seed)
        np.testing.assert_array_equal(st0, cam)
        np.testing.assert_allclose(st + p.DOFFocus * vv, cam + p.DOFFocus * v0, rtol=0, atol=1e-13)
        r = st - cam
        assert abs(np.dot(r, forward)) < 1e-15
        radii.append(np.linalg.norm(r))
        assert oracle.random_ints(s0, 2, 65536)[1] == s1
        np.testing.assert_array_equal(tint, [1.0, 1.0, 1.0])
    radii = np.array(radii) / rmax
    assert radii.max() <= 1.0 + 1e-12 and radii.max() > 0.98
    # sqrt of a uniform number: the radius is uniform over the disc's area, P(r < 1/2) = 1/4
    assert abs(np.mean(radii < 0.5) - 0.25) < 0.06


@pytest.mark.parametrize("hue, expect", [
    (0.0, (2, 0, 0)), (60.0, (2, 2, 0)), (120.0, (0, 2, 0)), (180.0, (0, 2, 2)), (240.0, (0, 0, 2)),
    (300.0, (2, 0, 2)), (30.0, (2, 1, 0)), (90.0, (1, 2, 0)), (210.0, (0, 1, 2)), (330.0, (2, 0, 1))])
def test_hsv2rgb_sectors(hue, expect):
    """Full saturation at value 2: the primaries and secondaries at multiples of 60 degrees, the
    halfway colours between them."""
    np.testing.assert_array_equal(oracle.hsv2rgb(hue, 1.0, 2.0), np.array(expect, np.float32))


def test_chromatic_aberration_tints_average_to_white_and_bend_radially():
    """The tints of the 3600 hues average to white (each channel is 2 over a third of the circle and
    ramps over two sixths); each sample's view vector is the undisturbed one pulled back along the
    camera's forward axis, so it stays in their common plane and turns away from that axis."""
    tints = np.array([oracle.hsv2rgb(np.float32(math.fmod(360.0 + h / 10.0 - 60.0, 360.0)), 1.0, 2.0)
                      for h in range(3600)], np.float64)
    np.testing.assert_allclose(tints.mean(axis=0), [1.0, 1.0, 1.0], atol=2e-3)
    p, q = _lens_scene(dof=False, ca=True), _lens_scene(dof=False)
    m, _, _ = oracle.frame_constants(p)
    forward = m[:, 1]
    seen = set()
    for seed in range(1, 3000, 7):
        _, vv, tint, _ = oracle.sample_ray(p, 5, 40, seed)
        _, v0, _, _ = oracle.sample_ray(q, 5, 40, seed)
        seen.add(tuple(np.round(tint, 3)))
        assert abs(np.linalg.norm(vv) - 1.0) < 1e-15
        assert abs(np.linalg.det(np.stack([vv, v0, forward]))) < 1e-15
        a = math.acos(min(1.0, np.dot(vv, forward)))
        a0 = math.acos(min(1.0, np.dot(v0, forward) / np.linalg.norm(v0)))
        assert a >= a0 - 1e-15
    assert len(seen) > 100


def test_soft_shadow_light_vectors_fill_the_cone():
    """Monte Carlo soft shadows: the unit light vector plus a point of the ball of radius R =
    tan(cone): every jittered vector is a unit vector within asin(R) of the light, the extreme is
    reached, and the jitter is symmetric about the light."""
    light = np.array([0.48, -0.6, 0.64])
    for cone_deg in (3.0, 12.0):
        R = math.tan(math.radians(cone_deg))
        out = oracle.soft_light(light, R, 11, 20000)
        np.testing.assert_allclose(np.linalg.norm(out, axis=1), 1.0, atol=1e-15)
        ang = np.arccos(np.clip(out @ light, -1.0, 1.0))
        assert ang.max() <= math.asin(R) + 1e-12
        assert ang.max() > 0.97 * math.asin(R)
        mean = out.mean(axis=0)
        np.testing.assert_allclose(mean / np.linalg.norm(mean), light, atol=3e-3 * R / 0.05)
    np.testing.assert_allclose(oracle.soft_light(light, 0.0, 5, 3), np.tile(light, (3, 1)), atol=1e-15)

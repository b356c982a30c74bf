"""GPU renders through the C ABI (libmbx.so) vs the FP64 oracle on identical scenes.

Bars (the oracle restates the reference CPU renderer in FP64):
  * hit/miss mask and per-pixel iteration count: equal on >= 99.9% of pixels (integers,
    compared exactly);
  * depth: FP32 marches within 1e-4 relative on >= 99.9% of the pixels both renders hit (the
    headline bar, asserted on the 1920x1080 Mandelbulb-8 frame itself); FP64 marches (Mandelbox,
    deep zoom) within 1e-6; a comparison without common hits fails;
  * RGB: within 1e-2 absolute on >= 99.9% of pixels.
"""
import numpy as np
import pytest

from mandelbulber2_amd import abi, scene
from tests.parity_util import compare

pytestmark = pytest.mark.gpu

GRID, PLAIN, PERSISTENT = 0, 1, 2


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


def _check(st, depth_tol=1e-4, iters=0.999):
    assert st["mask_agree"] >= 0.999, st
    assert st["iters_agree"] >= iters, st
    # ">= 99.9% of common hits within depth_tol": the 0.999 quantile of the relative error
    assert "depth_rel_p999" in st, ("no pixel hit in both renders", st)
    assert st["depth_rel_p999"] <= depth_tol, st
    assert st["rgb_within_1e-2"] >= 0.999, st


def test_scene_info_matches_oracle(rt, oracle_mod):
    for p in (scene.mandelbulb8(320, 240), scene.mandelbox_shadow_ao(320, 180, seed=3)):
        ctx = rt.Context(0)
        ctx.upload(p)
        m, fov, light, kind = ctx.scene_info()
        ctx.close()
        m_o, fov_o, light_o = oracle_mod.frame_constants(p)
        # the host derives the camera frame and light vector in the reference's operation order
        np.testing.assert_allclose(m, m_o, rtol=0, atol=1e-15)
        assert abs(fov - fov_o) <= 1e-15
        np.testing.assert_allclose(light, light_o, rtol=0, atol=1e-15)


def test_distance_matches_oracle(rt, oracle_mod):
    rs = np.random.default_rng(7)
    pts = rs.uniform(-1.3, 1.3, size=(4096, 3))
    p = scene.mandelbulb8(640, 480)
    ctx = rt.Context(0)
    ctx.upload(p)
    d, it = ctx.distance(pts, 1e-3)
    ctx.close()
    d_o, it_o = oracle_mod.distance(p, pts, 1e-3)
    outside = d_o > 1e-3
    assert (it == it_o).mean() >= 0.995
    rel = np.abs(d[outside] - d_o[outside]) / d_o[outside]
    assert np.quantile(rel, 0.99) < 1e-4, np.quantile(rel, 0.99)


@pytest.mark.parametrize("kernel", [GRID, PLAIN, PERSISTENT])
def test_mandelbulb8_small_parity(rt, oracle_mod, kernel):
    p = scene.mandelbulb8(160, 120)
    g = rt.render_to_host(p, kernel=kernel)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("mandelbulb8 160x120 kernel %d" % kernel, st)
    _check(st)


def test_mandelbulb8_640x480_parity(rt, oracle_mod):
    p = scene.mandelbulb8(640, 480)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("mandelbulb8 640x480", st)
    _check(st)
    # the work counters agree to within what the FP32 march changes; far-field steps (when
    # enabled) are the reference's DE evaluations the device proves unnecessary, one iteration each
    dev_evals = g["de_evals"] + g["far_steps"]
    assert abs(dev_evals - r["de_evals"]) / r["de_evals"] < 0.001
    assert abs(g["iterations"] + g["far_steps"] - r["iterations"]) / r["iterations"] < 0.001


def test_mandelbulb8_1080p_parity(rt, oracle_mod):
    """The benchmark's own frame: 1920x1080 Mandelbulb-8 from the default launch.  Mask and
    iteration counts equal on >= 99.9% of pixels, depth within 1e-4 relative on >= 99.9% of the
    common hits, RGB within 1e-2."""
    p = scene.mandelbulb8(1920, 1080)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("mandelbulb8 1920x1080", st)
    _check(st, depth_tol=1e-4)
    assert st["depth_within_1e-4"] >= 0.999, st


@pytest.mark.parametrize("scale", [2.0, -1.5])
def test_mandelbox_shadow_ao_parity(rt, oracle_mod, scale):
    """Mandelbox, delta DE (the reference's 4-orbit finite difference in FP64), hard shadow
    from the penetrating directional light and 16-ray AO."""
    p = scene.mandelbox_shadow_ao(960, 540, seed=1, scale=scale)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("mandelbox %g shadow+AO 960x540" % scale, st)
    _check(st, depth_tol=1e-6)
    assert g["de_evals"] == pytest.approx(r["de_evals"], rel=1e-4)


def test_menger_ifs_parity(rt, oracle_mod):
    p = scene.menger_ifs(160, 120)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("menger 160x120", st)
    _check(st)


@pytest.mark.parametrize("which", ["mandelbulb8", "mandelbox"])
def test_rotated_node_parity(rt, oracle_mod, which):
    """A rotated fractal node (world -> local R p before every DE evaluation)."""
    if which == "mandelbulb8":
        p = scene.mandelbulb8(128, 96)
    else:
        p = scene.mandelbox_shadow_ao(128, 72, seed=1)
    abi.set3(p.formulaRotation, (17.0, -33.0, 71.0))
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("rotated %s" % which, st)
    _check(st, depth_tol=1e-4 if which == "mandelbulb8" else 1e-6)


def test_mandelbulb_trig_power_parity(rt, oracle_mod):
    p = scene.mandelbulb8(128, 96)
    p.fractal.bulb.power = 5.5
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("mandelbulb 5.5 128x96", st)
    assert st["mask_agree"] >= 0.999
    assert st["iters_agree"] >= 0.995  # general power: asin/atan2/pow in FP32
    assert st["rgb_within_1e-2"] >= 0.999


def test_monte_carlo_gi_parity(rt, oracle_mod):
    """4 Monte Carlo samples per pixel with 2 GI bounces, shadows on: the per-pixel random
    stream is shared by the oracle and the kernel."""
    p = scene.mandelbulb8(96, 72)
    p.DOFMonteCarlo = 1
    p.DOFSamples = 4
    p.DOFMonteCarloGlobalIllumination = 1
    p.reflectionsMax = 2
    p.light.castShadows = 1
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("mandelbulb8 MC GI 96x72", st)
    _check(st)


def test_adaptive_monte_carlo_parity(rt, oracle_mod):
    """Adaptive Monte Carlo (render_worker.cpp:436-449): at most 24 samples, the stop allowed
    from the fifth; at a noise threshold of 0.02 the pixels of this scene stop anywhere from 5 to
    24 samples.  Against the oracle, and on the device alone: with the largest threshold every
    pixel stops at DOF_min_samples + 2, so the image equals a fixed 5-sample render bit for bit."""
    p = scene.mandelbox_shadow_ao(96, 64, seed=3)
    p.DOFMonteCarlo = 1
    p.DOFSamples = 24
    p.DOFMinSamples = 3
    p.DOFMaxNoise = 2.0
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("mandelbox adaptive MC 96x64", st)
    _check(st, depth_tol=1e-6)
    p.DOFMaxNoise = 100.0
    fixed = scene.mandelbox_shadow_ao(96, 64, seed=3)
    fixed.DOFMonteCarlo = 1
    fixed.DOFSamples = 5
    fixed.DOFMinSamples = 10
    a, b = rt.render_to_host(p), rt.render_to_host(fixed)
    np.testing.assert_array_equal(a["rgb"], b["rgb"])
    np.testing.assert_array_equal(a["depth"], b["depth"])


def _lens(p, dof=True, ca=True, soft=True):
    """Monte Carlo with the lens focused at the target (DOF_radius 10: a lens of 0.0015 x 10 x
    focus), chromatic aberration, and soft shadows through a 5 degree cone."""
    cam = np.array([p.camera[i] for i in range(3)])
    tgt = np.array([p.target[i] for i in range(3)])
    p.DOFMonteCarlo = 1
    p.DOFSamples = 8
    p.DOFEnabled = int(dof)
    p.DOFFocus = float(np.linalg.norm(tgt - cam)) * 0.9
    p.DOFRadius = 10.0
    p.DOFMonteCarloChromaticAberration = int(ca)
    p.DOFMonteCarloCACameraDispersion = 2.0
    p.monteCarloSoftShadows = int(soft)
    p.light.castShadows = 1
    p.light.softShadowCone = 5.0
    return p


@pytest.mark.parametrize("which", ["mandelbulb8", "mandelbulb_trig", "mandelbulb_fp64", "mandelbox", "generic"])
def test_monte_carlo_lens_parity(rt, oracle_mod, which):
    """Lens rays start off the camera (the march takes its thresholds from the distance to the
    camera point by point), chromatic aberration tints and bends each sample, and the soft-shadow
    light vector is jittered per sample: 8 samples per pixel against the oracle on the FP32
    Mandelbulb kinds, the FP64 Mandelbulb and Mandelbox, and the general kernel (Xenodreambuie)."""
    if which == "mandelbox":
        p = scene.mandelbox_shadow_ao(96, 64, seed=1)   # surface about 3 units from the lens
    else:
        p = scene.mandelbulb8(96, 72)
        if which == "mandelbulb_trig":
            p.fractal.bulb.power = 5.5
        if which == "generic":
            p.fractal.formula = abi.MBX_FORMULA_XENODREAMBUIE
    _lens(p)
    tune = {rt.MBX_TUNE_PRECISION: 2} if which == "mandelbulb_fp64" else None
    g = rt.render_to_host(p, tune=tune)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("%s lens + CA + MC soft shadows" % which, st)
    assert g["de_kind"] in {"mandelbulb8": (0,), "mandelbulb_trig": (1,), "mandelbulb_fp64": (5,),
                            "mandelbox": (2, 3), "generic": (6,)}[which]
    if which == "mandelbulb_trig":
        assert st["mask_agree"] >= 0.999 and st["iters_agree"] >= 0.995, st
        assert st["rgb_within_1e-2"] >= 0.999, st
    else:
        _check(st, depth_tol=1e-4 if which == "mandelbulb8" else 1e-6)


def test_lens_angle_draw_at_its_maximum(rt, oracle_mod):
    """Random(65536) draws 0 .. 65536 inclusive: a pixel whose lens angle draw is 65536 (frame seed
    and pixel found from the oracle's random stream) renders as the oracle's, FP64 march, 1 sample."""
    p = scene.mandelbulb8(64, 48)
    hits = np.flatnonzero(oracle_mod.render(p)["mask"].ravel())
    found = None
    for fs in range(1, 400):
        for pix in hits:
            _, s1 = oracle_mod.random_ints(oracle_mod.pixel_seed(int(pix), fs), 2, 1000)
            if oracle_mod.random_ints(s1, 2, 65536)[0][1] == 65536:
                found = (fs, int(pix))
                break
        if found:
            break
    assert found
    p.randomSeed = found[0]
    _lens(p, ca=False, soft=False)
    p.DOFSamples = 1
    g = rt.render_to_host(p, tune={rt.MBX_TUNE_PRECISION: 2})
    r = oracle_mod.render(p)
    y, x = divmod(found[1], p.imageWidth)
    assert r["mask"][y, x] and g["mask"][y, x]
    assert g["depth"][y, x] == r["depth"][y, x]
    np.testing.assert_array_equal(g["depth"], r["depth"])


def test_lens_frames_in_a_batch_equal_single_frames(rt):
    """A batch holding a lens frame runs every frame through the lens variant of the kernel; each
    frame equals its single-frame render (plain frames: the lite or full variant) bit for bit."""
    import torch
    frames = [_lens(scene.mandelbulb8(64, 48, seed=1)), scene.mandelbulb8(80, 40, seed=2),
              _lens(scene.mandelbulb8(48, 48, seed=3), dof=False, soft=False)]
    for kernel in (GRID, PERSISTENT):
        ctx = rt.Context(0)
        ctx.tune(rt.MBX_TUNE_KERNEL, kernel)
        bufs = [ctx.alloc_outputs(p) for p in frames]
        ctx.render_batch(frames, bufs)
        torch.cuda.synchronize()
        for p, b in zip(frames, bufs):
            single = rt.render_to_host(p)
            for key in ("rgb", "depth", "iters", "mask", "steps"):
                np.testing.assert_array_equal(b[key].cpu().numpy(), single[key], err_msg=key)
        ctx.close()


@pytest.mark.parametrize("powers", [(6.0, 7.0, 9.5), (8.0, 8.0, 8.0)])
@pytest.mark.parametrize("kernel", [GRID, PLAIN, PERSISTENT])
def test_batch_equals_single_frames(rt, powers, kernel):
    """Frames of different resolution, camera and power rendered in one batch (one grid launch,
    one plain launch per frame, or one persistent launch) equal their single-frame renders bit for
    bit.  (A batch mixing power 8 with other powers runs every frame through the general-power
    kernel, so the two cases are kept apart here.)"""
    import torch
    frames = []
    for k, (w, h) in enumerate([(96, 72), (64, 48), (120, 68)]):
        p = scene.mandelbulb8(w, h, seed=k)
        p.fractal.bulb.power = powers[k]
        abi.set3(p.camera, (0.3 * k, -3.0 - 0.2 * k, 0.5))
        frames.append(p)
    ctx = rt.Context(0)
    ctx.tune(rt.MBX_TUNE_KERNEL, kernel)
    bufs = [ctx.alloc_outputs(p) for p in frames]
    ctx.render_batch(frames, bufs)
    st = ctx.stats()
    torch.cuda.synchronize()
    assert st["rays"] == sum(p.imageWidth * p.imageHeight for p in frames)
    for p, b in zip(frames, bufs):
        single = rt.render_to_host(p)
        for key in ("rgb", "depth", "iters", "mask", "steps"):
            np.testing.assert_array_equal(b[key].cpu().numpy(), single[key], err_msg=key)
    ctx.close()


def test_animation_batch_parity(rt, oracle_mod):
    """The batched-animation workload (power sweep 2 -> 12, random-walk camera, mixed
    resolutions, 4 spp Monte Carlo GI of depth 2): all 64 frames at a quarter of their
    resolution (480x270 / 320x180 / 160x120) in one launch, each frame against the oracle."""
    import torch
    frames = scene.animation64(seed=3, frames=64, downscale=4)
    ctx = rt.Context(0)
    bufs = [ctx.alloc_outputs(p) for p in frames]
    ctx.render_batch(frames, bufs)
    torch.cuda.synchronize()
    worst = {"mask_agree": 1.0, "iters_agree": 1.0, "rgb_within_1e-2": 1.0, "depth_rel_p999": 0.0}
    for k, (p, b) in enumerate(zip(frames, bufs)):
        g = {key: t.cpu().numpy() for key, t in b.items() if not key.startswith("_")}
        r = oracle_mod.render(p)
        st = compare(g, r)
        if k % 8 == 0:
            _report("animation frame %d power %.2f %dx%d" % (k, p.fractal.bulb.power, p.imageWidth,
                                                             p.imageHeight), st)
        assert st["mask_agree"] >= 0.999, (k, st)
        assert st["iters_agree"] >= 0.995, (k, st)  # general power: asin/atan2/pow in FP32
        assert st["rgb_within_1e-2"] >= 0.995, (k, st)
        assert st["depth_rel_p999"] <= 1e-4, (k, st)
        for key in ("mask_agree", "iters_agree", "rgb_within_1e-2"):
            worst[key] = min(worst[key], st[key])
        worst["depth_rel_p999"] = max(worst["depth_rel_p999"], st["depth_rel_p999"])
    _report("animation 64 frames, worst frame", worst)
    ctx.close()


def test_mixed_kind_batch_parity(rt, oracle_mod):
    """A Mandelbox batch mixing rotated folds (general kernel) with plain frames (delta-DE
    kernel) renders every frame on the general kernel; each frame against the oracle."""
    import torch
    frames = []
    for k in range(3):
        p = scene.mandelbox_shadow_ao(96, 54, seed=k + 1)
        if k == 1:
            p.fractal.mandelbox.rotationsEnabled = 1
            p.fractal.mandelbox.rotation[0][0][0] = 11.0
            p.fractal.mandelbox.rotation[1][2][1] = -23.0
        frames.append(p)
    ctx = rt.Context(0)
    out = [ctx.alloc_outputs(p) for p in frames]
    ctx.render_batch(frames, out)
    torch.cuda.synchronize()
    for k, (p, b) in enumerate(zip(frames, out)):
        g = {key: t.cpu().numpy() for key, t in b.items() if not key.startswith("_")}
        st = compare(g, oracle_mod.render(p))
        _report("mixed mandelbox batch frame %d" % k, st)
        _check(st, depth_tol=1e-6)
    ctx.close()


def _histogram(values, edges):
    """Counts of values in [0, e0], (e0, e1], ..., (e_last, inf)."""
    counts = []
    lo = -1.0
    for e in edges:
        counts.append(int(((values > lo) & (values <= e)).sum()))
        lo = e
    counts.append(int((values > lo).sum()))
    return counts


def test_deep_zoom_mandelbox_parity(rt, oracle_mod):
    """Deep zoom: camera 1e-9 from a Mandelbox surface point found by a seeded FP64 surface
    search, constant DE threshold 1e-11, at the configuration's 1024x1024.  Ray positions, scan and orbits run in FP64 on
    the device, so against the all-FP64 oracle the depth (about 1e-9) agrees to 1e-6 relative;
    the per-pixel depth and RGB error histograms are printed."""
    p = scene.deep_zoom_mandelbox(1024, 1024, seed=5)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("deep zoom mandelbox 1024x1024", st)
    both = (g["mask"] > 0) & (r["mask"] > 0)
    rel = np.abs(g["depth"][both].astype(np.float64) - r["depth"][both]) / r["depth"][both]
    rgb = np.abs(g["rgb"].astype(np.float64) - r["rgb"]).max(axis=-1).ravel()
    edges = [0.0, 1e-9, 1e-7, 1e-5, 1e-3]
    print("[deep zoom depth rel err histogram] edges %s counts %s" % (edges, _histogram(rel, edges)))
    print("[deep zoom rgb abs err histogram] edges %s counts %s" % (edges, _histogram(rgb, edges)))
    print("[deep zoom] depth median %.4g, hits %d of %d" % (np.median(r["depth"][r["mask"] > 0]),
                                                           int(both.sum()), both.size), flush=True)
    _check(st, depth_tol=1e-6)


def test_graph_replays_match_direct_launch(rt):
    """The captured-graph launch, replayed several times with other device work in between,
    gives the direct launch's result every time."""
    import torch
    p = scene.mandelbulb8(128, 96)
    ctx = rt.Context(0)
    ctx.upload(p)
    ctx.tune(rt.MBX_TUNE_GRAPH, 0)
    ref = ctx.alloc_outputs()
    ctx.render(ref)
    ref_st = ctx.stats()
    ctx.tune(rt.MBX_TUNE_GRAPH, 1)
    out = ctx.alloc_outputs()
    for k in range(4):
        for t in out.values():
            if hasattr(t, "fill_"):
                t.fill_(7)
        junk = torch.randn(1 << 16, device="cuda").sum()  # eager work between replays
        ctx.render(out)
        st = ctx.stats()
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        for key in ("rgb", "depth", "iters", "mask", "normal", "steps"):
            assert torch.equal(out[key], ref[key]), (k, key)
        assert st["de_evals"] == ref_st["de_evals"] and st["iterations"] == ref_st["iterations"]
    ctx.close()


@pytest.mark.parametrize("case", ["bulb8", "power2", "power12", "mc_gi"])
def test_far_field_steps_leave_pixels_unchanged(rt, case):
    """March steps beyond the far-field radius taken without a DE evaluation give bit-identical
    planes to evaluating every step (the step is provably the clamped 3 there)."""
    import torch
    p = scene.mandelbulb8(160, 120)
    if case == "power2":
        p.fractal.bulb.power = 2.0
    elif case == "power12":
        p.fractal.bulb.power = 12.0
    elif case == "mc_gi":
        p = scene.mandelbulb8(96, 72)
        p.DOFMonteCarlo = 1
        p.DOFSamples = 4
        p.DOFMonteCarloGlobalIllumination = 1
        p.reflectionsMax = 2
    ctx = rt.Context(0)
    outs, stats = [], []
    for on in (0, 1):
        ctx.tune(rt.MBX_TUNE_FAR_FIELD, on)
        ctx.upload(p)
        b = ctx.alloc_outputs()
        ctx.render(b)
        stats.append(ctx.stats())
        outs.append(b)
    torch.cuda.synchronize()
    for key in ("rgb", "depth", "iters", "mask", "normal", "steps"):
        assert torch.equal(outs[0][key], outs[1][key]), key
    assert stats[0]["far_steps"] == 0 and stats[1]["far_steps"] > 0
    assert stats[1]["de_evals"] + stats[1]["far_steps"] == stats[0]["de_evals"]
    print("\n[far field %s] %d of %d DE evaluations skipped" % (
        case, stats[1]["far_steps"], stats[0]["de_evals"]), flush=True)
    ctx.close()


def _unit_ball(w, h, camera):
    """The power-2 Mandelbulb in Julia mode with c = 0: its filled set is the unit ball."""
    p = scene.mandelbulb8(w, h)
    p.fractal.bulb.power = 2.0
    p.fractal.juliaMode = 1
    abi.set3(p.fractal.juliaC, (0.0, 0.0, 0.0))
    abi.set3(p.camera, camera)
    abi.set3(p.target, (0.0, 0.0, 0.0))
    return p


@pytest.mark.parametrize("mode", ["iteration_threshold", "interior"])
def test_interior_and_iteration_threshold_parity(rt, oracle_mod, mode):
    """Iteration-threshold mode (the march stops where orbits reach maxiter, then backs off one
    threshold; the threshold drops the detail level) and interior mode (maxiter points measure
    the detail size, the march keeps 0.8 thresholds, normal taps invert the DE, shadows start one
    DE-factor threshold out) on the general kernel, with the hard shadow and fast AO on: the unit
    ball of the z^2 Julia set, then the Mandelbox of the shadow scene.

    Mask, iterations and depth: the FP64 bars.  Colour: in these modes the DE jumps between 0
    (or the detail size) and 1.01 thresholds at the maxiter set, and a shadow ray crawls along
    the surface in steps of 1.01 thresholds, so whether one of its samples lands inside the set
    changes under perturbations of the ray at the last bit.  The oracle shows it on itself: turning
    the light by 1e-13 degrees changes a few percent of its pixels (none in the ordinary mode).
    The kernel may differ from the oracle on at most twice the oracle's own fraction."""
    for p, name in ((_unit_ball(160, 120, (0.4, -2.6, 0.9)), "unit ball"),
                    (scene.mandelbox_shadow_ao(160, 120, seed=2), "mandelbox")):
        p.iterThreshMode = int(mode == "iteration_threshold")
        p.interiorMode = int(mode == "interior")
        p.detailLevel = 1.5
        p.light.castShadows = 1
        p.ambientOcclusionEnabled = 1
        p.ambientOcclusionMode = abi.MBX_AO_FAST
        g = rt.render_to_host(p)
        assert g["de_kind"] == 6
        r = oracle_mod.render(p)
        st = compare(g, r)
        p.light.rotation[0] += 1e-13
        r2 = oracle_mod.render(p)
        chaos = float((np.abs(r["rgb"] - r2["rgb"]).reshape(-1, 3).max(1) > 1e-2).mean())
        _report("%s %s 160x120 (oracle self-sensitivity %.4f)" % (name, mode, chaos), st)
        assert st["mask_agree"] >= 0.999 and st["iters_agree"] >= 0.995, st
        assert "depth_rel_p999" in st and st["depth_rel_p999"] <= 1e-6, st
        assert st["rgb_within_1e-2"] >= 1.0 - max(2.0 * chaos, 0.001), (st, chaos)


def test_unsupported_settings_are_rejected(rt):
    ctx = rt.Context(0)
    p = scene.mandelbulb8(64, 48)
    p.perspectiveType = 4
    with pytest.raises(rt.MbxError, match="perspective"):
        ctx.upload(p)
    p = scene.mandelbulb8(64, 48)
    p.perspectiveType = 1
    p.DOFMonteCarlo = 1
    p.DOFEnabled = 1
    with pytest.raises(rt.MbxError, match="fish-eye"):
        ctx.upload(p)
    p = scene.mandelbulb8(64, 48)
    p.antialiasingEnabled = 1
    p.antialiasingSize = 11
    with pytest.raises(rt.MbxError, match="antialiasing"):
        ctx.upload(p)
    p = scene.mandelbulb8(64, 48)
    p.antialiasingEnabled = 1
    p.basicFogEnabled = 1
    with pytest.raises(rt.MbxError, match="antialiasing"):
        ctx.upload(p)
    p = scene.mandelbulb8(64, 48)
    p.rngMode = abi.MBX_RNG_GLOBAL_SEQUENTIAL
    with pytest.raises(rt.MbxError):
        ctx.upload(p)
    ctx.close()


@pytest.mark.parametrize("kernel", [GRID, PLAIN, PERSISTENT])
def test_tile_shapes_give_identical_pixels(rt, kernel):
    """Tiles of 4x16, 16x4 and 32x2 pixels (one wave each) against the default 8x8 on a frame
    whose size is a multiple of none of them: every plane is bit-identical."""
    import torch
    p = scene.mandelbulb8(150, 101)
    ctx = rt.Context(0)
    ctx.tune(rt.MBX_TUNE_KERNEL, kernel)
    ctx.upload(p)
    ref = ctx.alloc_outputs()
    ctx.render(ref)
    ref_st = ctx.stats()
    for width in (4, 16, 32):
        ctx.tune(rt.MBX_TUNE_TILE_WIDTH, width)
        assert ctx.get_tune(rt.MBX_TUNE_TILE_WIDTH) == width
        out = ctx.alloc_outputs()
        ctx.render(out)
        st = ctx.stats()
        torch.cuda.synchronize()
        for key in ("rgb", "depth", "iters", "mask", "normal", "steps"):
            assert torch.equal(out[key], ref[key]), (width, key)
        assert st["de_evals"] == ref_st["de_evals"]
    with pytest.raises(rt.MbxError, match="tile width"):
        ctx.tune(rt.MBX_TUNE_TILE_WIDTH, 12)
    ctx.close()


def test_autotune_picks_a_launch_shape(rt):
    p = scene.mandelbulb8(256, 192)
    ctx = rt.Context(0)
    ctx.upload(p)
    bufs = ctx.alloc_outputs()
    waves, ms = ctx.autotune(bufs, reps=1)
    assert waves > 0 and ms > 0
    assert ctx.get_tune(rt.MBX_TUNE_WAVES) == waves
    assert ctx.get_tune(rt.MBX_TUNE_TILE_WIDTH) in (4, 8, 16, 32)
    ctx.render(bufs)
    assert ctx.stats()["rays"] == 256 * 192
    ctx.close()


def test_counters_report_one_launch(rt):
    """Each launch reports only its own work.  Direct grid launches accumulate into one of two
    counter banks and zero the other for the next launch; plain, graph and persistent launches
    clear their bank with a clear kernel.  Any order of modes, and several renders before one
    statistics read, give the single-launch counts."""
    p = scene.mandelbulb8(160, 120)
    ctx = rt.Context(0)
    ctx.upload(p)
    out = ctx.alloc_outputs()
    ref = None
    seq = [(GRID, 0), (GRID, 0), (GRID, 0), (PLAIN, 0), (GRID, 0), (GRID, 1), (GRID, 1), (GRID, 0),
           (GRID, 0), (PERSISTENT, 0), (GRID, 0), (PERSISTENT, 1), (GRID, 0)]
    for mode, graph in seq:
        ctx.tune(rt.MBX_TUNE_KERNEL, mode)
        ctx.tune(rt.MBX_TUNE_GRAPH, graph)
        for _ in range(3):
            ctx.render(out)
        st = ctx.stats()
        key = (st["de_evals"], st["iterations"], st["far_steps"])
        if ref is None:
            ref = key
            assert ref[0] > 0
        assert key == ref, (mode, graph, key, ref)
    ctx.close()


def test_untimed_renders_keep_counters(rt):
    """MBX_TUNE_TIMING 0 drops the context's event pair: the counters are unchanged and
    kernel_ms reads 0; switching timing back on reports a positive device time again."""
    p = scene.mandelbulb8(160, 120)
    ctx = rt.Context(0)
    ctx.upload(p)
    out = ctx.alloc_outputs()
    ctx.render(out)
    ref = ctx.stats()
    assert ref["kernel_ms"] > 0
    ctx.tune(rt.MBX_TUNE_TIMING, 0)
    for _ in range(3):
        ctx.render(out)
    st = ctx.stats()
    assert st["kernel_ms"] == 0
    assert (st["de_evals"], st["iterations"]) == (ref["de_evals"], ref["iterations"])
    ctx.tune(rt.MBX_TUNE_TIMING, 1)
    ctx.render(out)
    assert ctx.stats()["kernel_ms"] > 0
    ctx.close()

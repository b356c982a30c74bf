"""The general formula path (DE kind 6: every formula of the table, alone or in a hybrid sequence,
FP64 with a 4D orbit) on the GPU vs the FP64 oracle.

Bars as tests/test_gpu_parity.py: mask and iteration counts equal on >= 99.9% of pixels, RGB
within 1e-2 on >= 99.9%, depth within 1e-6 (99.9th percentile) where the formulas are plain
arithmetic; formulas and hybrids with trigonometric steps (Mandelbulb, Xenodreambuie) call the
device's FP64 libm-class functions, whose last-bit differences from the host's can move a hit by
a binary-search step: depth within 1e-3 on >= 99.9% and iterations >= 99.5% there.
"""
import numpy as np
import pytest

from mandelbulber2_amd import abi, scene
from tests.parity_util import compare

pytestmark = pytest.mark.gpu

GENERIC = 6


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


def _check(st, trig):
    assert st["mask_agree"] >= 0.999, st
    assert st["iters_agree"] >= (0.995 if trig else 0.999), st
    if trig:
        assert st["depth_within_1e-3"] >= 0.999, st
    else:
        assert st["depth_rel_p999"] <= 1e-6, st
    assert st["rgb_within_1e-2"] >= 0.999, st


def _base(formula, w=96, h=72, camera=(1.4, -2.8, 1.0)):
    p = scene.plain_material(scene.default_params(w, h))
    p.fractal.formula = formula
    p.fractal.formulaMaxiter = 60
    p.light.castShadows = 0
    abi.set3(p.camera, camera)
    return p


def _hybrid(slots, w=96, h=72, camera=(1.4, -2.8, 1.0)):
    """slots: (formula, iterations, weight, extra setter or None)"""
    p = _base(slots[0][0], w, h, camera)
    p.hybridCount = len(slots)
    conts = [p.fractal] + [p.hybridFractals[k] for k in range(len(slots) - 1)]
    for c, (formula, iters, weight, setter) in zip(conts, slots):
        c.formula = formula
        c.formulaIterations = iters
        c.formulaWeight = weight
        if setter:
            setter(c)
    return p


def _power(v):
    def set_power(c):
        c.bulb.power = v
    return set_power


def _menger_ifs(c):
    ifs = c.IFS
    ifs.scale = 3.0
    abi.set3(ifs.offset, (1.0, 1.0, 1.0))
    ifs.absX = ifs.absY = ifs.absZ = 1
    ifs.mengerSpongeMode = 1


SCENES = {
    "mandelbulb_power2": (lambda: _base(abi.MBX_FORMULA_MANDELBULB_POWER2), False),
    "hypercomplex": (lambda: _base(abi.MBX_FORMULA_HYPERCOMPLEX), False),
    "quaternion": (lambda: _base(abi.MBX_FORMULA_QUATERNION), False),
    "menger_sponge": (lambda: _base(abi.MBX_FORMULA_MENGER_SPONGE), False),
    "xenodreambuie": (lambda: _power_scene(abi.MBX_FORMULA_XENODREAMBUIE, 8.0), True),
    "box_fold_bulb_pow2": (lambda: _base(abi.MBX_FORMULA_BOX_FOLD_BULB_POW2, camera=(9.0, -18.0, 6.0)), False),
    "aexion": (lambda: _base(abi.MBX_FORMULA_AEXION), False),
    "hypercomplex_w": (lambda: _initial_w(), False),
    "hybrid_bulb_box": (lambda: _hybrid([(abi.MBX_FORMULA_MANDELBULB, 1, 1.0, _power(8.0)),
                                          (abi.MBX_FORMULA_MANDELBOX, 1, 1.0, None)],
                                         camera=(2.0, -4.0, 1.5)), True),
    "hybrid_menger_power2": (lambda: _hybrid([(abi.MBX_FORMULA_MENGER_SPONGE, 2, 1.0, None),
                                               (abi.MBX_FORMULA_MANDELBULB_POWER2, 1, 1.0, None)]), False),
    "hybrid_weighted": (lambda: _hybrid([(abi.MBX_FORMULA_MANDELBULB_POWER2, 1, 1.0, None),
                                          (abi.MBX_FORMULA_QUATERNION, 1, 0.6, None)]), False),
    "hybrid_ifs_box_julia": (lambda: _julia(_hybrid([(abi.MBX_FORMULA_KALEIDOSCOPIC_IFS, 1, 1.0, _menger_ifs),
                                                      (abi.MBX_FORMULA_MANDELBOX, 2, 1.0, None)],
                                                     camera=(3.0, -6.0, 2.0))), False),
}


def _power_scene(formula, power):
    p = _base(formula)
    p.fractal.bulb.power = power
    return p


def _initial_w():
    p = _base(abi.MBX_FORMULA_HYPERCOMPLEX)
    p.fractal.initialWAxis = 0.3
    return p


def _julia(p):
    p.fractal.juliaMode = 1
    abi.set3(p.fractal.juliaC, (0.2, -0.3, 0.1))
    return p


@pytest.mark.parametrize("name", sorted(SCENES))
def test_generic_parity(rt, oracle_mod, name):
    make, trig = SCENES[name]
    p = make()
    print("[case] %s" % name, flush=True)
    g = rt.render_to_host(p)
    assert g["de_kind"] == GENERIC
    r = oracle_mod.render(p)
    assert r["mask"].mean() > 0.05, "the scene shows the fractal"
    st = compare(g, r)
    _report(name, st)
    _check(st, trig)


def test_generic_palette_and_shading_parity(rt, oracle_mod):
    """A hybrid with the reference's default material: the hybrid colour index
    (colorMin, aux.color, r / DE), highlights and glow, with shadows."""
    p = scene.reference_material(SCENES["hybrid_menger_power2"][0]())
    p.light.castShadows = 1
    print("[case] generic reference material", flush=True)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    st = compare(g, r)
    _report("hybrid menger + power2 reference material", st)
    _check(st, False)


def test_generic_matches_specialised_kinds(rt):
    """A hybrid of one formula with itself is that formula: the general path and the specialised
    Mandelbox kernel agree on mask and iterations."""
    single = scene.mandelbox_shadow_ao(96, 54, seed=1)
    single.fractal.deltaDEMethod = 2  # analytic
    hyb = scene.copy_params(single)
    hyb.hybridCount = 2
    hyb.hybridFractals[0] = hyb.fractal
    a = rt.render_to_host(single)
    b = rt.render_to_host(hyb)
    assert a["de_kind"] != GENERIC and b["de_kind"] == GENERIC
    assert (a["mask"] == b["mask"]).mean() >= 0.999
    assert (a["iters"] == b["iters"]).mean() >= 0.999


def test_generic_settings_are_validated(rt):
    p = SCENES["hybrid_menger_power2"][0]()
    bad = scene.copy_params(p)
    bad.hybridCount = abi.MBX_MAX_FRACTALS + 1
    with pytest.raises(rt.MbxError, match="hybridCount"):
        rt.render_to_host(bad)
    bad = scene.copy_params(p)
    bad.fractal.formula = 103  # pseudo-Kleinian: not in the table
    with pytest.raises(rt.MbxError, match="not supported"):
        rt.render_to_host(bad)
    bad = scene.copy_params(p)
    bad.fractal.formula = 0
    bad.hybridFractals[0].formula = 0
    with pytest.raises(rt.MbxError, match="none"):
        rt.render_to_host(bad)


@pytest.mark.parametrize("kind", ["box_delta", "box_analytic", "hybrid", "rotated_folds"])
def test_mandelbox_main_rotation_parity(rt, oracle_mod, kind):
    """mandelbox_main_rotation_enabled (fractal_mandelbox.cpp:103: SetRotation2 of
    mandelbox_rotation_main between the offset and the scale) on the specialised FP64 Mandelbox
    kernels and in a hybrid on the general kernel."""
    if kind == "hybrid":
        p = _hybrid([(abi.MBX_FORMULA_MANDELBOX, 2, 1.0, None), (abi.MBX_FORMULA_MENGER_SPONGE, 1, 1.0, None)],
                    camera=(3.0, -6.0, 2.0))
        mb = p.fractal.mandelbox
    else:
        p = scene.mandelbox_shadow_ao(96, 54, seed=1)
        p.fractal.deltaDEMethod = 1 if kind == "box_delta" else 2
        mb = p.fractal.mandelbox
    mb.mainRotationEnabled = 1
    abi.set3(mb.rotationMain, (9.98, 3.995, 4.98))
    if kind == "rotated_folds":
        mb.rotationsEnabled = 1
        for fold, axis, angles in ((0, 0, (5.0, -7.0, 3.0)), (1, 0, (-4.0, 2.0, 6.0)), (0, 2, (3.0, 8.0, -2.0)),
                                   (1, 1, (0.0, 0.0, 0.0))):
            for k in range(3):
                mb.rotation[fold][axis][k] = angles[k]
    print("[case] mandelbox main rotation %s" % kind, flush=True)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    assert r["mask"].mean() > 0.05
    st = compare(g, r)
    _report("mandelbox main rotation %s" % kind, st)
    _check(st, False)
    if kind == "rotated_folds":
        assert g["de_kind"] == GENERIC
    plain = scene.copy_params(p)
    plain.fractal.mandelbox.mainRotationEnabled = 0
    plain.fractal.mandelbox.rotationsEnabled = 0
    assert (oracle_mod.render(plain)["iters"] != r["iters"]).mean() > 0.05, "the rotation changes the fractal"


def _advanced(p, variant):
    p.advancedQuality = 1
    if variant == "fixed_step":      # every step 0.02, horizon 133 steps away
        p.absMinMarchingStep = p.absMaxMarchingStep = 0.02
        p.relMinMarchingStep, p.relMaxMarchingStep = 0.0, 1e300
        p.maxRaymarchingSteps = 134
    else:                            # threshold clamp, relative limits, larger delta-DE offset
        p.detailSizeMax = 2e-3
        p.relMinMarchingStep = 0.2
        p.relMaxMarchingStep = 50.0
        p.absMaxMarchingStep = 0.5
        p.deltaDERelativeDelta = 0.004
    return p


@pytest.mark.parametrize("variant", ["fixed_step", "limits"])
@pytest.mark.parametrize("kind", ["bulb8", "mandelbox_delta", "hybrid"])
def test_advanced_quality_parity(rt, oracle_mod, kind, variant):
    """Advanced-quality march limits on the FP32 Mandelbulb kernel, the FP64 delta-DE Mandelbox
    kernel and the general kernel."""
    if kind == "bulb8":
        p = scene.mandelbulb8(96, 72)
        p.light.castShadows = 0
    elif kind == "mandelbox_delta":
        p = scene.mandelbox_shadow_ao(96, 54, seed=1)
        p.fractal.deltaDEMethod = 1
    else:
        p = SCENES["hybrid_menger_power2"][0]()
    p = _advanced(p, variant)
    print("[case] advanced quality %s %s" % (kind, variant), flush=True)
    g = rt.render_to_host(p)
    r = oracle_mod.render(p)
    assert r["mask"].mean() > 0.05
    st = compare(g, r)
    _report("advanced quality %s %s" % (kind, variant), st)
    if kind == "bulb8":   # FP32 march: depth within 1e-3
        assert st["mask_agree"] >= 0.999 and st["iters_agree"] >= 0.999, st
        assert st["depth_within_1e-3"] >= 0.999, st
        assert st["rgb_within_1e-2"] >= 0.999, st
    else:
        _check(st, False)

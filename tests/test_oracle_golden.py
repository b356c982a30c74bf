"""The CPU oracle against vectors produced by the reference's own formula and RNG sources
(tests/golden/formula_golden.json, written by tests/golden/make_golden.py from
oracle/_ref/formula_ref).  FP64 on both sides with the same libm: bit-exact."""
import json
import os

import numpy as np
import pytest

from mandelbulber2_amd import abi, scene

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "formula_golden.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def _hex(vals):
    return np.array([float.fromhex(v) for v in vals], np.float64)


def _group(records):
    groups = {}
    for r in records:
        groups.setdefault(tuple(r["params"]), []).append(r)
    return groups


def _run(oracle_mod, p, recs):
    z_in = np.stack([_hex(r["z_in"]) for r in recs])
    aux_in = np.stack([_hex(r["aux_in"]) for r in recs])
    z_out = np.stack([_hex(r["z_out"]) for r in recs])
    aux_out = np.stack([_hex(r["aux_out"]) for r in recs])
    z, aux = oracle_mod.formula_step(p, z_in, aux_in)
    return z, aux, z_out, aux_out


def test_mandelbulb_formula_bitexact(oracle_mod, golden):
    groups = _group(golden["records"]["mandelbulb"])
    assert len(groups) == 6
    for params, recs in groups.items():
        power, alpha_deg, beta_deg = (float.fromhex(v) for v in params)
        p = scene.default_params()
        p.fractal.formula = abi.MBX_FORMULA_MANDELBULB
        p.fractal.bulb.power = power
        # degrees, the .fract unit; both sides convert with the same M_PI_180 constant
        p.fractal.bulb.alphaAngleOffset = alpha_deg
        p.fractal.bulb.betaAngleOffset = beta_deg
        z, aux, z_ref, aux_ref = _run(oracle_mod, p, recs)
        np.testing.assert_array_equal(z, z_ref)
        np.testing.assert_array_equal(aux[:, 1], aux_ref[:, 1])  # DE


def test_mandelbox_formula_bitexact(oracle_mod, golden):
    groups = _group(golden["records"]["mandelbox"])
    assert len(groups) == 3
    for params, recs in groups.items():
        v = [float.fromhex(x) for x in params]
        p = scene.default_params()
        f = p.fractal
        f.formula = abi.MBX_FORMULA_MANDELBOX
        mb = f.mandelbox
        mb.scale, mb.foldingLimit, mb.foldingValue = v[0], v[1], v[2]
        mb.foldingSphericalMin, mb.foldingSphericalFixed = v[3], v[4]
        abi.set3(mb.offset, v[5:8])
        abi.set3(mb.colorFactor, v[8:11])
        mb.colorFactorSp1, mb.colorFactorSp2 = v[11], v[12]
        z, aux, z_ref, aux_ref = _run(oracle_mod, p, recs)
        np.testing.assert_array_equal(z, z_ref)
        np.testing.assert_array_equal(aux[:, 1:], aux_ref[:, 1:])  # DE and colour
        # both sphere-fold branches and the plain branch are exercised
        r2 = np.einsum("ij,ij->i", z_ref[:, :3] / v[0], z_ref[:, :3] / v[0])
        assert r2.min() < 10 and r2.max() > 0


def test_kifs_formula_bitexact(oracle_mod, golden):
    groups = _group(golden["records"]["kifs"])
    assert len(groups) == 3
    for params, recs in groups.items():
        v = [float.fromhex(x) for x in params]
        p = scene.default_params()
        f = p.fractal
        f.formula = abi.MBX_FORMULA_KALEIDOSCOPIC_IFS
        ifs = f.IFS
        ifs.scale = v[0]
        abi.set3(ifs.offset, v[1:4])
        abi.set3(ifs.edge, v[4:7])
        ifs.absX, ifs.absY, ifs.absZ = int(v[7]), int(v[8]), int(v[9])
        ifs.mengerSpongeMode = int(v[10])
        ifs.edgeEnabled = int(v[11])
        for i in range(abi.MBX_IFS_VECTOR_COUNT):
            ifs.enabled[i] = 0
        z, aux, z_ref, aux_ref = _run(oracle_mod, p, recs)
        np.testing.assert_array_equal(z, z_ref)
        np.testing.assert_array_equal(aux[:, 1], aux_ref[:, 1])


def test_park_miller_matches_reference(oracle_mod, golden):
    rng = golden["rng"]
    seq, s = oracle_mod.random_ints(1, 64)
    assert seq == rng["randomint"]
    seq2, s = oracle_mod.random_ints(s, 64, 1000)
    assert seq2 == rng["random1000"]
    seq3, s = oracle_mod.random_ints(s, 64, 20000)
    assert seq3 == rng["random20000"]
    assert s == rng["final_seed"]


def test_park_miller_minimal_standard_kat(oracle_mod):
    # Park & Miller (1988): from seed 1 the 10000th value of x <- 16807 x mod (2^31 - 1) is
    # 1043618065.
    seq, _ = oracle_mod.random_ints(1, 10000)
    assert seq[-1] == 1043618065
    x = 1
    for _ in range(50):
        x = (16807 * x) % 2147483647
    assert seq[49] == x


def _run_ex(oracle_mod, p, recs, iters=None, const_c=None):
    z_in = np.stack([_hex(r["z_in"]) for r in recs])
    aux_in = np.stack([_hex(r["aux_in"]) for r in recs])
    z, aux, cc = oracle_mod.formula_step_ex(p, z_in, aux_in, iters, const_c)
    return z, aux, cc, np.stack([_hex(r["z_out"]) for r in recs]), np.stack([_hex(r["aux_out"]) for r in recs])


@pytest.mark.parametrize("tag,formula", [("mandelbulbpower2", abi.MBX_FORMULA_MANDELBULB_POWER2),
                                         ("hypercomplex", abi.MBX_FORMULA_HYPERCOMPLEX),
                                         ("quaternion", abi.MBX_FORMULA_QUATERNION)])
def test_power2_formulas_bitexact(oracle_mod, golden, tag, formula):
    """z^2 in the bulb, hypercomplex and quaternion algebras, DE = 2 r DE (4D z)"""
    recs = golden["records"][tag]
    assert len(recs) == 48
    p = scene.default_params()
    p.fractal.formula = formula
    z, aux, _, z_ref, aux_ref = _run_ex(oracle_mod, p, recs)
    np.testing.assert_array_equal(z, z_ref)
    np.testing.assert_array_equal(aux[:, 1], aux_ref[:, 1])


def test_menger_sponge_formula_bitexact(oracle_mod, golden):
    groups = _group(golden["records"]["mengersponge"])
    assert len(groups) == 2
    for params, recs in groups.items():
        p = scene.default_params()
        p.fractal.formula = abi.MBX_FORMULA_MENGER_SPONGE
        p.fractal.transform.scale3 = float.fromhex(params[0])
        z, aux, _, z_ref, aux_ref = _run_ex(oracle_mod, p, recs)
        np.testing.assert_array_equal(z, z_ref)
        np.testing.assert_array_equal(aux[:, 1], aux_ref[:, 1])


def test_xenodreambuie_formula_bitexact(oracle_mod, golden):
    groups = _group(golden["records"]["xenodreambuie"])
    assert len(groups) == 3
    for params, recs in groups.items():
        power, alpha_deg, beta_deg = (float.fromhex(v) for v in params)
        p = scene.default_params()
        p.fractal.formula = abi.MBX_FORMULA_XENODREAMBUIE
        p.fractal.bulb.power = power
        p.fractal.bulb.alphaAngleOffset = alpha_deg
        p.fractal.bulb.betaAngleOffset = beta_deg
        z, aux, _, z_ref, aux_ref = _run_ex(oracle_mod, p, recs)
        np.testing.assert_array_equal(z, z_ref)
        np.testing.assert_array_equal(aux[:, 1], aux_ref[:, 1])


def test_box_fold_bulb_pow2_formula_bitexact(oracle_mod, golden):
    groups = _group(golden["records"]["boxfoldbulbpow2"])
    assert len(groups) == 2
    for params, recs in groups.items():
        ff, zf, en, s1, o1, o2 = (float.fromhex(v) for v in params)
        p = scene.default_params()
        p.fractal.formula = abi.MBX_FORMULA_BOX_FOLD_BULB_POW2
        t = p.fractal.transform
        t.foldingIntPowFoldFactor, t.foldingIntPowZFactor = ff, zf
        t.analyticDEEnabledFalse = int(en)
        t.analyticDEScale1, t.analyticDEOffset1, t.analyticDEOffset2 = s1, o1, o2
        z, aux, _, z_ref, aux_ref = _run_ex(oracle_mod, p, recs)
        np.testing.assert_array_equal(z, z_ref)
        np.testing.assert_array_equal(aux[:, 1], aux_ref[:, 1])


def test_aexion_formula_bitexact(oracle_mod, golden):
    """the first iteration (aux.i == 0) replaces the constant; later ones add it"""
    recs = golden["records"]["aexion"]
    groups = {}
    for r in recs:
        groups.setdefault(tuple(r["params"][:5]), []).append(r)
    assert len(groups) == 4
    for params, rs in groups.items():
        cadd, it, en, s1, o1 = (float.fromhex(v) for v in params)
        p = scene.default_params()
        p.fractal.formula = abi.MBX_FORMULA_AEXION
        t = p.fractal.transform
        t.aexionCadd, t.analyticDEEnabled, t.analyticDEScale1, t.analyticDEOffset1 = cadd, int(en), s1, o1
        cin = np.stack([_hex(r["params"][5:9]) for r in rs])
        z, aux, cc, z_ref, aux_ref = _run_ex(oracle_mod, p, rs, np.full(len(rs), int(it)), cin)
        np.testing.assert_array_equal(z, z_ref)
        np.testing.assert_array_equal(aux[:, 1], aux_ref[:, 1])
        np.testing.assert_array_equal(cc, np.stack([_hex(r["const_c_out"]) for r in rs]))


@pytest.mark.parametrize("which", ["boxfolding", "sphericalfolding"])
def test_global_foldings_bitexact(oracle_mod, golden, which):
    """BoxFolding / SphericalFolding (legacy_fractal_transforrms.cpp:43-101) against the
    reference's own functions: z, DE and colour bit for bit; the oracle also refreshes aux.r to
    the 4D length of the folded z, as Compute does after each folding (compute_fractal.cpp:192-202)."""
    for params, recs in _group(golden["records"][which]).items():
        limit, value, outer, inner = _hex(params)
        p = scene.mandelbulb8(8, 8)
        p.boxFolding = int(which == "boxfolding")
        p.sphericalFolding = int(which == "sphericalfolding")
        p.boxFoldingLimit, p.boxFoldingValue = limit, value
        p.sphericalFoldingOuter, p.sphericalFoldingInner = outer, inner
        z_in = np.stack([_hex(r["z_in"]) for r in recs])
        aux_in = np.stack([_hex(r["aux_in"]) for r in recs])
        z_out = np.stack([_hex(r["z_out"]) for r in recs])
        aux_out = np.stack([_hex(r["aux_out"]) for r in recs])
        z, aux = oracle_mod.global_foldings(p, z_in, aux_in)
        np.testing.assert_array_equal(z, z_out)
        np.testing.assert_array_equal(aux[:, 1:], aux_out[:, 1:])
        np.testing.assert_array_equal(aux[:, 0], np.sqrt((z_out * z_out).sum(1)))
        # the records exercise every branch
        folded = (z != z_in).any(1)
        assert folded.any() and (~folded).any()

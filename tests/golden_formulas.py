"""The formula golden vectors of tests/golden/formula_golden.json (written by
tests/golden/make_golden.py from the reference's own formula sources) as cases: for each formula
and parameter set, the mbx_fractal to run and the inputs / expected outputs.  The parameter
order of each record's "params" is the harness's (oracle/ref_harness.cpp)."""
import json
import os

import numpy as np

from mandelbulber2_amd import abi, scene

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "formula_golden.json")

# formula tag -> (formula id, whether its reference code calls libm trigonometry / pow)
FORMULAS = {
    "mandelbulb": (abi.MBX_FORMULA_MANDELBULB, True),
    "mandelbulbpower2": (abi.MBX_FORMULA_MANDELBULB_POWER2, False),
    "hypercomplex": (abi.MBX_FORMULA_HYPERCOMPLEX, False),
    "quaternion": (abi.MBX_FORMULA_QUATERNION, False),
    "mengersponge": (abi.MBX_FORMULA_MENGER_SPONGE, False),
    "mandelbox": (abi.MBX_FORMULA_MANDELBOX, False),
    "kifs": (abi.MBX_FORMULA_KALEIDOSCOPIC_IFS, False),
    "xenodreambuie": (abi.MBX_FORMULA_XENODREAMBUIE, True),
    "boxfoldbulbpow2": (abi.MBX_FORMULA_BOX_FOLD_BULB_POW2, False),
    "aexion": (abi.MBX_FORMULA_AEXION, False),
}


def load():
    with open(GOLDEN) as f:
        return json.load(f)


def hexs(vals):
    return np.array([float.fromhex(v) for v in vals], np.float64)


def _fractal(tag, v):
    """The container of one parameter set."""
    p = scene.default_params()
    f = p.fractal
    f.formula = FORMULAS[tag][0]
    if tag in ("mandelbulb", "xenodreambuie"):
        f.bulb.power, f.bulb.alphaAngleOffset, f.bulb.betaAngleOffset = v[0], v[1], v[2]   # degrees
    elif tag == "mengersponge":
        f.transform.scale3 = v[0]
    elif tag == "mandelbox":
        mb = f.mandelbox
        mb.scale, mb.foldingLimit, mb.foldingValue = v[0], v[1], v[2]
        mb.foldingSphericalMin, mb.foldingSphericalFixed = v[3], v[4]
        abi.set3(mb.offset, v[5:8])
        abi.set3(mb.colorFactor, v[8:11])
        mb.colorFactorSp1, mb.colorFactorSp2 = v[11], v[12]
    elif tag == "kifs":
        ifs = f.IFS
        ifs.scale = v[0]
        abi.set3(ifs.offset, v[1:4])
        abi.set3(ifs.edge, v[4:7])
        ifs.absX, ifs.absY, ifs.absZ = int(v[7]), int(v[8]), int(v[9])
        ifs.mengerSpongeMode = int(v[10])
        ifs.edgeEnabled = int(v[11])
        for i in range(abi.MBX_IFS_VECTOR_COUNT):
            ifs.enabled[i] = 0
    elif tag == "boxfoldbulbpow2":
        t = f.transform
        t.foldingIntPowFoldFactor, t.foldingIntPowZFactor = v[0], v[1]
        t.analyticDEEnabledFalse = int(v[2])
        t.analyticDEScale1, t.analyticDEOffset1, t.analyticDEOffset2 = v[3], v[4], v[5]
    elif tag == "aexion":
        t = f.transform
        t.aexionCadd, t.analyticDEEnabled, t.analyticDEScale1, t.analyticDEOffset1 = v[0], int(v[2]), v[3], v[4]
    return p


def cases(golden):
    """Yields (tag, params, iteration index, z_in, aux_in (r, DE, colour), const_c_in or None,
    z_out, aux_out, const_c_out or None) per parameter set."""
    for tag in FORMULAS:
        groups = {}
        for r in golden["records"][tag]:
            key = tuple(r["params"][:5]) if tag == "aexion" else tuple(r["params"])
            groups.setdefault(key, []).append(r)
        for key, recs in groups.items():
            v = hexs(key)
            it = int(v[1]) if tag == "aexion" else 0
            cin = np.stack([hexs(r["params"][5:9]) for r in recs]) if tag == "aexion" else None
            cout = np.stack([hexs(r["const_c_out"]) for r in recs]) if tag == "aexion" else None
            yield (tag, _fractal(tag, v), it, np.stack([hexs(r["z_in"]) for r in recs]),
                   np.stack([hexs(r["aux_in"]) for r in recs]), cin,
                   np.stack([hexs(r["z_out"]) for r in recs]), np.stack([hexs(r["aux_out"]) for r in recs]), cout)

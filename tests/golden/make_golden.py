"""Writes tests/golden/formula_golden.json from the reference's own formula and RNG sources.

Runs oracle/_ref/formula_ref, which `make -C oracle ref` compiles from
mandelbulber2/formula/definition/fractal_{mandelbulb,mandelbox,kaleidoscopic_ifs,
mandelbulb_power2,hypercomplex,quaternion,menger_sponge,xenodreambuie,box_fold_bulb_pow2,
aexion}.cpp, legacy_fractal_transforrms.cpp (the global foldings) and
mandelbulber2/src/common_math.cpp in the reference checkout (see
oracle/ref_harness.cpp).
The fixture holds inputs and outputs only, every double as an exact hex string.

    make -C oracle ref && python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "formula_ref")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    text = subprocess.check_output([exe]).decode()
    records = {"mandelbulb": [], "mandelbox": [], "kifs": [], "mandelbulbpower2": [],
               "hypercomplex": [], "quaternion": [], "mengersponge": [], "xenodreambuie": [],
               "boxfoldbulbpow2": [], "aexion": [], "boxfolding": [], "sphericalfolding": []}
    rng = {}
    for line in text.splitlines():
        tag, _, rest = line.partition(" ")
        if tag in records:
            parts = [s.split() for s in rest.split("|")]
            records[tag].append({"params": parts[0], "z_in": parts[1], "aux_in": parts[2],
                                 "z_out": parts[3], "aux_out": parts[4]})
        elif tag == "aexion_c":
            records["aexion"][-1]["const_c_out"] = rest.split()
        elif tag in ("randomint", "random1000", "random20000"):
            rng[tag] = [int(v) for v in rest.split()]
        elif tag == "final_seed":
            rng["final_seed"] = int(rest)
    out = {
        "source": "reference formula_code and Random() via oracle/_ref/formula_ref",
        "aux_fields": ["r", "DE", "color"],
        "mandelbulb_params": ["power", "alphaAngleOffset_deg", "betaAngleOffset_deg"],
        "mandelbox_params": ["scale", "foldingLimit", "foldingValue", "minR", "fixedR", "offset.x",
                             "offset.y", "offset.z", "color.x", "color.y", "color.z", "Sp1", "Sp2"],
        "kifs_params": ["scale", "offset.x", "offset.y", "offset.z", "edge.x", "edge.y", "edge.z",
                        "absX", "absY", "absZ", "menger", "edgeEnabled"],
        "mengersponge_params": ["scale3"],
        "xenodreambuie_params": ["power", "alphaAngleOffset_deg", "betaAngleOffset_deg"],
        "boxfoldbulbpow2_params": ["foldFactor", "zFactor", "analyticDE.enabledFalse",
                                   "analyticDE.scale1", "analyticDE.offset1", "analyticDE.offset2"],
        "foldings_params": ["boxLimit", "boxValue", "sphericalOuter", "sphericalInner"],
        "aexion_params": ["cadd", "aux.i", "analyticDE.enabled", "analyticDE.scale1",
                          "analyticDE.offset1", "const_c.x", "const_c.y", "const_c.z", "const_c.w"],
        "records": records,
        "rng": rng,
    }
    path = os.path.join(HERE, "formula_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print("wrote", path, {k: len(v) for k, v in records.items()})


if __name__ == "__main__":
    sys.exit(main())

"""The formula-plugin metadata at the C ABI (mbx_formula_count / mbx_formula_get) against the
values each formula's constructor sets in the reference, read by hand from
formula/definition/fractal_<name>.cpp (the constructor lines cited per row) with the enum values
of src/fractal_enums.h:47-99 and the formula ids of all_fractal_list_enums.hpp.  Host only."""
from mandelbulber2_amd import abi, runtime

# internalID, DEType, DEFunctionType, cpixelAddition, defaultBailout, DEAnalyticFunction,
# coloringFunction, internalName, nameInComboBox
ANALYTIC, DELTA = 0, 1                           # enumDEType (fractal_enums.h:70-75)
LINEAR, LOGARITHMIC = 1, 2                       # enumDEFunctionType (:47-60)
CPIXEL_ON, CPIXEL_OFF, CPIXEL_HAS = 0, 1, 2      # enumCPixelAddition (:62-68)
A_LINEAR, A_LOG, A_IFS = 1, 2, 3                 # enumDEAnalyticFunction (:77-88)
C_DEFAULT, C_ABOX, C_IFS = 0, 1, 2               # enumColoringFunction (:90-99)
HAND_READ = [
    (2, ANALYTIC, LOGARITHMIC, CPIXEL_ON, 10.0, A_LOG, C_DEFAULT, "mandelbulb", "Mandelbulb"),   # fractal_mandelbulb.cpp:18-26
    (3, ANALYTIC, LOGARITHMIC, CPIXEL_ON, 10.0, A_LOG, C_DEFAULT, "mandelbulb_power2",
     "Mandelbulb - Power 2"),                                                                    # fractal_mandelbulb_power2.cpp:18-26
    (4, ANALYTIC, LOGARITHMIC, CPIXEL_ON, 10.0, A_LOG, C_DEFAULT, "hypercomplex", "Hypercomplex"),  # fractal_hypercomplex.cpp:20-28
    (5, ANALYTIC, LOGARITHMIC, CPIXEL_ON, 10.0, A_LOG, C_DEFAULT, "quaternion", "Quaternion"),  # fractal_quaternion.cpp:18-26
    (7, ANALYTIC, LINEAR, CPIXEL_OFF, 10.0, A_IFS, C_IFS, "menger_sponge", "Menger sponge"),    # fractal_menger_sponge.cpp:19-27
    (8, ANALYTIC, LINEAR, CPIXEL_ON, 100.0, A_LINEAR, C_ABOX, "mandelbox", "Mandelbox"),        # fractal_mandelbox.cpp:20-28
    (10, ANALYTIC, LINEAR, CPIXEL_OFF, 100.0, A_IFS, C_IFS, "kaleidoscopic_ifs",
     "Kaleidoscopic IFS"),                                                                       # fractal_kaleidoscopic_ifs.cpp:19-27
    (11, ANALYTIC, LOGARITHMIC, CPIXEL_ON, 10.0, A_LOG, C_DEFAULT, "xenodreambuie", "Xenodreambuie"),  # fractal_xenodreambuie.cpp:19-27
    (16, DELTA, LINEAR, CPIXEL_ON, 10000.0, A_LINEAR, C_DEFAULT, "box_fold_bulb_pow2",
     "Box Fold Bulb Pow 2"),                                                                     # fractal_box_fold_bulb_pow2.cpp:17-25
    (19, ANALYTIC, LOGARITHMIC, CPIXEL_HAS, 10000.0, A_LOG, C_DEFAULT, "aexion", "Aexion"),     # fractal_aexion.cpp:18-26
]


def _row(info):
    return (info.internalID, info.DEType, info.DEFunctionType, info.cpixelAddition, info.defaultBailout,
            info.DEAnalyticFunction, info.coloringFunction, info.internalName.decode(),
            info.nameInComboBox.decode())


def test_formula_metadata_matches_the_constructors():
    got = [_row(i) for i in runtime.formulas()]
    assert got == HAND_READ


def test_formula_ids_and_enums_match_the_header():
    ids = sorted(getattr(abi, n) for n in dir(abi) if n.startswith("MBX_FORMULA_") and n != "MBX_FORMULA_NONE")
    assert ids == [r[0] for r in HAND_READ]
    assert (abi.MBX_CPIXEL_ENABLED_BY_DEFAULT, abi.MBX_CPIXEL_DISABLED_BY_DEFAULT, abi.MBX_CPIXEL_ALREADY_HAS) == \
        (CPIXEL_ON, CPIXEL_OFF, CPIXEL_HAS)
    assert (abi.MBX_COLORING_DEFAULT, abi.MBX_COLORING_ABOX, abi.MBX_COLORING_IFS) == (C_DEFAULT, C_ABOX, C_IFS)


def test_formula_get_rejects_bad_index():
    import ctypes
    L = runtime.load()
    info = abi.MbxFormulaInfo()
    assert L.mbx_formula_get(-1, ctypes.byref(info)) != 0
    assert L.mbx_formula_get(L.mbx_formula_count(), ctypes.byref(info)) != 0
    assert L.mbx_formula_get(0, None) != 0


def test_golden_cases_reproduce_on_the_oracle(oracle_mod):
    """The case table the device's formula test uses (tests/golden_formulas.py) gives the
    reference's outputs on the oracle, bit for bit: the parameter mapping is right."""
    import numpy as np
    from tests import golden_formulas as gf
    n = 0
    for tag, p, it, z_in, aux_in, cin, z_out, aux_out, cout in gf.cases(gf.load()):
        iters = np.full(len(z_in), it, np.int32) if cin is not None else None
        z, aux, cc = oracle_mod.formula_step_ex(p, z_in, aux_in, iters, cin)
        np.testing.assert_array_equal(z, z_out, err_msg=tag)
        np.testing.assert_array_equal(aux[:, 1], aux_out[:, 1], err_msg=tag)
        if tag == "mandelbox":
            np.testing.assert_array_equal(aux[:, 2], aux_out[:, 2], err_msg=tag)
        if cout is not None:
            np.testing.assert_array_equal(cc, cout, err_msg=tag)
        n += len(z_in)
    assert n >= 400

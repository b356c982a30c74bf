""".fract scene reader (mbx_load_fract) and parameter defaults (mbx_params_default).

Host logic only: these run without a GPU.  Every golden test runs twice, on the library's C++
reader and on the oracle's restatement (oracle/fract.py), with expected values read by hand
from the fixture text (tests/golden/fract, unmodified example scenes the reference's example
test loads) or written from the reference's closed forms:
  * the pre-2.21 field of view, atan(fov / 2) * 360 / pi (settings.cpp:1611-1627), and the
    53.13 -> 1.0 special case of the old default;
  * the pre-2.09 swap of delta_DE_function 0 <-> 2, the pre-2.07 "non-linear means
    logarithmic" rule and the pre-2.20 "preferred method means preferred function" rule
    (settings.cpp:1344-1357, 1500-1505, 1596-1606), each on an input where the rule changes
    the result;
  * toRGBFloat's c / 65536 (color_structures.hpp:106-109) on the fixtures' hexadecimal colours.
Then both readers must agree field for field on every fixture and, where the reference tree is
present, on every example scene it ships.
"""
import glob
import math
import os

import numpy as np
import pytest

from mandelbulber2_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(HERE, "golden", "fract")
EXAMPLES = "/root/reference/mandelbulber2/deploy/share/mandelbulber2"


def _read(name):
    with open(os.path.join(FIXTURES, name), "rb") as f:
        return f.read()


def _tolist(v):
    try:
        return [_tolist(x) for x in v]
    except TypeError:
        return v


class _CppResult:
    def __init__(self, p, info):
        self.p, self.info = p, info

    def __getitem__(self, path):
        obj = self.p
        for part in path.split("."):
            obj = obj[int(part)] if part.isdigit() else getattr(obj, part)
        return _tolist(obj)

    def flags(self):
        return set(abi.fract_flag_names(self.info.unsupported))

    def meta(self, k):
        return getattr(self.info, k)


class _OracleResult:
    def __init__(self, out, info):
        self.out, self.info = out, info

    def __getitem__(self, path):
        v = self.out[path]
        return int(v) if isinstance(v, bool) else _tolist(v)

    def flags(self):
        from oracle import fract as of
        return {k for k, v in of.FLAGS.items() if self.info["unsupported"] & v}

    def meta(self, k):
        return self.info[k]


class ReaderError(Exception):
    pass


def _cpp(text):
    from mandelbulber2_amd import runtime
    try:
        return _CppResult(*runtime.load_fract(text))
    except runtime.MbxError as e:
        raise ReaderError(str(e))


def _oracle(text):
    from oracle import fract as of
    if isinstance(text, bytes):
        text = text.decode("utf-8", errors="replace")
    try:
        return _OracleResult(*of.load(text))
    except of.FractError as e:
        raise ReaderError(str(e))


@pytest.fixture(params=["library", "oracle"])
def load(request):
    return _cpp if request.param == "library" else _oracle


def _rgb(r, g, b):
    return [float(np.float32(c / 65536.0)) for c in (r, g, b)]


def _hdr(version, body, fmt="only modified parameters"):
    return "# Mandelbulber settings file\n# version %s\n# %s\n[main_parameters]\n%s" % (version, fmt, body)


# ------------------------------------------------------------------------------------------
# defaults
# ------------------------------------------------------------------------------------------
def test_library_defaults_equal_python_defaults():
    from mandelbulber2_amd import runtime, scene
    assert bytes(runtime.params_default()) == bytes(scene.default_params(800, 600))


def test_empty_scene_is_the_defaults(load):
    r = load(_hdr("2.33", ""))
    from mandelbulber2_amd import scene
    d = _CppResult(scene.default_params(800, 600), None)
    for k in ("imageWidth", "imageHeight", "fov", "DEFactor", "detailLevel", "camera", "target",
              "topVector", "surfaceColor", "backgroundColor1", "fractal.formula",
              "fractal.formulaMaxiter", "fractal.bulb.power", "fractal.IFS.direction"):
        assert r[k] == d[k], k
    # a 2.33 file that names no light has none (the light list starts empty, settings.cpp:689-690)
    assert r["light.enabled"] == 0
    # palette colouring, specular highlights and glow are on by default and rendered
    assert r["useColorsFromPalette"] == 1 and r["specular"] == 5.0 and r["glowEnabled"] == 1
    assert r.flags() == set()


# ------------------------------------------------------------------------------------------
# fixtures: values read from the file text
# ------------------------------------------------------------------------------------------
def test_mandelbulb001(load):
    r = load(_read("mandelbulb001.fract"))
    assert r.meta("fileVersion") == 2.21 and r.meta("format") == 1 and r.meta("errors") == 0
    assert r["imageWidth"] == 1920 and r["imageHeight"] == 1080
    assert r["camera"] == [-0.1588900287502842, -1.18529433988786, 0.1644950620698425]
    assert r["target"] == [-0.6034600287502843, 0.4738634876330364, 0.624747793594753]
    assert r["topVector"] == [0.06698729810778062, -0.2499999999999999, 0.9659258262890683]
    assert r["fov"] == 57.62          # saved by 2.21: already degrees
    assert r["DEFactor"] == 0.645654 and r["detailLevel"] == 5.0
    assert r["fractal.formula"] == abi.MBX_FORMULA_MANDELBULB     # no formula line: the default
    assert r["fractal.bulb.power"] == 9.0 and r["fractal.formulaMaxiter"] == 250
    # pre-2.25 file: light #1 is created with its defaults (settings.cpp:1650-1654)
    assert r["light.enabled"] == 1 and r["light.rotation"] == [-45.0, 45.0, 0.0]
    assert r["light.intensity"] == 1.0
    assert r["ambientOcclusionEnabled"] == 1 and r["ambientOcclusionMode"] == abi.MBX_AO_SCREEN_SPACE
    assert r["ambientOcclusionQuality"] == 7 and r["SSAORandomMode"] == 0
    # screen-space AO is rendered (mbx_render_ssao); iteration fog is not
    assert r["iterFogEnabled"] == 1 and r["basicFogEnabled"] == 0 and r["volFogEnabled"] == 0
    assert r.flags() == {"fog"}
    # image adjustments: float members of sImageAdjustments (fractparams.cpp:155-159)
    assert r["image.brightness"] == float(np.float32(1.28))
    assert r["image.gamma"] == float(np.float32(0.6899999999999999))
    assert r["image.contrast"] == 1.0 and r["image.saturation"] == 1.0 and r["image.hdr"] == 0
    # mat1_surface_color_gradient "0 fd6029 999 698403 1999 fff59b ... 9000 aba53c": 10 stops and
    # the first colour again at 1.0, positions as float(n) / 10000
    assert r["surfaceGradientCount"] == 11
    pos = r["surfaceGradientPosition"]
    assert pos[0] == 0.0 and pos[1] == float(np.float32(999) / np.float32(10000)) and pos[10] == 1.0
    assert r["surfaceGradientColor"][0] == [0xfd, 0x60, 0x29] == r["surfaceGradientColor"][10]
    assert r["surfaceGradientColor"][9] == [0xab, 0xa5, 0x3c]


def test_mandelbox001(load):
    r = load(_read("mandelbox001.fract"))
    assert r["fractal.formula"] == abi.MBX_FORMULA_MANDELBOX
    assert r["fractal.mandelbox.scale"] == -1.73
    assert r["detailLevel"] == 2.0
    assert r["surfaceColor"] == _rgb(0x0b00, 0x2400, 0x4c00)
    assert r["light.enabled"] == 1 and r["light.intensity"] == 4.0
    assert r["light.rotation"] == [-45.0, 45.0, 0.0]
    assert r["reflectionsMax"] == 3
    # the volumetric distance fog is rendered: its values as the file writes them, fog colour 1
    # at its default (initparameters.cpp:247)
    assert r["volFogEnabled"] == 1 and r["volFogDensity"] == 0.125
    assert r["volFogColour1Distance"] == 0.00192060288419113
    assert r["volFogColour2Distance"] == 0.00384120576838226
    assert r["volFogDistanceFactor"] == 0.00384120576838226
    assert r["volFogDistanceFromSurface"] == 7.8125e-18
    assert r["volFogColour"] == [list(_rgb(30000, 30000, 30000)), list(_rgb(0xff00, 0, 0)),
                                 list(_rgb(0x5500, 0, 0))]
    assert r["distanceFogShadows"] == 0 and r["volumetricLightDEFactor"] == 1.0
    assert r.flags() == {"reflections"}
    assert r["image.brightness"] == 2.0 and r["image.hdr"] == 1
    assert r["useColorsFromPalette"] == 0 and r["specular"] == 1.0 and r["metallic"] == 0


def test_ifs001_pre_2_25_lights(load):
    r = load(_read("IFS_001.fract"))
    assert r["fractal.formula"] == abi.MBX_FORMULA_KALEIDOSCOPIC_IFS
    assert r["fractal.IFS.scale"] == 3.0 and r["fractal.IFS.offset"] == [1.0, 1.0, 1.0]
    assert (r["fractal.IFS.absX"], r["fractal.IFS.absY"], r["fractal.IFS.absZ"]) == (1, 1, 1)
    assert r["fractal.IFS.mengerSpongeMode"] == 1
    assert r["fractal.IFS.enabled"] == [0, 0, 0, 0, 0, 1, 1, 1, 0]
    assert r["fractal.IFS.direction"][5] == [1.0, -1.0, 0.0]
    assert r["fractal.IFS.direction"][7] == [0.39, 1.0, -1.0]
    # main_light_enable -> light1_enabled; aux_light_enabled_1 -> light2_enabled
    assert r["light.enabled"] == 0
    assert r["ambientOcclusion"] == 1.04 and r["ambientOcclusionQuality"] == 3
    assert r["ambientOcclusionMode"] == abi.MBX_AO_MULTIPLE_RAYS
    assert {"lights", "ao_lightmap"} <= r.flags() and "postprocess" not in r.flags()
    assert r["image.hdr"] == 1


def test_pre_2_07_keyframe_file(load):
    """Version 2.05: fov absent -> the old default 53.13 means 1.0 -> 2 atan(1/2) in degrees."""
    r = load(_read("keyframe_anim_mandelbox_boxes.fract"))
    assert r.meta("fileVersion") == 2.05
    assert r["fov"] == pytest.approx(2.0 * math.degrees(math.atan(0.5)), rel=1e-15)
    assert r["fov"] == pytest.approx(53.13010235415598, rel=1e-15)
    assert r["light.rotation"] == [-76.43000000000001, 30.96, 0.0]
    assert r["light.relativePosition"] == 0
    assert r["fractal.formula"] == abi.MBX_FORMULA_MANDELBOX
    assert r["fractal.mandelbox.scale"] == -1.5
    assert r["fractal.IFS.rotationEnabled"] == 1      # settings.cpp:1507-1513
    assert r["backgroundColor3"] == _rgb(0x3d00, 0x1c00, 0x500)
    assert r["ambientOcclusion"] == 0.45
    assert {"animation", "postprocess"} <= r.flags()
    assert not {"fog", "ao_screen_space"} & r.flags()
    assert r["volFogEnabled"] == 1 and r["volFogColour1Distance"] == 0.7762942609158674
    assert r["volFogColour"][2] == list(_rgb(0xbc00, 0x9700, 0x5b00))


def test_mandelbulb_power_8_global_constant_factor(load):
    r = load(_read("mandelbulb_power_8.fract"))
    assert r["fractal.bulb.power"] == 8.0
    assert r["fractal.bulb.alphaAngleOffset"] == 22.5 and r["fractal.bulb.betaAngleOffset"] == 22.5
    # pre-2.35 single fractal: the main-section constant factor becomes the node's
    assert r["fractal.constantFactor"] == [1.1, 1.1, 1.1]
    assert r["DEFactor"] == 0.3 and r["fov"] == 28.07 and r["detailLevel"] == 4.0
    assert r["light.rotation"] == [-45.0, -45.0, 0.0] and r["light.intensity"] == 0.3
    assert r["background3ColorsEnable"] == 0 and r["backgroundColor1"] == [0.0, 0.0, 0.0]
    assert "ao_lightmap" in r.flags() and "postprocess" not in r.flags() and "antialiasing" not in r.flags()
    assert r["antialiasingEnabled"] == 1 and r["antialiasingSize"] == 2
    assert r["image.gamma"] == float(np.float32(0.7))
    assert "palette" not in r.flags()


def test_menger_sponge_and_mandelbox_22(load):
    r = load(_read("menger_sponge_001.fract"))
    assert r["fractal.IFS.rotations"][6] == [45.0, 0.0, 0.0]
    assert r["viewDistanceMax"] == 9.46211700403319 and r["DEFactor"] == 0.81298828125
    assert "lights" in r.flags()                      # light1_volumetric
    r = load(_read("mandelbox_22.fract"))
    assert r["fractal.mandelbox.foldingSphericalMin"] == 0.1
    assert r["light.rotation"] == [-27.61, 21.51, 0.0] and r["light.intensity"] == 5.6
    assert r["ambientOcclusionMode"] == abi.MBX_AO_FAST
    assert not {"ao_lightmap", "ao_screen_space"} & r.flags()


def test_boolean_scene_is_reported(load):
    r = load(_read("boolean001.fract"))
    assert {"hybrid", "transform"} <= r.flags()


def _containers(r):
    n = r["hybridCount"]
    return [(r["fractal.formula"], r["fractal.formulaIterations"])] + [
        (r["hybridFractals.%d.formula" % k], r["hybridFractals.%d.formulaIterations" % k])
        for k in range(n - 1)]


def test_node_tree_hybrid(load):
    # 2.35 object tree: a hybrid node (type 10) whose children are the fractal nodes 2-4;
    # node n holds fractal n - 1 ([fractal_1] formula 8 iterations 2, [fractal_2] 7 / 3,
    # [fractal_3] 3 with the default single iteration)
    r = load(_read("hybrid002.fract"))
    assert r.meta("fileVersion") == 2.35
    assert "hybrid" not in r.flags()
    assert r["hybridCount"] == 3
    assert _containers(r) == [(abi.MBX_FORMULA_MANDELBOX, 2), (abi.MBX_FORMULA_MENGER_SPONGE, 3),
                              (abi.MBX_FORMULA_MANDELBULB_POWER2, 1)]
    assert r["fractal.mandelbox.scale"] == -0.4
    assert r["repeatFrom"] == 1


def test_legacy_hybrids(load):
    # formula_N / formula_iterations_N of the main section, [fractal_N] parameters; containers
    # with formula none (hybrid14's fractal_3 and fractal_4) leave the sequence
    r = load(_read("hybrid003.fract"))
    assert "hybrid" not in r.flags()
    assert _containers(r) == [(8, 2), (7, 3), (3, 1)]
    r = load(_read("hybrid16.fract"))
    assert _containers(r) == [(8, 3), (10, 3), (5, 1)]
    assert r["fractal.juliaMode"] == 1
    assert r["fractal.mandelbox.scale"] == -1.5 and r["hybridFractals.0.IFS.scale"] == 2.62
    r = load(_read("hybrid14.fract"))
    assert r.meta("fileVersion") == 2.05
    assert _containers(r) == [(5, 20), (7, 250)]
    assert r["hybridFractals.0.bulb.power"] == 8.0 and r["DEThresh"] == 0.03


def test_advanced_quality_and_transform_parameters(load):
    r = load(_read("hybrid22_foldingIntPow.fract"))
    assert _containers(r) == [(8, 2), (8, 2), (abi.MBX_FORMULA_BOX_FOLD_BULB_POW2, 40)]
    assert r["fractal.mandelbox.scale"] == 1.4 and r["hybridFractals.0.mandelbox.scale"] == 1.7
    assert r["hybridFractals.1.transform.foldingIntPowZFactor"] == 10.0
    assert r["hybridFractals.1.transform.foldingIntPowFoldFactor"] == 2.0   # default
    assert r["fractal.deltaDEFunction"] == 1
    assert r["advancedQuality"] == 1 and r["deltaDERelativeDelta"] == 0.005
    assert r["constantDEThreshold"] == 1 and r["DEThresh"] == 5e-05
    assert r["absMaxMarchingStep"] == 3.0 and r["relMinMarchingStep"] == 0.01   # defaults


# ------------------------------------------------------------------------------------------
# synthetic files: parser rules
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("text,what", [
    ("", "header"),
    ("# settings\n# version 2.2\n# only modified parameters\n", "header"),
    ("# Mandelbulber settings file\n# 2.2\n# only modified parameters\n", "version"),
    ("# Mandelbulber settings file\n# version two\n# only modified parameters\n", "version"),
    ("# Mandelbulber settings file\n# version 2.36\n# only modified parameters\n", "newer"),
    ("# Mandelbulber settings file\n# version 2.2\n# application settings\n", "application"),
    ("# Mandelbulber settings file\n# version 2.2\n# something\n", "format"),
])
def test_bad_headers_are_rejected(load, text, what):
    with pytest.raises(ReaderError):
        load(text)


def test_three_malformed_lines_pass_four_fail(load):
    bad = "no_semicolon 1\nsemicolon;before\n;x y\n"
    r = load(_hdr("2.33", bad + "fov 40;\n"))
    assert r.meta("errors") == 3 and r["fov"] == 40.0
    with pytest.raises(ReaderError):
        load(_hdr("2.33", bad + "another_bad\nfov 40;\n"))


def test_value_rules(load):
    r = load(_hdr("2.33", "\n".join([
        "fov 12,5;",                     # comma decimal
        "'DE_factor' '0.25';",           # quotes stripped
        "detail_level 2;", "detail_level 3;",   # last line wins
        "constant_DE_threshold 1;",      # only "true" is true
        "interior_mode true;",
        "image_width 3;",                # clamped to the registered minimum 5
        "view_distance_max 0;",          # clamped to 1e-15
        "background_color_2 1ffff 10 zz;",  # hex, clamped to 65535; bad parts read 0
        "camera 1 2;",                   # a short vector keeps the default's third part
        "some_unknown_parameter 7;",     # not an error here
        "perspective_type fish_eye;",    # enumeration by name
    ]) + "\n"))
    assert r["fov"] == 12.5 and r["DEFactor"] == 0.25 and r["detailLevel"] == 3.0
    assert r["constantDEThreshold"] == 0 and r["interiorMode"] == 1
    assert r["imageWidth"] == 5 and r["viewDistanceMax"] == 1e-15
    assert r["backgroundColor2"] == _rgb(65535, 0x10, 0)
    assert r["camera"] == [1.0, 2.0, 2.0]
    assert r["perspectiveType"] == 1 and r.meta("errors") == 0
    assert "perspective" not in r.flags() and "interior" not in r.flags()


@pytest.mark.parametrize("version,fov,ptype,want", [
    ("2.20", "1,2", "0", math.degrees(math.atan(0.6)) * 2.0),
    ("2.20", "53,13", "0", math.degrees(math.atan(0.5)) * 2.0),
    ("2.20", "0,5", "fish_eye", 90.0),
    ("2.20", "0,25", "2", 90.0),
    ("2.20", "0,5", "3", 90.0),
    ("2.21", "1,2", "0", 1.2),
])
def test_pre_2_21_field_of_view(load, version, fov, ptype, want):
    r = load(_hdr(version, "fov %s;\nperspective_type %s;\n" % (fov, ptype)))
    assert r["fov"] == pytest.approx(want, rel=1e-14)


@pytest.mark.parametrize("version,body,want", [
    ("2.05", "delta_DE_method 1;", 2),                    # <= 2.06: non-linear -> logarithmic
    ("2.05", "delta_DE_method 1;\ndelta_DE_function 1;", 1),
    ("2.08", "delta_DE_method 1;\ndelta_DE_function 0;", 2),  # < 2.09: 0 <-> 2
    ("2.08", "delta_DE_method 1;\ndelta_DE_function 2;", 0),
    ("2.10", "delta_DE_method 1;\ndelta_DE_function 2;", 2),
    ("2.10", "delta_DE_function 1;", 0),                  # < 2.20, preferred method -> preferred
    ("2.20", "delta_DE_function 1;", 1),
    ("2.05", "linear_DE_mode true;\ndelta_DE_method 1;", 1),  # <= 2.06 rename
])
def test_delta_de_function_migrations(load, version, body, want):
    assert load(_hdr(version, body + "\n"))["fractal.deltaDEFunction"] == want


def test_legacy_fractal_lines_and_transforms(load):
    r = load(_hdr("2.20", "formula_1 8;\nformula_rotation_1 10 20 30;\nN 40;\n"
                  "[fractal_1]\nmandelbox_scale 3;\n"))
    assert r["fractal.formula"] == abi.MBX_FORMULA_MANDELBOX
    assert r["fractal.mandelbox.scale"] == 3.0
    assert r["formulaRotation"] == [10.0, 20.0, 30.0] and r["fractal.formulaMaxiter"] == 40
    # a non-default global fractal_rotation overrides every fractal's (settings.cpp:2370-2379)
    r = load(_hdr("2.20", "formula_1 8;\nformula_rotation_1 10 20 30;\nfractal_rotation 1 2 3;\n"))
    assert r["formulaRotation"] == [1.0, 2.0, 3.0]
    r = load(_hdr("2.20", "fractal_position 0 0 1;\n"))
    assert "transform" in r.flags()


def test_legacy_boolean_single_fractal(load):
    body = ("boolean_operators true;\nformula_1 10;\njulia_mode_1 true;\n"
            "julia_c_1 0,1 0,2 0,3;\nN 33;\nformula_maxiter_1 77;\n")
    r = load(_hdr("2.28", body))
    assert r["fractal.formula"] == abi.MBX_FORMULA_KALEIDOSCOPIC_IFS
    assert r["fractal.juliaMode"] == 1 and r["fractal.juliaC"] == [0.1, 0.2, 0.3]
    assert r["fractal.formulaMaxiter"] == 33           # before 2.29 the node takes N
    assert load(_hdr("2.29", body))["fractal.formulaMaxiter"] == 77
    assert "hybrid" not in r.flags()


def test_objects_tree_node(load):
    body = ("node_0001_definition mandelbox 1,1,1,0,1;\nnode_0001_rotation 10 20 30;\n"
            "node_0001_formula_maxiter 40;\nnode_0001_julia_mode true;\nN 99;\n"
            "[fractal_1]\nformula 8;\nmandelbox_scale -2;\n")
    r = load(_hdr("2.35", body))
    assert r["fractal.formula"] == abi.MBX_FORMULA_MANDELBOX
    assert r["formulaRotation"] == [10.0, 20.0, 30.0]
    assert r["fractal.formulaMaxiter"] == 40 and r["fractal.juliaMode"] == 1
    assert r["fractal.mandelbox.scale"] == -2.0
    r = load(_hdr("2.35", body.replace("N 99;", "node_0001_enabled false;")))
    assert "disabled" in r.flags()


def test_light_versions(load):
    r = load(_hdr("2.24", "main_light_alpha 10;\nmain_light_beta 20;\nmain_light_colour ff00 0 0;\n"))
    assert r["light.rotation"] == [10.0, 20.0, 0.0] and r["light.color"] == _rgb(0xff00, 0, 0)
    r = load(_hdr("2.25", "light1_rotation 1 2 3;\n"))
    assert r["light.enabled"] == 1 and r["light.rotation"] == [1.0, 2.0, 3.0]
    assert load(_hdr("2.30", "fov 30;\n"))["light.enabled"] == 0
    # pre-2.28 iteration fog divided light #1 by its boost, which pre-2.13 files had as 100
    r = load(_hdr("2.12", "iteration_fog_enable true;\nlight1_intensity 2;\n"))
    assert r["light.intensity"] == pytest.approx(0.02, rel=1e-15)
    r = load(_hdr("2.20", "iteration_fog_enable true;\niteration_fog_brightness_boost 4;\n"
                          "light1_intensity 2;\n"))
    assert r["light.intensity"] == 0.5
    r = load(_hdr("2.24", "aux_light_enabled_1 true;\n"))
    assert "lights" in r.flags()


def test_pre_2_15_material_defaults(load):
    # a material first named by a pre-2.15 file gets specular 1, width 1 and metallic off
    # (settings.cpp:2107-2119); the file's own lines still win
    r = load(_hdr("2.14", "mat1_shading 0,5;\n"))
    assert r["shading"] == 0.5
    assert (r["specular"], r["specularWidth"], r["metallic"]) == (1.0, 1.0, 0)
    r = load(_hdr("2.14", "mat1_specular 0;\n"))
    assert r["specular"] == 0.0
    r = load(_hdr("2.15", "mat1_shading 0,5;\n"))
    assert (r["specular"], r["specularWidth"], r["metallic"]) == (5.0, 0.05, 1)
    r = load(_hdr("2.14", "fov 30;\n"))          # no material line: InitMaterialParams defaults
    assert r["specular"] == 5.0


def _stops(r):
    n = r["surfaceGradientCount"]
    return [(r["surfaceGradientPosition"][i], tuple(r["surfaceGradientColor"][i])) for i in range(n)]


def test_gradient_strings(load):
    # unsorted stops are sorted; the first colour is repeated at 1.0
    r = load(_hdr("2.33", "mat1_surface_color_gradient 5000 00ff00 0 ff0000;\n"))
    assert _stops(r) == [(0.0, (255, 0, 0)), (0.5, (0, 255, 0)), (1.0, (0, 255, 0))]
    r = load(_hdr("2.33", "mat1_surface_color_gradient 0 ff0000 5000 00ff00;\n"))
    assert _stops(r) == [(0.0, (255, 0, 0)), (0.5, (0, 255, 0)), (1.0, (255, 0, 0))]
    # a single token: white at both ends (color_gradient.cpp:311-319)
    r = load(_hdr("2.33", "mat1_surface_color_gradient 0;\n"))
    assert _stops(r) == [(0.0, (255, 255, 255)), (1.0, (255, 255, 255))]


def test_pre_2_19_palette_migration(load):
    # surface_color_palette -> gradient with evenly spread positions (settings.cpp:1375-1396),
    # then offset / size and speed * 10 / size with size = tokens / 2 (settings.cpp:1563-1593)
    r = load(_hdr("2.18", "mat1_surface_color_palette ff0000 00ff00 0000ff;\n"
                          "mat1_coloring_speed 3;\nmat1_coloring_palette_offset 6;\n"))
    assert _stops(r) == [(0.0, (255, 0, 0)), (float(np.float32(3333) / np.float32(10000)), (0, 255, 0)),
                         (float(np.float32(6666) / np.float32(10000)), (0, 0, 255)), (1.0, (255, 0, 0))]
    assert r["coloringSpeed"] == pytest.approx(3.0 * 10.0 / 3.0, rel=1e-15)
    assert r["coloringPaletteOffset"] == pytest.approx(2.0, rel=1e-15)


# ------------------------------------------------------------------------------------------
# the two readers agree
# ------------------------------------------------------------------------------------------
def _agree(text):
    try:
        a = _cpp(text)
    except ReaderError:
        a = None
    try:
        b = _oracle(text)
    except ReaderError:
        b = None
    assert (a is None) == (b is None)
    if a is None:
        return 0
    for k in b.out:
        av, bv = a[k], b[k]
        assert av == bv, (k, av, bv)
    for k in ("fileVersion", "format", "lines", "errors", "formula", "unsupported"):
        assert a.meta(k) == b.meta(k), k
    return 1


@pytest.mark.parametrize("name", sorted(os.listdir(FIXTURES)) if os.path.isdir(FIXTURES) else [])
def test_readers_agree_on_fixtures(name):
    if not name.endswith(".fract"):
        pytest.skip("not a scene")
    assert _agree(_read(name)) == 1


@pytest.mark.skipif(not os.path.isdir(EXAMPLES), reason="reference example scenes not present")
def test_readers_agree_on_every_example():
    files = sorted(glob.glob(os.path.join(EXAMPLES, "**", "*.fract"), recursive=True))
    assert len(files) > 100
    loaded = 0
    for f in files:
        with open(f, "rb") as fh:
            loaded += _agree(fh.read())
    assert loaded >= len(files) - 2     # the 1.x-format defaults file is rejected by both


def test_global_foldings_and_slow_shading(load):
    """folded mender sponge.fract: box_folding true, box_folding_limit 0,1, box_folding_value 0,2,
    spherical_folding true (outer / inner radii at their defaults 1 and 0.5, initparameters.cpp:199-204);
    aexion05.fract: slow_shading true with N 47."""
    r = load(_read("folded_mender_sponge.fract"))
    assert r["boxFolding"] == 1 and r["sphericalFolding"] == 1
    assert r["boxFoldingLimit"] == 0.1 and r["boxFoldingValue"] == 0.2
    assert r["sphericalFoldingOuter"] == 1.0 and r["sphericalFoldingInner"] == 0.5
    assert r["slowShading"] == 0
    r = load(_read("aexion05.fract"))
    assert r["slowShading"] == 1 and r["fractal.formulaMaxiter"] == 47
    assert r["boxFolding"] == 0 and r["sphericalFolding"] == 0
    r = load(_hdr("2.33", "spherical_folding true;\nspherical_folding_outer 1,5;\nspherical_folding_inner 0.25;\n"))
    assert r["sphericalFolding"] == 1 and r["sphericalFoldingOuter"] == 1.5 and r["sphericalFoldingInner"] == 0.25


def test_all_lights_factors_and_main_n(load):
    """cLight (light.cpp:77-87) scales light 1's intensity, visibility and size by the
    all_lights_ factors; the main section's N (initparameters.cpp:130, default 250) is kept apart
    from the node's maxiter for slow shading (shader_calculate_normals.cpp:103)."""
    body = ("light1_is_defined true;\nlight1_intensity 1,5;\nlight1_visibility 2;\nlight1_size 0,25;\n"
            "all_lights_intensity 2;\nall_lights_visibility 0,5;\nall_lights_size 4;\nN 47;\n")
    r = load(_hdr("2.33", body))
    assert r["light.intensity"] == 3.0 and r["light.visibility"] == 1.0 and r["light.size"] == 1.0
    assert r["N"] == 47 and r["fractal.formulaMaxiter"] == 47      # pre-2.35: N is the maxiter
    r = load(_hdr("2.35", body))
    assert r["N"] == 47 and r["fractal.formulaMaxiter"] == 250     # 2.35: the node's own maxiter
    r = load(_hdr("2.33", "light1_is_defined true;\n"))
    assert r["light.intensity"] == 1.0 and r["N"] == 250
    r = load(_hdr("2.27", "iteration_fog_enable true;\niteration_fog_brightness_boost 4;\n"
                          "light1_is_defined true;\nlight1_intensity 2;\nall_lights_intensity 3;\n"))
    assert r["light.intensity"] == 1.5       # (2 / 4) * 3: settings.cpp:1700-1712, then cLight


@pytest.mark.parametrize("body,flag", [
    ("DOF_monte_carlo true;\nMC_soft_shadows_enable true;\n", None),
    ("DOF_monte_carlo true;\nDOF_MC_CA_enable true;\n", None),
    ("DOF_monte_carlo true;\nDOF_enabled true;\n", None),     # the Monte Carlo lens: rendered
    ("DOF_enabled true;\n", "postprocess"),                     # the post-processing blur
    ("nebula_mode true;\n", "nebula"),
])
def test_monte_carlo_effect_and_nebula_flags(load, body, flag):
    got = load(_hdr("2.33", body)).flags()
    assert ({"mc_effects", "nebula", "postprocess"} & got) == ({flag} if flag else set())


def test_monte_carlo_lens_aberration_and_soft_shadows(load):
    """DOF_focus / DOF_radius / DOF_MC_CA_camera_dispersion with their limits
    (initparameters.cpp:323-338) and the two switches."""
    r = load(_hdr("2.33", "DOF_monte_carlo true;\nDOF_enabled true;\nDOF_focus 2,5;\nDOF_radius 300;\n"
                          "DOF_MC_CA_enable true;\nDOF_MC_CA_camera_dispersion 0,5;\nMC_soft_shadows_enable true;\n"))
    assert r["DOFEnabled"] == 1 and r["DOFFocus"] == 2.5 and r["DOFRadius"] == 200.0
    assert r["DOFMonteCarloChromaticAberration"] == 1 and r["DOFMonteCarloCACameraDispersion"] == 0.5
    assert r["monteCarloSoftShadows"] == 1
    r = load(_hdr("2.33", ""))
    assert (r["DOFEnabled"], r["DOFFocus"], r["DOFRadius"], r["DOFMonteCarloCACameraDispersion"]) == (0, 6.0, 10.0, 1.0)

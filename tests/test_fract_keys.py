"""Key coverage of the .fract reader: every parameter line of every shipped example scene is
either consumed by the reader, covered by an effect flag that the reader sets for that scene, or
shown to have no effect on a still frame of the reference's CPU renderer.

The reader (oracle/fract.py, which the library's reader matches field for field and flag for
flag on every example: test_fract.py) records where each line went and which names it looked up.
A line the reader did not look up must match one rule below.  Each rule names the reference code
that decides the key's role:
  * "inert" keys are read by no part of the CPU render of one frame (user-interface state,
    animation, export, OpenCL-only settings), or only by a path that the scene turns off;
  * "gated" keys act only through a feature switch; when the scene turns the switch on, the
    reader must have set the named flag.
A key matching no rule fails the test: a new silently dropped key cannot go unnoticed.

Fractal-section keys are judged per container: a container the scene's fractal node does not
render is inert; one whose formula is outside the device's table must carry the `formula` flag;
for a table formula, a key the formula's iteration reads (FORMULA_READS, from the formula
sources' `fractal->` fields and fractal.cpp's key names) must have been consumed.
"""
import glob
import os
import re

import pytest

from oracle import fract as of

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(HERE, "golden", "fract")
EXAMPLES = "/root/reference/mandelbulber2/deploy/share/mandelbulber2"


def _v(c, name, default=None):
    """A container value without marking it read."""
    p = dict.get(c, name)
    return default if p is None else p.value


# ---- formulas the device renders: the file keys each iteration reads ----------------------
_BULB = {"power", "alpha_angle_offset", "beta_angle_offset"}        # fractal.cpp:54-56
_MBOX = {"mandelbox_scale", "mandelbox_folding_limit", "mandelbox_folding_value",
         "mandelbox_folding_min_radius", "mandelbox_folding_fixed_radius", "mandelbox_offset",
         "mandelbox_rotation_main", "mandelbox_color", "mandelbox_color_R", "mandelbox_color_Sp1",
         "mandelbox_color_Sp2", "mandelbox_rotations_enabled", "mandelbox_main_rotation_enabled"} | \
    {"mandelbox_rotation_%s_%d" % (s, i) for s in ("neg", "pos") for i in (1, 2, 3)}  # :60-80
_IFS = {"IFS_scale", "IFS_rotation", "IFS_rotation_enabled", "IFS_offset", "IFS_edge",
        "IFS_edge_enabled", "IFS_abs_x", "IFS_abs_y", "IFS_abs_z", "IFS_menger_sponge_mode"} | \
    {"IFS_%s_%d" % (k, i) for k in ("direction", "rotations", "distance", "intensity", "enabled")
     for i in range(9)}                                                  # :96-114
FORMULA_READS = {
    2: _BULB,                                     # fractal_mandelbulb.cpp
    3: set(),                                     # fractal_mandelbulb_power2.cpp
    4: set(),                                     # fractal_hypercomplex.cpp
    5: set(),                                     # fractal_quaternion.cpp
    7: {"transf_scale_3", "fold_color_aux_color_enabled_false", "fold_color_difs_0000",
        "fold_color_start_iterations_A", "fold_color_stop_iterations_A"},  # fractal_menger_sponge.cpp
    8: _MBOX,                                     # fractal_mandelbox.cpp
    10: _IFS,                                     # fractal_kaleidoscopic_ifs.cpp
    11: _BULB,                                    # fractal_xenodreambuie.cpp
    16: {"analyticDE_enabled_false", "analyticDE_offset_1", "analyticDE_offset_2",
         "analyticDE_scale_1", "boxfold_bulbpow2_folding_factor",
         "boxfold_bulbpow2_z_factor"},            # fractal_box_fold_bulb_pow2.cpp
    19: {"cadd", "analyticDE_enabled", "analyticDE_offset_1", "analyticDE_scale_1"},  # fractal_aexion.cpp
}
# the Menger sponge's fold colour feeds aux.color, which only palette colouring reads
# (fractal_menger_sponge.cpp:40-55); the reader flags `palette` for it
_MENGER_COLOR = {"fold_color_aux_color_enabled_false", "fold_color_difs_0000",
                 "fold_color_start_iterations_A", "fold_color_stop_iterations_A"}


# ---- main-section rules ------------------------------------------------------------------
INERT = "inert"


def _gate(flag, switch=None):
    """Gated by `switch(par, node)` (True when the key can act; node holds the rendered node's
    material); the reader must set `flag`."""
    return ("gate", flag, switch)


def _on(*names):
    return lambda par, m: any(_v(par, n) for n in names)


def _mc(par, m):
    return bool(_v(par, "DOF_monte_carlo"))


MAIN_RULES = [
    # user-interface, animation, export and OpenCL-only state
    (r"(keyframe|flight|anim|animsound|frames_per|voxel|mesh_export)_.*", INERT),
    (r"camera_rotation|camera_distance_to_target", INERT),   # PrepareMainVectors derives both
                                                             # from camera/target/top (render_worker.cpp:541-549)
    (r"image_proportion", INERT),                            # interface state (initparameters.cpp:68)
    (r"legacy_coordinate_system", INERT),                    # manipulations.cpp:97 (camera moves only)
    (r"view_distance_min", INERT),                           # render_worker.cpp:331 starts at 0
    (r"analityc_DE_mode|aux_light_intensity|aux_light_place_behind", INERT),  # 1.x settings
    (r"antialiasing_ocl_depth", INERT),                      # OpenCL engine only
    (r"MC_denoiser_.*", INERT),                              # read by the OpenCL engine only
    (r"(mat|light)\d+_is_defined", INERT),                   # registry markers
    (r"mat\d+_name", INERT),
    (r"boolean_operator_\d+", _gate("hybrid")),              # legacy boolean trees
    # pre-2.35 selectors the migration to the objects tree resolves (settings.cpp:2551-2760):
    # the boolean path takes its node from the name_N temporaries, not from these
    (r"formula_\d+|hybrid_fractal_enable|julia_c", INERT),
    # feature switches
    # the fog-limiting primitives are commented out in the reference
    # (shader_volumetric.cpp:229-238, 250-259, 336-342)
    (r"(basic_fog|distance_fog|iteration_fog)_primitive", INERT),
    # basic and distance fog are read; iteration fog and clouds act only when switched on
    (r"(iteration_fog|clouds)_.*", _gate("fog", _on("iteration_fog_enable", "clouds_enable"))),
    # the scattering colours and the GI of volumetrics light only iteration fog, clouds, shadowed
    # fog and volumetric lights (shader_volumetric.cpp:364-468; the reader flags the first three,
    # and volumetric lights raise the lights flag)
    (r"rayleigh_scattering_.*|MC_fog_illumination", _gate("fog", lambda p, m: _v(p, "iteration_fog_enable") or
                                                          _v(p, "clouds_enable") or
                                                          (_v(p, "basic_fog_enabled") and _v(p, "basic_fog_cast_shadows")) or
                                                          (_v(p, "volumetric_fog_enabled") and _v(p, "distance_fog_shadows")))),
    (r"(fake_lights)_.*", _gate("lights", _on("fake_lights_enabled"))),
    (r"(random_lights)_.*", _gate("lights", _on("random_lights_group"))),
    (r"primitive_.*", _gate("primitives")),
    (r"limit_(min|max)", _gate("transform", _on("limits_enabled"))),
    (r"stereo_.*", _gate("perspective", _on("stereo_enabled"))),
    (r"(file_background|background_rotation|background_v_scale|background_h_scale|"
     r"background_texture_offset_[xy]|textured_background_map_type)",
     _gate("textures", _on("textured_background"))),
    (r"file_envmap|env_mapping_.*", _gate("textures", _on("env_mapping_enable"))),
    (r"file_lightmap", _gate("ao_lightmap", lambda p, m: _v(p, "ambient_occlusion_enabled") and
                             _v(p, "ambient_occlusion_mode") == 1)),
    (r"hdr_blur_.*", _gate("postprocess", _on("hdr_blur_enabled"))),
    (r"post_chromatic_aberration_.*", _gate("postprocess", _on("post_chromatic_aberration_enabled"))),
    (r"DOF_(blur_opacity|number_of_passes|HDR|max_radius)",  # the post-processing DOF only
     _gate("postprocess", lambda p, m: _v(p, "DOF_enabled") and not _v(p, "DOF_monte_carlo"))),
    # the aberration's dispersion gain bends refraction only (render_worker.cpp:1105-1110)
    # (of the rendered node's material; the switch's second argument holds the node)
    (r"DOF_MC_CA_dispersion_gain", _gate("reflections", lambda p, node: _mc(p, node) and
                                         _v(p, "DOF_MC_CA_enable") and _v(p, "raytraced_reflections") and
                                         (_v(p, "mat%d_reflectance" % node["material"], 0.0) > 0.0 or
                                          _v(p, "mat%d_transparency_of_surface" % node["material"], 0.0) > 0.0))),
    (r"MC_global_illumination_volumetric|volumetric_light_DE_Factor",
     _gate("glow", lambda p, m: _v(p, "glow_enabled") and _v(p, "glow_intensity") > 0.0)),
    (r"nebula_mode", _gate("nebula", _on("nebula_mode"))),
    (r"nebula_.*", _gate("nebula", _on("nebula_mode"))),
    (r"all_lights_.*", INERT),           # consumed with light 1; without it, other lights are flagged
]


def _light_rule(par, n, key, flags):
    """lightN_key: lights other than #1 act only when enabled (lights flag); light 1's position,
    target, cone and decay act only for non-directional types, its volumetric visibility only
    with `volumetric` (light.cpp:60-120; the reader flags both)."""
    if key in ("is_defined",):
        return None
    if n != 1:
        if not _v(par, "light%d_enabled" % n, False) or "lights" in flags:
            return None
        return "light %d enabled without the lights flag" % n
    if not _v(par, "light1_enabled", False) or "lights" in flags:
        return None
    if key in ("position", "target", "use_target_point", "cone_angle", "cone_soft_angle",
               "decayFunction", "repeat_texture", "file_texture"):
        return None if _v(par, "light1_type", 0) == 0 else "light 1 " + key
    if key in ("volumetric_visibility",):
        return None if not _v(par, "light1_volumetric") else "light 1 " + key
    return "light 1 key %s not consumed" % key


_MAT_RULES = [
    # ray-traced reflections and refraction (render_worker.cpp:1420-1470)
    (r"fresnel_reflectance|reflections_color.*|transparency_.*|reflectance",
     lambda par, mp, key: _v(par, "raytraced_reflections") and (
         _v(par, mp + "reflectance", 0.0) > 0.0 or _v(par, mp + "transparency_of_surface", 0.0) > 0.0),
     "reflections"),
    (r"iridescence_.*", lambda par, mp, key: _v(par, mp + "iridescence_enabled"), "material"),
    (r"surface_roughness|rough_surface", lambda par, mp, key: _v(par, mp + "rough_surface"), "material"),
    (r"perlin_noise_.*", lambda par, mp, key: _v(par, mp + "perlin_noise_enable"), "material"),
    (r"subsurface_.*", lambda par, mp, key: _v(par, mp + "subsurface_scattering"), "material"),
    (r"(file|use)_.*texture|texture_.*|.*_texture_.*|normal_map_.*|displacement_.*",
     lambda par, mp, key: any(_v(par, mp + "use_%s_texture" % t) for t in (
         "color", "diffusion", "displacement", "luminosity", "normal_map", "reflectance",
         "transparency", "roughness", "transparency_alpha")), "textures"),
    # a gradient acts with palette colours and its own switch (material.cpp:360-380)
    (r"(specular|diffuse|luminosity|roughness|reflectance|transparency)_gradient(_enable)?",
     lambda par, mp, key: _v(par, mp + "use_colors_from_palette") and
     _v(par, mp + key.split("_gradient")[0] + "_gradient_enable"), "palette"),
    # the extra colouring options act with palette colours, in the extra-colour branch
    # (fractal_coloring.cpp:118-290) or, for hybrids, with the extra options switch (:305-314)
    (r"fractal_coloring_.*",
     lambda par, mp, key: _v(par, mp + "use_colors_from_palette") and (
         _v(par, mp + "fractal_coloring_extra_color_enabled_false") or
         _v(par, mp + "fractal_coloring_extra_color_options_false")), "palette"),
    (r"inside_coloring", None, None),                  # read by the OpenCL engine only
]


def _material_rule(par, mp, key, flags):
    for pat, switch, flag in _MAT_RULES:
        if re.fullmatch(pat, key):
            if switch is None or not switch(par, mp, key) or flag in flags:
                return None
            return "material key %s acts without the %s flag" % (key, flag)
    return "material key %s not consumed" % key


def check_scene(text):
    """Returns the list of lines that the reader neither consumed nor covered."""
    trace = {}
    try:
        _, info = of.load(text, trace)
    except of.FractError:
        return []
    par, fract, legacy = trace["containers"]
    read_main = set(getattr(par, "read", ()))
    read_fract = [set(getattr(f, "read", ())) for f in fract]
    flags = {k for k, v in of.FLAGS.items() if info["unsupported"] & v}
    node = trace["node"] or {}
    used = set(node.get("members") or ([node["object"]] if node else []))
    material = node.get("material", 1)
    problems = []
    for section, raw, target in trace["lines"]:
        kind = target[0]
        if kind == "skip":
            continue
        if kind == "legacy":
            # julia/maxiter temporaries feed only a boolean scene's single node
            # (settings.cpp:2128-2172, 2551-2620)
            if target[1] in legacy.read or not trace["boolean"] or "hybrid" in flags:
                continue
            problems.append("legacy %s" % raw)
            continue
        if kind == "fractal":
            idx, name = target[1], target[2]
            if name in read_fract[idx]:
                continue
        else:
            name = target[1]
            if section != "main_parameters":
                idx = of._qtoint(section[8:]) - 1 if "fractal" in section else -1
                if not 0 <= idx < 9:
                    continue
            elif kind == "main" and name in read_main:
                continue
        if kind == "fractal" or section != "main_parameters":
            if idx + 1 not in used or "formula" in flags or "disabled" in flags or "hybrid" in flags:
                continue
            fo = _v(fract[idx], "formula")
            reads = FORMULA_READS.get(fo)
            if reads is None:
                problems.append("formula %d of container %d is not flagged" % (fo, idx + 1))
            elif name in reads:
                if fo == 7 and name in _MENGER_COLOR and (
                        "palette" in flags or not _v(par, "mat%d_use_colors_from_palette" % material)):
                    continue
                problems.append("fractal %d key %s (formula %d) not consumed" % (idx + 1, name, fo))
            continue
        m = re.fullmatch(r"light(\d+)_(.*)", name)
        if m:
            p = _light_rule(par, int(m.group(1)), m.group(2), flags)
            if p:
                problems.append(p)
            continue
        m = re.fullmatch(r"mat(\d+)_(.*)", name)
        if m and m.group(2) != "is_defined" and m.group(2) != "name":
            if int(m.group(1)) != material:
                continue       # a material no rendered object uses (primitives are flagged)
            p = _material_rule(par, "mat%d_" % material, m.group(2), flags)
            if p:
                problems.append(p)
            continue
        for pat, rule in MAIN_RULES:
            if re.fullmatch(pat, name):
                if rule == INERT:
                    break
                _, flag, switch = rule
                if flag in flags or (switch is not None and not switch(par, {"material": material})):
                    break
                problems.append("%s acts without the %s flag" % (name, flag))
                break
        else:
            problems.append("main key %s matches no rule" % name)
    return problems


@pytest.mark.parametrize("name", sorted(n for n in os.listdir(FIXTURES) if n.endswith(".fract")))
def test_fixture_keys_covered(name):
    with open(os.path.join(FIXTURES, name), encoding="utf-8", errors="replace") as f:
        assert check_scene(f.read()) == []


@pytest.mark.skipif(not os.path.isdir(EXAMPLES), reason="reference example scenes not present")
def test_every_example_key_covered():
    files = sorted(glob.glob(os.path.join(EXAMPLES, "**", "*.fract"), recursive=True))
    assert len(files) > 100
    bad = {}
    for fn in files:
        with open(fn, encoding="utf-8", errors="replace") as f:
            p = check_scene(f.read())
        if p:
            bad[os.path.relpath(fn, EXAMPLES)] = p
    assert not bad, "%d scenes: %s" % (len(bad), list(bad.items())[:5])


def test_unknown_key_is_reported():
    text = "# Mandelbulber settings file\n# version 2.33\n# only modified parameters\n" \
           "[main_parameters]\nsome_future_effect 1;\n"
    assert check_scene(text) == ["main key some_future_effect matches no rule"]


def test_gated_key_without_its_flag_is_reported(monkeypatch):
    """A reader that forgot the flag: the colour of enabled fake lights then fails."""
    text = "# Mandelbulber settings file\n# version 2.33\n# only modified parameters\n" \
           "[main_parameters]\nfake_lights_enabled true;\nfake_lights_color 0 0 0;\n"
    assert check_scene(text) == []
    real = of.load

    def forgetful(t, trace=None):
        out, info = real(t, trace)
        return out, dict(info, unsupported=info["unsupported"] & ~of.FLAGS["lights"])
    monkeypatch.setattr(of, "load", forgetful)
    assert check_scene(text) == ["fake_lights_color acts without the lights flag"]

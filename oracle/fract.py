"""Oracle for the .fract scene reader: a plain-Python restatement of how cSettings loads a
Mandelbulber settings file (mandelbulber2/src/settings.cpp) into typed parameter containers
(cParameterContainer / cOneParameter / cMultiVal: parameters.cpp, one_parameter.cpp:51-150,
multi_val.cpp:87-123), restricted to the parameters the raymarch path reads.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of mbx_load_fract; never by the
product.  Unlike the library's reader (which keeps raw strings per section and types them on
read), this one keeps typed containers that are set line by line, as the reference does:
  * InitParams / InitFractalParams / InitMaterialParams / InitLightParams / InitNodeParams
    register (type, default, limits) (initparameters.cpp line numbers below);
  * Decode walks the sections (settings.cpp:742-905), DecodeOneLine splits and renames each line
    (settings.cpp:1015-1245, Compatibility 1262-1492) and Set stores the value with the
    parameter's type and limits;
  * Compatibility2 (settings.cpp:1494-1754) and the migration to the objects tree
    (settings.cpp:2324-2402, 2551-2760) then run on the containers.
Every function cites the reference lines it follows.  Parity pins: tests/test_fract.py checks
hand-read values of example files and the closed forms of the version migrations.
"""
import math
import struct

# ------------------------------------------------------------------------------------------
# typed values (cMultiVal)
# ------------------------------------------------------------------------------------------
BOOL, INT, DOUBLE, VEC3, RGB, STRING = "bool", "int", "double", "vec3", "rgb", "string"


def _qtoint(s, base=10):
    """QString::toInt: the whole trimmed string must parse, else 0."""
    s = s.strip()
    try:
        v = int(s, base)
    except ValueError:
        return 0
    return v if -2 ** 31 <= v < 2 ** 31 else 0


def _qtodouble(s):
    """QLocale::toDouble of a value after everyLocaleDouble (settings.cpp:2063-2072)."""
    s = s.strip().replace(",", ".")
    try:
        return float(s)
    except ValueError:
        return 0.0


class Param:
    def __init__(self, typ, default, lo=None, hi=None, enum=None):
        self.typ, self.default, self.lo, self.hi, self.enum = typ, default, lo, hi, enum
        self.value = default

    def set_text(self, text):
        """cOneParameter::Set(QString) -> cMultiVal::Store + LimitValue."""
        t = self.typ
        if t == BOOL:
            self.value = text == "true"        # DecodeOneLine turns "true" into 1, anything else 0
        elif t == INT:
            if self.enum is not None and text in self.enum:
                v = self.enum.index(text)      # GetIndexByEnumeration
            else:
                v = _qtoint(text)
            self.value = self._limit(v)
        elif t == DOUBLE:
            self.value = self._limit(_qtodouble(text))
        elif t == VEC3:
            parts = text.split(" ")
            v = list(self.value)
            for j in range(min(len(parts), 3)):
                v[j] = _qtodouble(parts[j])
            self.value = tuple(v)
        elif t == RGB:
            parts = text.split(" ")
            v = list(self.value)
            for j in range(min(len(parts), 3)):
                v[j] = min(max(_qtoint(parts[j], 16), 0), 65535)
            self.value = tuple(v)
        else:
            self.value = text

    def _limit(self, v):
        if self.lo is not None and v < self.lo:
            return self.lo
        if self.hi is not None and v > self.hi:
            return self.hi
        return v


class Container(dict):
    """A parameter container; `read` collects the names the reader looked up, which the
    key-coverage test compares with the keys a scene sets."""

    def add(self, name, *spec, **kw):
        self[name] = Param(*spec, **kw)

    def get(self, name, default=None):
        self.__dict__.setdefault("read", set()).add(name)
        p = dict.get(self, name)
        return default if p is None else p.value

    def setv(self, name, value):
        self[name].value = value


# ------------------------------------------------------------------------------------------
# registrations (initparameters.cpp)
# ------------------------------------------------------------------------------------------
def init_params():
    """InitParams, the registrations this oracle needs (initparameters.cpp:64-402)."""
    c = Container()
    c.add("image_width", INT, 800, 5, 65535)                        # :64
    c.add("image_height", INT, 600, 5, 65535)                       # :65
    c.add("antialiasing_enabled", BOOL, False)                      # :69
    c.add("antialiasing_size", INT, 2, 1, 10)                       # :70
    c.add("camera", VEC3, (3.0, -6.0, 2.0))                         # :113
    c.add("target", VEC3, (0.0, 0.0, 0.0))                          # :114
    c.add("camera_top", VEC3, (-0.1277753, 0.2555506, 0.958314))    # :116
    c.add("fov", DOUBLE, 53.13, 0.0, 720.0)                         # :120
    c.add("perspective_type", INT, 0, enum=["three_point", "fish_eye", "equirectangular",
                                            "fish_eye_cut"])         # :121-122
    c.add("sweet_spot_horizontal_angle", DOUBLE, 0.0, -180.0, 180.0)  # :124
    c.add("sweet_spot_vertical_angle", DOUBLE, 0.0, -90.0, 90.0)    # :125
    c.add("julia_mode", BOOL, False)                                # :128
    c.add("julia_c", VEC3, (0.0, 0.0, 0.0))                         # :129
    c.add("N", INT, 250, 1, 99999)                                  # :130
    c.add("minN", INT, 1, 0, 99999)                                 # :131
    c.add("fractal_constant_factor", VEC3, (1.0, 1.0, 1.0))         # :132
    c.add("detail_level", DOUBLE, 1.0, 1e-8, 1e8)                   # :133
    c.add("DE_thresh", DOUBLE, 0.01, 1e-15, 1e5)                    # :135
    c.add("smoothness", DOUBLE, 1.0, 1e-15, 1e15)                   # :136
    c.add("iteration_threshold_mode", BOOL, False)                  # :137
    c.add("DE_factor", DOUBLE, 1.0, 1e-15, 1e15)                    # :139
    c.add("view_distance_max", DOUBLE, 50.0, 1e-15, 1e15)           # :141
    c.add("limits_enabled", BOOL, False)                            # :145
    c.add("interior_mode", BOOL, False)                             # :147
    c.add("constant_DE_threshold", BOOL, False)                     # :148
    c.add("hybrid_fractal_enable", BOOL, False)                     # :149
    c.add("repeat_from", INT, 1, 1, 9)                              # :151
    c.add("linear_DE_offset", DOUBLE, 0.0)                          # :157
    c.add("bailout", DOUBLE, 100.0, 1.0, 1e15)                      # :150
    c.add("delta_DE_function", INT, 0, 0, 6)                        # :152-153
    c.add("delta_DE_method", INT, 0, 0, 2)                          # :154
    c.add("use_default_bailout", BOOL, True)                        # :155
    c.add("initial_waxis", DOUBLE, 0.0)                             # :156
    c.add("advanced_quality", BOOL, False)                          # :159
    c.add("abs_max_marching_step", DOUBLE, 3.0, 1e-15, 1e15)        # :160
    c.add("abs_min_marching_step", DOUBLE, 1e-15, 1e-15, 1e15)      # :161
    c.add("rel_max_marching_step", DOUBLE, 1e4, 1e-15, 1e15)        # :162
    c.add("rel_min_marching_step", DOUBLE, 1e-2, 1e-15, 1e15)       # :163
    c.add("detail_size_max", DOUBLE, 1.0)                           # :164
    c.add("detail_size_min", DOUBLE, 1e-12)                         # :165
    c.add("deltade_relative_delta", DOUBLE, 0.01, 1e-15, 1e6)       # :166
    c.add("max_raymarching_steps", INT, 10000, 1, 100000000)        # :167
    c.add("stereo_enabled", BOOL, False)                            # :170
    c.add("brightness", DOUBLE, 1.0, 0.0, 1e15)                     # :208
    c.add("contrast", DOUBLE, 1.0, 0.0, 1e15)                       # :209
    c.add("gamma", DOUBLE, 1.0, 0.0, 1e15)                          # :210
    c.add("hdr", BOOL, False)                                       # :211
    c.add("slow_shading", BOOL, False)                              # :140
    c.add("box_folding", BOOL, False)                               # :199
    c.add("box_folding_limit", DOUBLE, 1.0)                         # :200
    c.add("box_folding_value", DOUBLE, 2.0)                         # :201
    c.add("spherical_folding", BOOL, False)                         # :202
    c.add("spherical_folding_outer", DOUBLE, 1.0)                   # :203
    c.add("spherical_folding_inner", DOUBLE, 0.5)                   # :204
    c.add("saturation", DOUBLE, 1.0, 0.0, 1e15)                     # :212
    c.add("ambient_occlusion", DOUBLE, 1.0, 0.0, 1e15)              # :213
    c.add("ambient_occlusion_quality", INT, 4, 1, 10)               # :214
    c.add("ambient_occlusion_fast_tune", DOUBLE, 1.0, 1e-5, 1e5)    # :215
    c.add("ambient_occlusion_enabled", BOOL, False)                 # :216
    c.add("ambient_occlusion_mode", INT, 2)                         # :217-218 (AOModeScreenSpace)
    c.add("ambient_occlusion_color", RGB, (65535, 65535, 65535))    # :219
    c.add("ao_light_map_rotation", DOUBLE, 0.0, -180.0, 180.0)      # :220
    c.add("SSAO_random_mode", BOOL, False)                          # :221
    c.add("glow_enabled", BOOL, True)                               # :222
    c.add("glow_intensity", DOUBLE, 0.2, 0.0, 1e15)                 # :223
    c.add("glow_color_1", RGB, (40984, 44713, 49490))               # :240
    c.add("glow_color_2", RGB, (57192, 60888, 62408))               # :241
    c.add("textured_background", BOOL, False)                       # :224
    c.add("background_brightness", DOUBLE, 1.0, 0.0, 1000.0)        # :227
    c.add("background_gamma", DOUBLE, 1.0, 0.0, 1000.0)             # :228
    c.add("raytraced_reflections", BOOL, False)                     # :236
    c.add("reflections_max", INT, 5, 0, 1000)                       # :237
    c.add("env_mapping_enable", BOOL, False)                        # :238
    c.add("background_color_1", RGB, (0, 38306, 65535))             # :242
    c.add("background_color_2", RGB, (65535, 65535, 65535))         # :243
    c.add("background_color_3", RGB, (0, 10000, 500))               # :244
    c.add("background_3_colors_enable", BOOL, True)                 # :245
    c.add("fog_color_1", RGB, (30000, 30000, 30000))                # :247
    c.add("fog_color_2", RGB, (0, 30000, 65535))                    # :248
    c.add("fog_color_3", RGB, (65535, 65535, 65535))                # :249
    c.add("volumetric_fog_enabled", BOOL, False)                    # :251
    c.add("volumetric_fog_density", DOUBLE, 0.5, 0.0, 1e15)         # :252
    c.add("volumetric_fog_colour_1_distance", DOUBLE, 1.0, 0.0, 1e15)  # :253
    c.add("volumetric_fog_colour_2_distance", DOUBLE, 2.0, 0.0, 1e15)  # :254
    c.add("volumetric_fog_distance_factor", DOUBLE, 1.0, 0.0, 1e15)    # :255
    c.add("volumetric_fog_distance_from_surface", DOUBLE, 1e-15, 0.0, 1e15)  # :256-257
    c.add("distance_fog_shadows", BOOL, False)                      # :258
    c.add("iteration_fog_enable", BOOL, False)                      # :261
    c.add("iteration_fog_brightness_boost", DOUBLE, 1.0, 0.0, 1e6)  # :270
    c.add("clouds_enable", BOOL, False)                             # :274
    c.add("hdr_blur_enabled", BOOL, False)                          # :299
    c.add("post_chromatic_aberration_enabled", BOOL, False)         # :303
    c.add("fill_light_color", RGB, (0, 0, 0))                       # :309
    c.add("all_lights_intensity", DOUBLE, 1.0, 0.0, 100000.0)       # :310
    c.add("all_lights_visibility", DOUBLE, 1.0, 0.0, 100000.0)      # :311
    c.add("all_lights_size", DOUBLE, 1.0, 0.0, 100000.0)            # :312
    c.add("basic_fog_enabled", BOOL, False)                         # :315
    c.add("basic_fog_visibility", DOUBLE, 20.0, 0.0, 1e15)          # :316
    c.add("basic_fog_color", RGB, (59399, 61202, 65535))            # :317
    c.add("basic_fog_cast_shadows", BOOL, False)                    # :319
    c.add("DOF_enabled", BOOL, False)                               # :322
    c.add("DOF_focus", DOUBLE, 6.0, 0.0, 10000.0)                   # :323
    c.add("DOF_radius", DOUBLE, 10.0, 0.0, 200.0)                   # :324
    c.add("DOF_MC_CA_camera_dispersion", DOUBLE, 1.0, 1e-15, 1000.0)  # :338
    c.add("DOF_monte_carlo", BOOL, False)                           # :329
    c.add("DOF_samples", INT, 100)                                  # :330
    c.add("DOF_min_samples", INT, 10)                               # :331
    c.add("DOF_max_noise", DOUBLE, 1.0, 0.00001, 100.0)             # :332
    c.add("DOF_MC_global_illumination", BOOL, False)                # :333
    c.add("MC_global_illumination_volumetric", BOOL, False)         # :334
    c.add("DOF_MC_CA_enable", BOOL, False)                          # :336
    c.add("MC_soft_shadows_enable", BOOL, False)                    # :339
    c.add("MC_GI_radiance_limit", DOUBLE, 10.0, 0.001, 1e10)        # :340
    c.add("volumetric_light_DE_Factor", DOUBLE, 1.0, 1e-15, 1e15)   # :350
    c.add("random_lights_group", BOOL, False)                       # :360
    c.add("fake_lights_enabled", BOOL, False)                       # :374
    c.add("nebula_mode", BOOL, False)                               # :404
    c.add("fractal_position", VEC3, (0.0, 0.0, 0.0))                # :400
    c.add("fractal_rotation", VEC3, (0.0, 0.0, 0.0))                # :401
    c.add("repeat", VEC3, (0.0, 0.0, 0.0))                          # :402
    return c


def init_fractal(index):
    """InitFractalParams for container fractal<index> (initparameters.cpp:675-768)."""
    c = Container()
    c.add("formula", INT, 2 if index == 0 else 0)                   # :693-700
    c.add("formula_iterations", INT, 1, 1, 65536)
    c.add("formula_weight", DOUBLE, 1.0, 0.0, 1.0)
    c.add("formula_start_iteration", INT, 0, 0, 65536)
    c.add("formula_stop_iteration", INT, 250, 0, 65536)
    c.add("julia_mode", BOOL, False)                                # :705
    c.add("julia_c", VEC3, (0.0, 0.0, 0.0))
    c.add("fractal_constant_factor", VEC3, (1.0, 1.0, 1.0))
    c.add("initial_waxis", DOUBLE, 0.0)
    c.add("formula_material_id", INT, 1)                            # :710
    c.add("fractal_enable", BOOL, True)                             # :711
    c.add("dont_add_c_constant", BOOL, False)                        # :713
    c.add("check_for_bailout", BOOL, True)
    c.add("formula_maxiter", INT, 250, 1, 9999)                     # :715
    c.add("power", DOUBLE, 9.0)                                     # :717
    c.add("alpha_angle_offset", DOUBLE, 0.0)
    c.add("beta_angle_offset", DOUBLE, 0.0)
    c.add("gamma_angle_offset", DOUBLE, 0.0)
    c.add("IFS_scale", DOUBLE, 2.0)                                 # :725
    c.add("IFS_rotation", VEC3, (0.0, 0.0, 0.0))
    c.add("IFS_rotation_enabled", BOOL, False)
    c.add("IFS_offset", VEC3, (1.0, 0.0, 0.0))
    c.add("IFS_edge", VEC3, (0.0, 0.0, 0.0))
    c.add("IFS_edge_enabled", BOOL, False)
    c.add("IFS_abs_x", BOOL, False)
    c.add("IFS_abs_y", BOOL, False)
    c.add("IFS_abs_z", BOOL, False)
    c.add("IFS_menger_sponge_mode", BOOL, False)                    # :734
    for i in range(9):                                              # :736-743
        c.add("IFS_direction_%d" % i, VEC3, (1.0, 0.0, 0.0))
        c.add("IFS_rotations_%d" % i, VEC3, (0.0, 0.0, 0.0))
        c.add("IFS_distance_%d" % i, DOUBLE, 0.0)
        c.add("IFS_intensity_%d" % i, DOUBLE, 1.0)
        c.add("IFS_enabled_%d" % i, BOOL, False)
    c.add("mandelbox_scale", DOUBLE, 2.0)                           # :746
    c.add("mandelbox_folding_limit", DOUBLE, 1.0)
    c.add("mandelbox_folding_value", DOUBLE, 2.0)
    c.add("mandelbox_folding_min_radius", DOUBLE, 0.5)
    c.add("mandelbox_folding_fixed_radius", DOUBLE, 1.0)
    c.add("mandelbox_offset", VEC3, (0.0, 0.0, 0.0))                # :752
    c.add("mandelbox_rotation_main", VEC3, (0.0, 0.0, 0.0))
    for i in range(1, 4):                                           # :754-759
        c.add("mandelbox_rotation_neg_%d" % i, VEC3, (0.0, 0.0, 0.0))
        c.add("mandelbox_rotation_pos_%d" % i, VEC3, (0.0, 0.0, 0.0))
    c.add("mandelbox_color", VEC3, (0.03, 0.05, 0.07))              # :761
    c.add("mandelbox_color_R", DOUBLE, 0.0)                         # :764
    c.add("mandelbox_color_Sp1", DOUBLE, 0.2)
    c.add("mandelbox_color_Sp2", DOUBLE, 0.2)
    c.add("mandelbox_main_rotation_enabled", BOOL, False)           # :767
    c.add("mandelbox_rotations_enabled", BOOL, False)               # :768
    c.add("cadd", DOUBLE, -1.3)                                     # :722
    c.add("boxfold_bulbpow2_folding_factor", DOUBLE, 2.0)           # :784
    c.add("boxfold_bulbpow2_z_factor", DOUBLE, 5.0)                 # :785
    c.add("analyticDE_enabled", BOOL, True)                         # :904
    c.add("analyticDE_enabled_false", BOOL, False)                  # :905
    c.add("analyticDE_scale_1", DOUBLE, 1.0)                        # :906
    c.add("analyticDE_offset_1", DOUBLE, 1.0)                       # :909
    c.add("analyticDE_offset_2", DOUBLE, 1.0)                       # :910
    c.add("fold_color_aux_color_enabled_false", BOOL, False)        # :917
    c.add("transf_scale_3", DOUBLE, 3.0)                            # :1010
    return c


def init_material(c, m):
    """InitMaterialParams (initparameters.cpp:1600-1810), the parameters read here."""
    p = "mat%d_" % m
    c.add(p + "is_defined", BOOL, False)
    c.add(p + "surface_color", RGB, (50000, 50000, 50000))
    c.add(p + "shading", DOUBLE, 1.0, 0.0, 1e15)
    c.add(p + "luminosity", DOUBLE, 0.0, 0.0, 1e15)
    c.add(p + "luminosity_color", RGB, (65535, 65535, 65535))
    c.add(p + "use_colors_from_palette", BOOL, True)                 # :1780-1781
    c.add(p + "surface_gradient_enable", BOOL, True)
    c.add(p + "surface_color_gradient", STRING, DEFAULT_GRADIENT)
    c.add(p + "coloring_speed", DOUBLE, 1.0, 0.0, 1e15)
    c.add(p + "coloring_palette_offset", DOUBLE, 0.0, 0.0, 256.0)
    c.add(p + "fractal_coloring_algorithm", INT, 0, 0, 4)
    c.add(p + "fractal_coloring_color_preV215_false", BOOL, False)
    c.add(p + "fractal_coloring_color_4D_enabled_false", BOOL, False)
    c.add(p + "fractal_coloring_extra_color_enabled_false", BOOL, False)
    c.add(p + "fractal_coloring_extra_color_options_false", BOOL, False)
    c.add(p + "fractal_coloring_sphere_radius", DOUBLE, 1.0, 0.0, 1e20)
    c.add(p + "fractal_coloring_line_direction", VEC3, (1.0, 0.0, 0.0))
    for g in ("specular", "diffuse", "luminosity", "roughness", "reflectance", "transparency"):
        c.add(p + g + "_gradient_enable", BOOL, False)
    c.add(p + "specular", DOUBLE, 5.0, 0.0, 1e15)
    c.add(p + "specular_width", DOUBLE, 0.05, 1e-10, 1e15)
    c.add(p + "specular_metallic", DOUBLE, 2.0, 0.0, 1e15)
    c.add(p + "specular_metallic_width", DOUBLE, 1.0, 1e-10, 1e15)
    c.add(p + "specular_metallic_roughness", DOUBLE, 0.1, 0.0, 1e4)
    c.add(p + "specular_plastic_enable", BOOL, True)
    c.add(p + "specular_color", RGB, (65535, 65535, 65535))
    c.add(p + "luminosity_emissive", DOUBLE, 1.0, 0.0, 1e15)
    c.add(p + "metallic", BOOL, True)
    c.add(p + "reflectance", DOUBLE, 0.0, 0.0, 1e15)
    c.add(p + "transparency_of_surface", DOUBLE, 0.0, 0.0, 1.0)
    for flag in ("iridescence_enabled", "rough_surface", "subsurface_scattering",
                 "perlin_noise_enable", "use_color_texture",
                 "use_diffusion_texture", "use_displacement_texture", "use_luminosity_texture",
                 "use_normal_map_texture", "use_reflectance_texture", "use_transparency_texture"):
        c.add(p + flag, BOOL, False)


DEFAULT_GRADIENT = ("0 fd6029 1000 698403 2000 fff59b 3000 f5bd22 4000 0b5e87 5000 c68876 "
                    "6000 a51c64 7000 3b9fee 8000 d4ffd4 9000 aba53c")


def gradient_stops(text, limit=64):
    """cColorGradient::SetColorsFromString then SortGradient (color_gradient.cpp:271-354): a list
    of (position, (r, g, b)) with 8-bit channels, sorted by position (stable)."""
    import numpy as np
    parts = text.split(" ")
    if len(parts) < 2:
        stops = [(0.0, (255, 255, 255)), (1.0, (255, 255, 255))]
    else:
        stops = []
        position = 0.0
        for i, part in enumerate(parts):
            if not part:
                continue
            if i % 2 == 0:
                position = float(np.float32(_qtoint(part)) / np.float32(10000.0))
            else:
                h = _qtoint(part, 16)
                stops.append((position, (h // 65536, (h // 256) % 256, h % 256)))
                if i == 1:
                    stops.append((1.0, stops[0][1]))
    stops = sorted(stops, key=lambda s: s[0])
    return stops[:limit]


def init_light(c, n):
    """InitLightParams (initparameters.cpp:1841-1960), the parameters read here."""
    p = "light%d_" % n
    c.add(p + "is_defined", BOOL, False)
    c.add(p + "enabled", BOOL, n == 1)
    c.add(p + "cast_shadows", BOOL, True)
    c.add(p + "penetrating", BOOL, True)
    c.add(p + "relative_position", BOOL, n == 1)
    c.add(p + "volumetric", BOOL, False)
    c.add(p + "intensity", DOUBLE, 1.0 if n == 1 else 0.0, 0.0, 1e10)  # only light 1 is read
    c.add(p + "visibility", DOUBLE, 1.0, 0.0, 1e10)
    c.add(p + "size", DOUBLE, 0.5, 0.0, 1e10)
    c.add(p + "soft_shadow_cone", DOUBLE, 1.0, 0.0, 1e10)
    c.add(p + "contour_sharpness", DOUBLE, 1.0, 0.0, 1e10)
    c.add(p + "rotation", VEC3, (-45.0, 45.0, 0.0) if n == 1 else (0.0, 0.0, 0.0))
    c.add(p + "alpha", DOUBLE, -45.0 if n == 1 else 0.0)
    c.add(p + "beta", DOUBLE, 45.0 if n == 1 else 0.0)
    c.add(p + "color", RGB, (65535, 65535, 65535))
    c.add(p + "type", INT, 0 if n == 1 else 1, enum=["directional", "point", "spot",
                                                     "projection", "beam"])


def init_node(c, node_id):
    """InitNodeParams (initparameters.cpp:2102-2125)."""
    p = "node_%04d_" % node_id
    c.add(p + "definition", STRING, "mandelbulb 1,1,1,0,1")
    c.add(p + "position", VEC3, (0.0, 0.0, 0.0))
    c.add(p + "rotation", VEC3, (0.0, 0.0, 0.0))
    c.add(p + "scale", DOUBLE, 1.0)
    c.add(p + "repeat", VEC3, (0.0, 0.0, 0.0))
    c.add(p + "material", INT, 1)
    c.add(p + "enabled", BOOL, True)
    c.add(p + "julia_mode", BOOL, False)
    c.add(p + "julia_c", VEC3, (0.0, 0.0, 0.0))
    c.add(p + "fractal_constant_factor", VEC3, (1.0, 1.0, 1.0))
    c.add(p + "initial_waxis", DOUBLE, 0.0)
    c.add(p + "formula_maxiter", INT, 250, 1, 9999)


# ------------------------------------------------------------------------------------------
# Compatibility (settings.cpp:1262-1492)
# ------------------------------------------------------------------------------------------
_MAT_071 = ["shading", "specular", "reflect", "transparency_of_surface", "transparency_of_interior",
            "transparency_index_of_refraction", "transparency_interior_color",
            "fresnel_reflectance", "coloring_random_seed", "coloring_saturation", "coloring_speed",
            "coloring_palette_size", "coloring_palette_offset", "fractal_color",
            "surface_color_palette", "fractal_coloring_algorithm",
            "fractal_coloring_sphere_radius", "fractal_coloring_line_direction"]
_MAT_071_TO = {"reflect": "mat1_reflectance", "fractal_color": "mat1_use_colors_from_palette"}

_LIGHT_225 = [("main_light_intensity", "light1_intensity"),
              ("main_light_visibility_size", "light1_size"),
              ("main_light_contour_sharpness", "light1_contour_sharpness"),
              ("main_light_alpha", "light1_alpha"), ("main_light_beta", "light1_beta"),
              ("main_light_colour", "light1_color"), ("penetrating_lights", "light1_penetrating"),
              ("shadows_enabled", "light1_cast_shadows"),
              ("shadows_cone_angle", "light1_soft_shadow_cone"),
              ("main_light_enable", "light1_enabled"),
              ("main_light_position_relative", "light1_relative_position"),
              ("main_light_volumetric_intensity", "light1_volumetric_visibility"),
              ("main_light_volumetric_enabled", "light1_volumetric")]


def compatibility(ver, name, value):
    if ver <= 2.01:                                                          # :1265-1283
        if "aux_light_predefined" in name:
            name = name.replace("aux_light_predefined", "aux_light")
        if name == "volumetric_light_intensity_0":
            name = "main_light_volumetric_intensity"
        elif name == "volumetric_light_enabled_0":
            name = "main_light_volumetric_enabled"
        elif "volumetric_light" in name:
            name = name.replace("volumetric_light", "aux_light_volumetric")
    if ver <= 2.04 and name == "fractal_constant_factor":                   # :1286-1293
        value = " ".join([value] * 3)
    if ver <= 2.06 and name == "linear_DE_mode":                            # :1296-1303
        name, value = "delta_DE_function", "1"
    if ver <= 2.071 and name in _MAT_071:                                   # :1306-1341
        name = _MAT_071_TO.get(name, "mat1_" + name)
    if ver < 2.09 and name == "delta_DE_function":                          # :1344-1357
        value = {"0": "2", "2": "0"}.get(value, value)
    if ver < 2.12:                                                           # :1360-1363
        name = name.replace("gpu_", "opencl_")
    if ver < 2.19:                                                           # :1375-1403
        if "surface_color_palette" in name:
            name = name.replace("surface_color_palette", "surface_color_gradient")
            parts = value.split(" ")
            count = len(parts)
            if parts and len(parts[-1]) < 6:
                count -= 1
            step = 1.0 / count
            value = " ".join("%d %s" % (int(i * step * 10000.0), parts[i]) for i in range(count))
        if "luminosity_color_thesame" in name:
            name = name.replace("luminosity_color_thesame", "luminosity_gradient_enable")
        if any(k in name for k in ("coloring_palette_size", "coloring_random_seed",
                                   "coloring_saturation")):
            name = "skip"
    if ver < 2.25:                                                           # :1406-1477
        if name == "main_light_visibility":
            name = "light1_visibility"
        for old, new in _LIGHT_225:
            if old in name:
                name = name.replace(old, new)
        if "aux_light" in name:
            if "visibility_size" in name:
                name = name.replace("aux_light_visibility_size", "light2_size")
            elif "visibility" in name:
                name = name.replace("aux_light_visibility", "light2_visibility")
            elif "aux_light_place_behind" not in name:
                parts = name.split("_")
                if len(parts) >= 4:
                    prefix = "light%d" % (_qtoint(parts[-1]) + 1)
                    what = parts[2]
                    if what in ("intensity", "position", "enabled"):
                        name = "%s_%s" % (prefix, what)
                    elif what == "colour":
                        name = prefix + "_color"
                    elif what == "volumetric" and len(parts) >= 5:
                        if parts[3] == "intensity":
                            name = prefix + "_volumetric_visibility"
                        elif parts[3] == "enabled":
                            name = prefix + "_volumetric"
    return name, value


# ------------------------------------------------------------------------------------------
# the loader
# ------------------------------------------------------------------------------------------
class FractError(Exception):
    pass


FLAGS = {
    "hybrid": 0x1, "primitives": 0x2, "transform": 0x4, "palette": 0x8, "specular": 0x10,
    "glow": 0x20, "fog": 0x40, "reflections": 0x80, "lights": 0x100, "textures": 0x200,
    "ao_screen_space": 0x400, "ao_lightmap": 0x800, "postprocess": 0x1000, "mc_adaptive": 0x2000,
    "perspective": 0x4000, "formula": 0x8000, "animation": 0x10000, "material": 0x20000,
    "interior": 0x40000, "disabled": 0x80000, "antialiasing": 0x100000, "mc_effects": 0x200000,
    "nebula": 0x400000,
}

_LEGACY_MAIN_INDEXED = ("julia_mode", "julia_c", "fractal_constant_factor", "formula_maxiter",
                        "initial_waxis")


class _LegacyMain(dict):
    """The name_N temporaries of a pre-2.35 file, recording which ones the migration reads."""

    def __init__(self):
        super().__init__()
        self.read = set()

    def get(self, name, default=None):
        self.read.add(name)
        return dict.get(self, name, default)

    def __contains__(self, name):
        self.read.add(name)
        return dict.__contains__(self, name)


def load(text, trace=None):
    """Returns (fields, info): fields maps mbx_params field paths to values, info the header and
    the effect flags.  Raises FractError where mbx_load_fract returns MBX_ERR_INVALID.

    With a dict as `trace`, also records where every parameter line went: trace["lines"] holds
    (section, name as written, target) with target one of ("main", name), ("fractal", index,
    name), ("legacy", name), ("skip", name) or ("dropped", name) for a name no container
    registers; trace["containers"] holds the main container, the nine fractal containers and
    the legacy temporaries, whose `read` sets name what the reader looked up; trace["node"] is
    the fractal node the scene resolved to and trace["boolean"] whether the file is a pre-2.35
    boolean scene."""
    lines = text.strip().replace("\r\n", "\n").replace("\r", "\n").split("\n")
    # DecodeHeader (settings.cpp:540-601)
    if len(lines) < 3:
        raise FractError("no header")
    if "Mandelbulber settings file" not in lines[0]:
        raise FractError("wrong header")
    pos = lines[1].rfind("version")
    if pos < 0:
        raise FractError("no version")
    try:
        ver = float(lines[1][pos + 7:].strip())
    except ValueError:
        raise FractError("wrong version number")
    if ver > 2.35:
        raise FractError("newer version")
    if "all parameters" in lines[2]:
        fmt = 0
    elif "only modified parameters" in lines[2]:
        fmt = 1
    elif "application settings" in lines[2]:
        raise FractError("application settings")
    else:
        raise FractError("format not specified")

    par = init_params()
    fract = [init_fractal(i) for i in range(9)]
    if trace is not None:
        trace["lines"] = []
    if ver < 2.35:
        # temporaries the reference injects for pre-2.35 files (settings.cpp:2128-2214)
        par.add("boolean_operators", BOOL, False)
        par.add("formula_material_id", INT, 1)
        for f in fract:
            f.add("formula_position", VEC3, (0.0, 0.0, 0.0))
            f.add("formula_rotation", VEC3, (0.0, 0.0, 0.0))
            f.add("formula_repeat", VEC3, (0.0, 0.0, 0.0))
            f.add("formula_scale", DOUBLE, 1.0)
    materials, lights, nodes = set(), set(), set()
    legacy_main = _LegacyMain()   # name_N temporaries of settings.cpp:2128-2172 that a file set
    primitives = animation = False
    n_lines = n_errors = 0

    # node ids defined in the text (settings.cpp:698-726)
    for ln in lines:
        t = ln.strip()
        if t.startswith("node_") and "_definition" in t:
            parts = t.split("_")
            if len(parts) >= 3 and _qtoint(parts[1]) > 0:
                nodes.add(_qtoint(parts[1]))
    for nid in (sorted(nodes) or [1]):
        init_node(par, nid)

    def note(target):
        if trace is not None:
            trace["lines"].append((section, raw_name, target))

    def decode(container, line, main):
        nonlocal n_lines, primitives, raw_name
        line = line.strip()
        if not line:
            return True
        sp, sc = line.find(" "), line.find(";")
        if sp < 0 or sc < 0 or sc <= sp:
            return False
        name = line[:sp].strip()
        value = line[sp + 1:sc].strip()
        for q in ('"', "'"):
            if len(name) >= 2 and name[0] == q and name[-1] == q:
                name = name[1:-1].strip()
            if len(value) >= 2 and value[0] == q and value[-1] == q:
                value = value[1:-1].strip()
        n_lines += 1
        raw_name = name
        if main and ver < 2.35:
            # TryResolveLegacyFractalParam (settings.cpp:2268-2322)
            us = name.rfind("_")
            if us > 0:
                tail = name[us + 1:]
                base = name[:us]
                if tail.strip().lstrip("+-").isdigit() and _qtoint(tail) >= 1:
                    k = _qtoint(tail)
                    temp = base in _LEGACY_MAIN_INDEXED or (base == "dont_add_c_constant" and k <= 8)
                    if temp:
                        legacy_main[name] = value
                        note(("legacy", name))
                        return True
                    if base != "formula_maxiter" and k <= 9 and base in fract[k - 1]:
                        n2, v2 = compatibility(ver, base, value)
                        if n2 != "skip" and n2 in fract[k - 1]:
                            fract[k - 1][n2].set_text(v2)
                            note(("fractal", k - 1, n2))
                        else:
                            note(("skip", n2))
                        return True
        name, value = compatibility(ver, name, value)
        if name == "skip":
            note(("skip", raw_name))
            return True
        if main:
            if name.startswith("primitive_"):
                primitives = True
            if name.startswith("mat") and "_" in name:
                idx = name[3:name.index("_")]
                if idx.isdigit() and int(idx) not in materials:
                    m = int(idx)
                    init_material(par, m)
                    materials.add(m)
                    if ver < 2.15:                       # PreCompatibilityMaterials :2107-2119
                        par.setv("mat%d_metallic" % m, False)
                        par.setv("mat%d_specular" % m, 1.0)
                        par.setv("mat%d_specular_width" % m, 1.0)
            if name.startswith("light") and "_" in name:
                idx = name[5:name.index("_")]
                if idx.isdigit() and int(idx) not in lights:
                    init_light(par, int(idx))
                    lights.add(int(idx))
        if name in container:
            container[name].set_text(value)
            note(("main", name) if main else
                 ("fractal", [id(f) for f in fract].index(id(container)), name))
        else:
            note(("dropped", name))
        return True

    section = ""
    raw_name = ""
    for l in range(3, len(lines)):
        line = lines[l]
        if len(line) >= 2 and line[0] == "[" and line[-1] == "]":
            section = line[1:-1]
            continue
        if section == "description" or line == "":
            continue
        ok = True
        if section == "main_parameters":
            ok = decode(par, line, True)
        elif "fractal" in section:
            idx = _qtoint(section[8:]) - 1
            if 0 <= idx < 9:
                ok = decode(fract[idx], line, False)
        elif section in ("frames", "keyframes"):
            animation = True
        if not ok:
            n_errors += 1
            if n_errors > 3:
                raise FractError("too many errors")

    if not materials:                                            # CheckIfMaterialsAreDefined
        init_material(par, 1)

    # ---- Compatibility2 (settings.cpp:1494-1754) ----
    if ver <= 2.06:
        if par.get("delta_DE_function") != 1:
            par.setv("delta_DE_function", 2)
        for f in fract[:4]:
            f.setv("IFS_rotation_enabled", True)
    if ver <= 2.12 and par.get("iteration_fog_enable"):
        par.setv("iteration_fog_brightness_boost", 100.0)
    if ver < 2.20 and par.get("delta_DE_method") == 0:
        par.setv("delta_DE_function", 0)
    if ver < 2.21:
        fov = par.get("fov")
        if fov == 53.13:
            fov = 1.0
        pt = par.get("perspective_type")
        deg = 0.0
        if pt == 0:
            deg = math.atan(fov / 2.0) * 360.0 / math.pi
        elif pt in (1, 3):
            deg = fov * 180.0
        elif pt == 2:
            deg = fov * 360.0
        par["fov"].value = par["fov"]._limit(deg)
    if ver < 2.22:
        for f in fract:
            if any(c != 0.0 for c in f.get("IFS_edge")):
                f.setv("IFS_edge_enabled", True)
    if ver < 2.25:
        if 1 in lights:
            par.setv("light1_rotation", (par.get("light1_alpha"), par.get("light1_beta"), 0.0))
        else:
            init_light(par, 1)
            lights.add(1)
    if ver < 2.28 and 1 in lights and par.get("iteration_fog_enable"):
        par.setv("light1_intensity",
                 par.get("light1_intensity") / par.get("iteration_fog_brightness_boost"))

    flags = 0
    # ---- the objects tree: one fractal node ----
    node = None
    if ver < 2.35:
        # MigrateLegacyParamsToFractal (settings.cpp:2324-2402), MigrateToObjectsTree (:2551-2760)
        for k in range(1, 9):
            if legacy_main.get("dont_add_c_constant_%d" % k) == "true":
                fract[k - 1].setv("dont_add_c_constant", True)
        if par.get("formula_material_id") not in (None, 1):
            for f in fract:
                f.setv("formula_material_id", par.get("formula_material_id"))
        for glob, per in (("fractal_rotation", "formula_rotation"),
                          ("fractal_position", "formula_position"), ("repeat", "formula_repeat")):
            if par.get(glob) != (0.0, 0.0, 0.0):
                for f in fract:
                    f[per] = Param(VEC3, par.get(glob))
        hybrid = par.get("hybrid_fractal_enable")
        boolean = par.get("boolean_operators")
        enabled = [i + 1 for i, f in enumerate(fract) if f.get("formula") != 0]

        def transform(f):
            return dict(rotation=f.get("formula_rotation", (0.0, 0.0, 0.0)),
                        position=f.get("formula_position", (0.0, 0.0, 0.0)),
                        repeat=f.get("formula_repeat", (0.0, 0.0, 0.0)),
                        scale=f.get("formula_scale", 1.0),
                        material=f.get("formula_material_id"))

        if boolean:
            if len(enabled) == 1:
                k = enabled[0]
                f = fract[k - 1]
                node = dict(object=k, **transform(f))
                lm = legacy_main
                node["julia_mode"] = lm.get("julia_mode_%d" % k) == "true"
                node["julia_c"] = _vec(lm.get("julia_c_%d" % k), (0.0, 0.0, 0.0))
                node["fractal_constant_factor"] = _vec(lm.get("fractal_constant_factor_%d" % k),
                                                       (1.0, 1.0, 1.0))
                node["initial_waxis"] = _qtodouble(lm["initial_waxis_%d" % k]) \
                    if "initial_waxis_%d" % k in lm else 0.0
                mi = _qtoint(lm["formula_maxiter_%d" % k]) if "formula_maxiter_%d" % k in lm else 250
                node["formula_maxiter"] = min(max(mi, 1), 9999)
                if ver < 2.29:
                    node["formula_maxiter"] = par.get("N")
                node["enabled"] = f.get("fractal_enable")
            else:
                flags |= FLAGS["disabled"] if not enabled else FLAGS["hybrid"]
        elif not hybrid:
            f = fract[0]
            node = dict(object=1, **transform(f))
            node.update(julia_mode=par.get("julia_mode"), julia_c=par.get("julia_c"),
                        fractal_constant_factor=par.get("fractal_constant_factor"),
                        initial_waxis=par.get("initial_waxis"), formula_maxiter=par.get("N"),
                        enabled=f.get("fractal_enable"))
        else:
            # the hybrid node (settings.cpp:2690-2760) over the containers whose formula is not
            # none; fractal_enable false takes a child out of the render list
            members = [k for k in enabled if fract[k - 1].get("fractal_enable")]
            if not members:
                flags |= FLAGS["disabled"]
            node = dict(object=members[0] if members else (enabled[0] if enabled else 1),
                        rotation=par.get("fractal_rotation"), position=par.get("fractal_position"),
                        repeat=par.get("repeat"), scale=1.0, material=1,
                        julia_mode=par.get("julia_mode"), julia_c=par.get("julia_c"),
                        fractal_constant_factor=par.get("fractal_constant_factor"),
                        initial_waxis=par.get("initial_waxis"), formula_maxiter=par.get("N"),
                        enabled=True, members=members)
        if node is None and enabled:
            node = _default_node(enabled[0])
        if primitives:
            flags |= FLAGS["primitives"]
    else:
        defs = []
        for nid in sorted(nodes) or [1]:
            p = "node_%04d_" % nid
            parts = par.get(p + "definition").split(",")
            typ = _qtoint(parts[2]) if len(parts) >= 5 else 0
            parent = _qtoint(parts[3]) if len(parts) >= 5 else 0
            obj = _qtoint(parts[4]) if len(parts) >= 5 else 0
            defs.append((nid, typ, obj, par.get(p + "enabled"), parent))
        frac = [d for d in defs if d[1] == 1]
        hyb = [d for d in defs if d[1] == 10]
        if any(d[1] == 2 for d in defs):
            flags |= FLAGS["primitives"]
        # one hybrid node at the root whose children are all the fractal nodes: a hybrid sequence
        hybrid_tree = (len(hyb) == 1 and hyb[0][4] == 0 and frac and
                       all(d[4] == hyb[0][0] for d in frac) and
                       not any(d[1] not in (1, 2, 10) for d in defs))
        if not hybrid_tree and (len(frac) > 1 or any(d[1] not in (1, 2) for d in defs)):
            flags |= FLAGS["hybrid"]
        pick = hyb[0] if hybrid_tree else next((d for d in frac if d[3]), frac[0] if frac else None)
        if pick is None:
            flags |= FLAGS["disabled"]
        else:
            members = [d[2] for d in frac if d[3]] if hybrid_tree else None
            p = "node_%04d_" % pick[0]
            node = dict(object=(members[0] if members else pick[2]), rotation=par.get(p + "rotation"),
                        position=par.get(p + "position"), repeat=par.get(p + "repeat"),
                        scale=par.get(p + "scale"), material=par.get(p + "material"),
                        julia_mode=par.get(p + "julia_mode"), julia_c=par.get(p + "julia_c"),
                        fractal_constant_factor=par.get(p + "fractal_constant_factor"),
                        initial_waxis=par.get(p + "initial_waxis"),
                        formula_maxiter=par.get(p + "formula_maxiter"), enabled=pick[3],
                        members=members)
            if hybrid_tree and not members:
                flags |= FLAGS["disabled"]
    if node is None:
        node = _default_node(1)
    if not node["enabled"]:
        flags |= FLAGS["disabled"]
    if node["position"] != (0.0, 0.0, 0.0) or node["repeat"] != (0.0, 0.0, 0.0) or \
            node["scale"] != 1.0 or par.get("limits_enabled"):
        flags |= FLAGS["transform"]

    # ---- parameter block ----
    out = {}
    g = par.get
    out["imageWidth"], out["imageHeight"] = g("image_width"), g("image_height")
    out["camera"], out["target"], out["topVector"] = g("camera"), g("target"), g("camera_top")
    out["perspectiveType"], out["fov"] = g("perspective_type"), g("fov")
    if g("stereo_enabled") or (g("perspective_type") != 0 and g("DOF_monte_carlo") and g("DOF_enabled")):
        flags |= FLAGS["perspective"]
    out["sweetSpotHAngle"] = g("sweet_spot_horizontal_angle")
    out["sweetSpotVAngle"] = g("sweet_spot_vertical_angle")
    out["minN"], out["detailLevel"], out["DEThresh"] = g("minN"), g("detail_level"), g("DE_thresh")
    out["smoothness"], out["DEFactor"] = g("smoothness"), g("DE_factor")
    out["iterThreshMode"] = int(g("iteration_threshold_mode"))
    out["viewDistanceMax"] = g("view_distance_max")
    out["interiorMode"] = int(g("interior_mode"))
    out["constantDEThreshold"] = int(g("constant_DE_threshold"))
    out["maxRaymarchingSteps"] = g("max_raymarching_steps")
    out["advancedQuality"] = int(g("advanced_quality"))
    for k, v in (("absMaxMarchingStep", "abs_max_marching_step"), ("absMinMarchingStep", "abs_min_marching_step"),
                 ("relMaxMarchingStep", "rel_max_marching_step"), ("relMinMarchingStep", "rel_min_marching_step"),
                 ("detailSizeMax", "detail_size_max"), ("detailSizeMin", "detail_size_min"),
                 ("deltaDERelativeDelta", "deltade_relative_delta")):
        out[k] = g(v)
    # sImageAdjustments (fractparams.cpp:155-159): floats
    for k in ("brightness", "contrast", "gamma", "saturation"):
        out["image." + k] = struct.unpack("f", struct.pack("f", g(k)))[0]
    out["image.hdr"] = int(g("hdr"))
    # sCommonParams::foldings (fractparams.cpp:346-351), slow shading (fractparams.cpp:218)
    out["boxFolding"] = int(g("box_folding"))
    out["boxFoldingLimit"] = g("box_folding_limit")
    out["boxFoldingValue"] = g("box_folding_value")
    out["sphericalFolding"] = int(g("spherical_folding"))
    out["sphericalFoldingOuter"] = g("spherical_folding_outer")
    out["sphericalFoldingInner"] = g("spherical_folding_inner")
    out["slowShading"] = int(g("slow_shading"))
    out["ambientOcclusion"] = g("ambient_occlusion")
    out["ambientOcclusionQuality"] = g("ambient_occlusion_quality")
    out["ambientOcclusionFastTune"] = g("ambient_occlusion_fast_tune")
    out["ambientOcclusionEnabled"] = int(g("ambient_occlusion_enabled"))
    out["ambientOcclusionMode"] = g("ambient_occlusion_mode")
    out["ambientOcclusionColor"] = _rgbf(g("ambient_occlusion_color"))
    out["SSAORandomMode"] = int(g("SSAO_random_mode"))
    # screen-space AO (mode 2) is rendered by mbx_render_ssao after the render
    if g("ambient_occlusion_enabled") and g("ambient_occlusion_mode") == 1:
        flags |= FLAGS["ao_lightmap"]
    out["backgroundBrightness"] = g("background_brightness")
    out["backgroundGamma"] = g("background_gamma")
    for i in (1, 2, 3):
        out["backgroundColor%d" % i] = _rgbf(g("background_color_%d" % i))
    out["background3ColorsEnable"] = int(g("background_3_colors_enable"))
    out["reflectionsMax"] = g("reflections_max")
    out["fillLightColor"] = _rgbf(g("fill_light_color"))
    out["DOFMonteCarlo"] = int(g("DOF_monte_carlo"))
    out["DOFSamples"] = g("DOF_samples")
    out["DOFMonteCarloGlobalIllumination"] = int(g("DOF_MC_global_illumination"))
    out["monteCarloGIRadianceLimit"] = g("MC_GI_radiance_limit")
    out["DOFMinSamples"] = g("DOF_min_samples")         # the adaptive stop (render_worker.cpp:436-449)
    out["DOFMaxNoise"] = g("DOF_max_noise")
    out["aoLightMapRotation"] = g("ao_light_map_rotation")
    # Monte Carlo soft shadows jitter the light vector (shader_aux_shadow.cpp:76-87);
    # the Monte Carlo lens, chromatic aberration and soft shadows (fractparams.cpp:116-130, 180)
    out["DOFEnabled"] = int(g("DOF_enabled"))
    out["DOFFocus"] = g("DOF_focus")
    out["DOFRadius"] = g("DOF_radius")
    out["DOFMonteCarloChromaticAberration"] = int(g("DOF_MC_CA_enable"))
    out["DOFMonteCarloCACameraDispersion"] = g("DOF_MC_CA_camera_dispersion")
    out["monteCarloSoftShadows"] = int(g("MC_soft_shadows_enable"))
    if g("nebula_mode"):                    # the OpenCL nebula engine renders these (render_job.cpp:664-673)
        flags |= FLAGS["nebula"]
    out["N"] = g("N")                       # params->N: slow shading's pseudo-distance (fractparams.cpp:189)

    m = node["material"]
    mp = "mat%d_" % m
    if mp + "surface_color" not in par:
        init_material(par, m)
    out["surfaceColor"] = _rgbf(g(mp + "surface_color"))
    out["shading"], out["luminosity"] = g(mp + "shading"), g(mp + "luminosity")
    out["luminosityColor"] = _rgbf(g(mp + "luminosity_color"))
    out["luminosityEmissive"] = g(mp + "luminosity_emissive")
    out["useColorsFromPalette"] = int(g(mp + "use_colors_from_palette"))
    out["surfaceGradientEnable"] = int(g(mp + "surface_gradient_enable"))
    out["coloringAlgorithm"] = g(mp + "fractal_coloring_algorithm")
    out["coloringPreV215"] = int(g(mp + "fractal_coloring_color_preV215_false"))
    out["coloring4D"] = int(g(mp + "fractal_coloring_color_4D_enabled_false"))
    speed, offset = g(mp + "coloring_speed"), g(mp + "coloring_palette_offset")
    gtext = g(mp + "surface_color_gradient")
    if ver < 2.19:                                                   # Compatibility2 :1563-1593
        size = len(gtext.split(" ")) // 2
        offset = par[mp + "coloring_palette_offset"]._limit(offset / size)
        speed = par[mp + "coloring_speed"]._limit(speed * 10.0 / size)
    out["coloringSpeed"], out["coloringPaletteOffset"] = speed, offset
    out["coloringSphereRadius"] = g(mp + "fractal_coloring_sphere_radius")
    out["coloringLineDirection"] = g(mp + "fractal_coloring_line_direction")
    stops = gradient_stops(gtext)
    out["surfaceGradientCount"] = len(stops)
    out["surfaceGradientPosition"] = [s[0] for s in stops] + [0.0] * (64 - len(stops))
    out["surfaceGradientColor"] = [[float(c) for c in s[1]] for s in stops] + \
        [[0.0, 0.0, 0.0]] * (64 - len(stops))
    if g(mp + "use_colors_from_palette") and (
            g(mp + "fractal_coloring_extra_color_enabled_false") or
            any(g(mp + k + "_gradient_enable") for k in ("specular", "diffuse", "luminosity",
                                                         "roughness", "reflectance", "transparency"))):
        flags |= FLAGS["palette"]
    out["specularPlasticEnable"] = int(g(mp + "specular_plastic_enable"))
    out["metallic"] = int(g(mp + "metallic"))
    out["specular"] = g(mp + "specular")
    out["specularWidth"] = g(mp + "specular_width")
    out["specularMetallic"] = g(mp + "specular_metallic")
    out["specularMetallicWidth"] = g(mp + "specular_metallic_width")
    out["specularMetallicRoughness"] = g(mp + "specular_metallic_roughness")
    out["specularColor"] = _rgbf(g(mp + "specular_color"))
    if g("raytraced_reflections") and (g(mp + "reflectance") > 0.0 or
                                       g(mp + "transparency_of_surface") > 0.0):
        flags |= FLAGS["reflections"]
    if any(g(mp + k) for k in ("iridescence_enabled", "rough_surface", "subsurface_scattering",
                                "perlin_noise_enable")):
        flags |= FLAGS["material"]
    if any(g(mp + k) for k in ("use_color_texture", "use_diffusion_texture",
                                "use_displacement_texture", "use_luminosity_texture",
                                "use_normal_map_texture", "use_reflectance_texture",
                                "use_transparency_texture")):
        flags |= FLAGS["textures"]

    L = {}
    if 1 in lights:
        # cLight (light.cpp:77-87): intensity, visibility and size scale by the all_lights_ factors
        L = dict(enabled=int(g("light1_enabled")), castShadows=int(g("light1_cast_shadows")),
                 penetrating=int(g("light1_penetrating")),
                 relativePosition=int(g("light1_relative_position")), rotation=g("light1_rotation"),
                 intensity=g("light1_intensity") * g("all_lights_intensity"),
                 visibility=g("light1_visibility") * g("all_lights_visibility"),
                 size=g("light1_size") * g("all_lights_size"),
                 softShadowCone=g("light1_soft_shadow_cone"),
                 contourSharpness=g("light1_contour_sharpness"), color=_rgbf(g("light1_color")))
        if g("light1_enabled") and (g("light1_type") != 0 or g("light1_volumetric")):
            flags |= FLAGS["lights"]
    else:
        L = dict(enabled=0)
    for n in lights:
        if n >= 2 and g("light%d_enabled" % n):
            flags |= FLAGS["lights"]
    if g("random_lights_group") or g("fake_lights_enabled"):
        flags |= FLAGS["lights"]
    for k, v in L.items():
        out["light." + k] = v

    out["glowEnabled"] = int(g("glow_enabled"))
    out["glowIntensity"] = g("glow_intensity")
    out["glowColor1"] = _rgbf(g("glow_color_1"))
    out["glowColor2"] = _rgbf(g("glow_color_2"))
    if g("glow_enabled") and g("glow_intensity") > 0.0 and (
            g("volumetric_light_DE_Factor") != 1.0 or
            (g("DOF_monte_carlo") and g("DOF_MC_global_illumination") and
             g("MC_global_illumination_volumetric"))):
        flags |= FLAGS["glow"]
    # basic and distance fog are rendered (lit by (1, 1, 1)); iteration fog, clouds, shadows in
    # fog and the volumetric shader's own march (a DE factor other than 1) are not
    out["basicFogEnabled"] = int(g("basic_fog_enabled"))
    out["basicFogVisibility"] = g("basic_fog_visibility")
    out["basicFogColor"] = _rgbf(g("basic_fog_color"))
    out["basicFogCastShadows"] = int(g("basic_fog_cast_shadows"))
    out["volFogEnabled"] = int(g("volumetric_fog_enabled"))
    out["volFogDensity"] = g("volumetric_fog_density")
    out["volFogColour1Distance"] = g("volumetric_fog_colour_1_distance")
    out["volFogColour2Distance"] = g("volumetric_fog_colour_2_distance")
    out["volFogDistanceFactor"] = g("volumetric_fog_distance_factor")
    out["volFogDistanceFromSurface"] = g("volumetric_fog_distance_from_surface")
    out["volFogColour"] = [_rgbf(g("fog_color_%d" % k)) for k in (1, 2, 3)]
    out["distanceFogShadows"] = int(g("distance_fog_shadows"))
    out["iterFogEnabled"] = int(g("iteration_fog_enable"))
    out["cloudsEnabled"] = int(g("clouds_enable"))
    out["volumetricLightDEFactor"] = g("volumetric_light_DE_Factor")
    fog_on = g("basic_fog_enabled") or g("volumetric_fog_enabled")
    if g("iteration_fog_enable") or g("clouds_enable") or \
            (g("basic_fog_enabled") and g("basic_fog_cast_shadows")) or \
            (g("volumetric_fog_enabled") and g("distance_fog_shadows")) or \
            (fog_on and g("volumetric_light_DE_Factor") != 1.0):
        flags |= FLAGS["fog"]
    if g("textured_background") or g("env_mapping_enable"):
        flags |= FLAGS["textures"]
    # DOF without Monte Carlo is the post-processing blur; with it, the lens of the pixel loop
    if g("hdr_blur_enabled") or (g("DOF_enabled") and not g("DOF_monte_carlo")) or \
            g("post_chromatic_aberration_enabled"):
        flags |= FLAGS["postprocess"]
    if animation:
        flags |= FLAGS["animation"]
    out["antialiasingEnabled"] = int(g("antialiasing_enabled"))
    out["antialiasingSize"] = g("antialiasing_size")
    lens_fx = (g("DOF_monte_carlo") and (g("DOF_enabled") or g("DOF_MC_CA_enable"))) or \
        g("basic_fog_enabled") or g("volumetric_fog_enabled")
    if lens_fx and g("perspective_type") != 0:
        flags |= FLAGS["perspective"]
    if lens_fx and g("antialiasing_enabled"):
        flags |= FLAGS["antialiasing"]

    TABLE = (2, 3, 4, 5, 7, 8, 10, 11, 16, 19)  # the formulas the device path renders

    def container(prefix, f):
        """sFractal of one container (fractal.cpp:41-120) as mbx_fractal fields"""
        nonlocal flags
        fo_ = f.get("formula")
        out[prefix + "formula"] = fo_
        out[prefix + "formulaIterations"] = f.get("formula_iterations")
        out[prefix + "formulaWeight"] = f.get("formula_weight")
        out[prefix + "formulaStartIteration"] = f.get("formula_start_iteration")
        out[prefix + "formulaStopIteration"] = f.get("formula_stop_iteration")
        out[prefix + "checkForBailout"] = int(f.get("check_for_bailout"))
        out[prefix + "dontAddCConstant"] = int(f.get("dont_add_c_constant"))
        out[prefix + "bulb.power"] = f.get("power")
        out[prefix + "bulb.alphaAngleOffset"] = f.get("alpha_angle_offset")
        out[prefix + "bulb.betaAngleOffset"] = f.get("beta_angle_offset")
        mb = {"scale": "mandelbox_scale", "foldingLimit": "mandelbox_folding_limit",
              "foldingValue": "mandelbox_folding_value",
              "foldingSphericalMin": "mandelbox_folding_min_radius",
              "foldingSphericalFixed": "mandelbox_folding_fixed_radius",
              "offset": "mandelbox_offset", "rotationMain": "mandelbox_rotation_main",
              "colorFactor": "mandelbox_color", "colorFactorSp1": "mandelbox_color_Sp1",
              "colorFactorSp2": "mandelbox_color_Sp2"}
        for k, v in mb.items():
            out[prefix + "mandelbox." + k] = f.get(v)
        out[prefix + "mandelbox.rotationsEnabled"] = int(f.get("mandelbox_rotations_enabled"))
        out[prefix + "mandelbox.mainRotationEnabled"] = int(f.get("mandelbox_main_rotation_enabled"))
        out[prefix + "mandelbox.colorFactorR"] = f.get("mandelbox_color_R")
        out[prefix + "mandelbox.rotation"] = [[f.get("mandelbox_rotation_%s_%d" % (s_, i))
                                               for i in (1, 2, 3)] for s_ in ("neg", "pos")]
        out[prefix + "IFS.scale"] = f.get("IFS_scale")
        out[prefix + "IFS.offset"] = f.get("IFS_offset")
        out[prefix + "IFS.rotation"] = f.get("IFS_rotation")
        out[prefix + "IFS.edge"] = f.get("IFS_edge")
        out[prefix + "IFS.direction"] = [f.get("IFS_direction_%d" % i) for i in range(9)]
        out[prefix + "IFS.rotations"] = [f.get("IFS_rotations_%d" % i) for i in range(9)]
        out[prefix + "IFS.distance"] = [f.get("IFS_distance_%d" % i) for i in range(9)]
        out[prefix + "IFS.intensity"] = [f.get("IFS_intensity_%d" % i) for i in range(9)]
        out[prefix + "IFS.enabled"] = [int(f.get("IFS_enabled_%d" % i)) for i in range(9)]
        for k, v in (("absX", "IFS_abs_x"), ("absY", "IFS_abs_y"), ("absZ", "IFS_abs_z"),
                     ("mengerSpongeMode", "IFS_menger_sponge_mode"),
                     ("rotationEnabled", "IFS_rotation_enabled"), ("edgeEnabled", "IFS_edge_enabled")):
            out[prefix + "IFS." + k] = int(f.get(v))
        for k, v in (("scale3", "transf_scale_3"), ("aexionCadd", "cadd"),
                     ("foldingIntPowFoldFactor", "boxfold_bulbpow2_folding_factor"),
                     ("foldingIntPowZFactor", "boxfold_bulbpow2_z_factor"),
                     ("analyticDEScale1", "analyticDE_scale_1"), ("analyticDEOffset1", "analyticDE_offset_1"),
                     ("analyticDEOffset2", "analyticDE_offset_2")):
            out[prefix + "transform." + k] = f.get(v)
        out[prefix + "transform.analyticDEEnabled"] = int(f.get("analyticDE_enabled"))
        out[prefix + "transform.analyticDEEnabledFalse"] = int(f.get("analyticDE_enabled_false"))
        if fo_ != 0 and fo_ not in TABLE:
            flags |= FLAGS["formula"]
        if fo_ == 7 and g(mp + "use_colors_from_palette") and f.get("fold_color_aux_color_enabled_false"):
            flags |= FLAGS["palette"]
        return fo_

    f = fract[node["object"] - 1] if 1 <= node["object"] <= 9 else init_fractal(1)
    fo = container("fractal.", f)
    members = node.get("members")
    if members:
        out["hybridCount"] = len(members)
        for k, obj in enumerate(members[1:]):
            container("hybridFractals.%d." % k, fract[obj - 1])
        out["repeatFrom"] = g("repeat_from")
        if g(mp + "use_colors_from_palette") and g(mp + "fractal_coloring_extra_color_options_false"):
            flags |= FLAGS["palette"]
    else:
        out["hybridCount"] = 0
    out["linearDEOffset"] = g("linear_DE_offset")
    out["fractal.formulaMaxiter"] = node["formula_maxiter"]
    out["fractal.deltaDEMethod"] = g("delta_DE_method")
    out["fractal.deltaDEFunction"] = g("delta_DE_function")
    out["fractal.useDefaultBailout"] = int(g("use_default_bailout"))
    out["fractal.bailout"] = g("bailout")
    out["fractal.juliaMode"] = int(node["julia_mode"])
    out["fractal.juliaC"] = node["julia_c"]
    out["fractal.constantFactor"] = node["fractal_constant_factor"]
    out["fractal.initialWAxis"] = node["initial_waxis"]
    if fo == 0:
        flags |= FLAGS["disabled"]
    out["formulaRotation"] = node["rotation"]
    info = dict(fileVersion=ver, format=fmt, lines=n_lines, errors=n_errors, formula=fo,
                unsupported=flags)
    if trace is not None:
        trace.update(containers=(par, fract, legacy_main), node=node,
                     boolean=ver < 2.35 and bool(par.get("boolean_operators")))
    return out, info


def _default_node(obj):
    return dict(object=obj, rotation=(0.0, 0.0, 0.0), position=(0.0, 0.0, 0.0),
                repeat=(0.0, 0.0, 0.0), scale=1.0, material=1, julia_mode=False,
                julia_c=(0.0, 0.0, 0.0), fractal_constant_factor=(1.0, 1.0, 1.0),
                initial_waxis=0.0, formula_maxiter=250, enabled=True)


def _vec(text, default):
    if text is None:
        return default
    p = Param(VEC3, default)
    p.set_text(text)
    return p.value


def _rgbf(c):
    """toRGBFloat (color_structures.hpp:106-109): float(c / 65536)."""
    import numpy as np
    return tuple(float(np.float32(v / 65536.0)) for v in c)

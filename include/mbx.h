/*
 * mbx.h - C ABI of the MI355X (gfx950) distance-estimation raymarcher.
 *
 * The library (libmbx.so) replaces the per-pixel hot path of Mandelbulber2:
 *   RayMarching()        mandelbulber2/src/render_worker.cpp:708-951
 *   CalculateDistance()  mandelbulber2/src/calculate_distance.cpp:60-64, 533-763
 *   Compute<Mode>()      mandelbulber2/src/compute_fractal.cpp:67-654
 *   FormulaCode()        mandelbulber2/formula/definition/fractal_{mandelbulb,mandelbox,
 *                        kaleidoscopic_ifs}.cpp
 * plus the normal, Lambert light, hard-shadow, ambient-occlusion, background and Monte Carlo
 * GI shaders that RayRecursion() runs on each hit (render_worker.cpp:969-1819).
 *
 * Conventions: plain C linkage, POD structs, caller-owned buffers, int status codes
 * (0 = success, negative = error; the message is read with mbx_last_error()).  No C++
 * exception crosses this boundary.  One context per device stream; calls on one context are
 * serialised by the caller.
 *
 * Parameter names follow the reference's .fract keys and sParamRender / sFractal field names
 * (mandelbulber2/src/fractparams.cpp:80-356, mandelbulber2/src/fractal.cpp:41-120).
 */
#ifndef MBX_H_
#define MBX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBX_ABI_VERSION 9

/* array sizes as in mandelbulber2/src/fractal.h:52-54 */
#define MBX_IFS_VECTOR_COUNT 9
#define MBX_MANDELBOX_FOLDS 2
/* upper bound of explicitly supplied ambient-occlusion directions */
#define MBX_MAX_AO_VECTORS 1024 /* PrepareAOVectors keeps at most ~660 at the largest quality, 10 */
/* upper bound of the stops of a colour gradient (the first colour counts twice: at its position
 * and again at 1.0) */
#define MBX_MAX_GRADIENT_STOPS 64
/* fractal containers of a hybrid sequence (NUMBER_OF_FRACTALS, mandelbulber2/src/fractal.h:52) */
#define MBX_MAX_FRACTALS 9
/* most frames of one mbx_render_batch_noalloc call (the size of its static device scratch) */
#define MBX_NOALLOC_MAX_FRAMES 64

/* status codes */
#define MBX_OK 0
#define MBX_ERR_INVALID -1
#define MBX_ERR_UNSUPPORTED -2
#define MBX_ERR_HIP -3
#define MBX_ERR_NOMEM -4
#define MBX_ERR_IO -5

/* enumFractalFormula ids (mandelbulber2/formula/definition/all_fractal_list_enums.hpp:45-62) */
enum mbx_formula {
	MBX_FORMULA_NONE = 0,
	MBX_FORMULA_MANDELBULB = 2,
	MBX_FORMULA_MANDELBULB_POWER2 = 3,
	MBX_FORMULA_HYPERCOMPLEX = 4,
	MBX_FORMULA_QUATERNION = 5,
	MBX_FORMULA_MENGER_SPONGE = 7,
	MBX_FORMULA_MANDELBOX = 8,
	MBX_FORMULA_KALEIDOSCOPIC_IFS = 10,
	MBX_FORMULA_XENODREAMBUIE = 11,
	MBX_FORMULA_BOX_FOLD_BULB_POW2 = 16,
	MBX_FORMULA_AEXION = 19
};

/* enumDEType (mandelbulber2/src/fractal_enums.h:70-75) */
enum mbx_de_type { MBX_DE_ANALYTIC = 0, MBX_DE_DELTA = 1 };

/* enumDEFunctionType (fractal_enums.h:47-61): used by the delta-DE finish */
enum mbx_de_function { MBX_DE_FUNCTION_LINEAR = 1, MBX_DE_FUNCTION_LOGARITHMIC = 2 };

/* enumDEAnalyticFunction (fractal_enums.h:77-88): used by the analytic-DE finish */
enum mbx_analytic_function {
	MBX_ANALYTIC_LINEAR = 1,
	MBX_ANALYTIC_LOGARITHMIC = 2,
	MBX_ANALYTIC_IFS = 3
};

/* enumCPixelAddition (fractal_enums.h:62-68): whether the formula adds c by default */
enum mbx_cpixel_addition {
	MBX_CPIXEL_ENABLED_BY_DEFAULT = 0,
	MBX_CPIXEL_DISABLED_BY_DEFAULT = 1,
	MBX_CPIXEL_ALREADY_HAS = 2
};

/* enumColoringFunction (fractal_enums.h:90-99): the colour index of a single formula */
enum mbx_coloring_function {
	MBX_COLORING_DEFAULT = 0,
	MBX_COLORING_ABOX = 1,
	MBX_COLORING_IFS = 2,
	MBX_COLORING_AMAZING_SURF = 3,
	MBX_COLORING_DONUT = 4
};

/* enumAOMode (mandelbulber2/src/ao_modes.h:42-44); screen-space AO is a post effect over the
 * finished planes (mbx_render_ssao), the render itself takes no AO term in that mode */
enum mbx_ao_mode { MBX_AO_FAST = 0, MBX_AO_MULTIPLE_RAYS = 1, MBX_AO_SCREEN_SPACE = 2 };

/* RNG stream layout.  PER_PIXEL: one Park-Miller stream per pixel, seeded from
 * (pixel index, random_seed); identical on CPU and GPU.  GLOBAL_SEQUENTIAL: one stream for the
 * whole image in row-major order starting from seed 1, as the single-threaded reference renders
 * (mandelbulber2/src/common_math.cpp:47-72); CPU oracle only. */
enum mbx_rng_mode { MBX_RNG_PER_PIXEL = 0, MBX_RNG_GLOBAL_SEQUENTIAL = 1 };

/* sFractalMandelbulb (mandelbulber2/src/fractal.h:227-234); angles in degrees as in .fract */
typedef struct mbx_fractal_mandelbulb {
	double power;            /* "power" */
	double alphaAngleOffset; /* "alpha_angle_offset", degrees */
	double betaAngleOffset;  /* "beta_angle_offset", degrees */
} mbx_fractal_mandelbulb;

/* sFractalMandelbox (fractal.h:185-219); rotations in degrees as in .fract */
typedef struct mbx_fractal_mandelbox {
	double scale;                 /* "mandelbox_scale" */
	double foldingLimit;          /* "mandelbox_folding_limit" */
	double foldingValue;          /* "mandelbox_folding_value" */
	double foldingSphericalMin;   /* "mandelbox_folding_min_radius" */
	double foldingSphericalFixed; /* "mandelbox_folding_fixed_radius" */
	double offset[3];             /* "mandelbox_offset" */
	double rotationMain[3];       /* "mandelbox_rotation_main" */
	double rotation[MBX_MANDELBOX_FOLDS][3][3]; /* [0]="mandelbox_rotation_neg_i",
	                                              [1]="mandelbox_rotation_pos_i" */
	double colorFactor[3];        /* "mandelbox_color" */
	double colorFactorSp1;        /* "mandelbox_color_Sp1" */
	double colorFactorSp2;        /* "mandelbox_color_Sp2" */
	double colorFactorR;          /* "mandelbox_color_R" */
	int32_t rotationsEnabled;     /* "mandelbox_rotations_enabled" */
	int32_t mainRotationEnabled;  /* "mandelbox_main_rotation_enabled" */
} mbx_fractal_mandelbox;

/* sFractalIFS (fractal.h:155-174) */
typedef struct mbx_fractal_ifs {
	double scale;                              /* "IFS_scale" */
	double offset[3];                          /* "IFS_offset" */
	double rotation[3];                        /* "IFS_rotation", degrees */
	double edge[3];                            /* "IFS_edge" */
	double direction[MBX_IFS_VECTOR_COUNT][3]; /* "IFS_direction_i" (normalised on upload) */
	double rotations[MBX_IFS_VECTOR_COUNT][3]; /* "IFS_rotations_i", degrees */
	double distance[MBX_IFS_VECTOR_COUNT];     /* "IFS_distance_i" */
	double intensity[MBX_IFS_VECTOR_COUNT];    /* "IFS_intensity_i" */
	int32_t enabled[MBX_IFS_VECTOR_COUNT];     /* "IFS_enabled_i" */
	int32_t absX, absY, absZ;                  /* "IFS_abs_x/y/z" */
	int32_t mengerSpongeMode;                  /* "IFS_menger_sponge_mode" */
	int32_t rotationEnabled;                   /* "IFS_rotation_enabled" */
	int32_t edgeEnabled;                       /* "IFS_edge_enabled" */
} mbx_fractal_ifs;

/* Parameters of the other formulas on the path (sFractal members, fractal.cpp:93-363) */
typedef struct mbx_fractal_transform {
	double scale3;                /* "transf_scale_3" (Menger sponge scale) */
	double aexionCadd;            /* "cadd" */
	double foldingIntPowFoldFactor; /* "boxfold_bulbpow2_folding_factor" */
	double foldingIntPowZFactor;  /* "boxfold_bulbpow2_z_factor" */
	double analyticDEScale1;      /* "analyticDE_scale_1" */
	double analyticDEOffset1;     /* "analyticDE_offset_1" */
	double analyticDEOffset2;     /* "analyticDE_offset_2" */
	int32_t analyticDEEnabled;    /* "analyticDE_enabled" */
	int32_t analyticDEEnabledFalse; /* "analyticDE_enabled_false" */
} mbx_fractal_transform;

/* One fractal container (sFractal) with its hybrid-sequence fields, plus, in mbx_params.fractal,
 * the node data that cHybridFractalSequences::CollectSequenceData takes from the rendered node
 * (mandelbulber2/src/hybrid_fractal_sequences.cpp:181-287, 377-503): maxiter, Julia, constant
 * factor, initial w, bailout and DE mode.  In a hybrid's further containers
 * (mbx_params.hybridFractals) only the formula, its parameters and the sequence fields are read. */
typedef struct mbx_fractal {
	int32_t formula;            /* enum mbx_formula ("formula") */
	int32_t formulaMaxiter;     /* "N" / node formula_maxiter */
	int32_t deltaDEMethod;      /* "delta_DE_method": 0 preferred, 1 force delta, 2 force analytic */
	int32_t deltaDEFunction;    /* "delta_DE_function": 0 preferred, 1 linear, 2 logarithmic */
	int32_t useDefaultBailout;  /* "use_default_bailout" */
	int32_t juliaMode;          /* "julia_mode" */
	int32_t dontAddCConstant;   /* "dont_add_c_constant" */
	int32_t checkForBailout;    /* "check_for_bailout" (hybrid sequences; a single formula always checks) */
	double bailout;             /* "bailout" (used when useDefaultBailout == 0) */
	double juliaC[3];           /* "julia_c" */
	double constantFactor[3];   /* "fractal_constant_factor" */
	double initialWAxis;        /* "initial_waxis" */
	int32_t formulaIterations;  /* "formula_iterations" (hybrid sequences) */
	int32_t formulaStartIteration; /* "formula_start_iteration" */
	int32_t formulaStopIteration;  /* "formula_stop_iteration" */
	int32_t reserved0;
	double formulaWeight;       /* "formula_weight" (hybrid sequences) */
	mbx_fractal_mandelbulb bulb;
	mbx_fractal_mandelbox mandelbox;
	mbx_fractal_ifs IFS;
	mbx_fractal_transform transform;
} mbx_fractal;

/* Light #1 (cLight, mandelbulber2/src/light.cpp:51-195).  Only the directional type
 * is on the path. */
typedef struct mbx_light {
	int32_t enabled;          /* "light1_enabled" */
	int32_t castShadows;      /* "light1_cast_shadows" */
	int32_t penetrating;      /* "light1_penetrating" */
	int32_t relativePosition; /* "light1_relative_position" */
	double rotation[3];       /* "light1_rotation", degrees */
	double intensity;         /* "light1_intensity" */
	double visibility;        /* "light1_visibility" */
	double size;              /* "light1_size" */
	double softShadowCone;    /* "light1_soft_shadow_cone", degrees */
	double contourSharpness;  /* "light1_contour_sharpness" */
	double color[3];          /* "light1_color" as 0..1 */
} mbx_light;

/* sImageAdjustments (image_adjustments.h:38-47; fractparams.cpp:155-159): the post-processing
 * cImage::CalculatePixel applies when it turns the float image into 16-bit RGB
 * (cimage.cpp:317-361).  Defaults 1, 1, 1, 1, off (initparameters.cpp:208-212). */
typedef struct mbx_image_adjustments {
	float brightness;                 /* "brightness" */
	float contrast;                   /* "contrast" */
	float gamma;                      /* "gamma" */
	float saturation;                 /* "saturation" */
	int32_t hdr;                      /* "hdr": tanh tone mapping */
	int32_t reserved;
} mbx_image_adjustments;

/* Render parameters: the sParamRender subset on the path
 * (mandelbulber2/src/fractparams.cpp:80-356), with material #1 and light #1. */
typedef struct mbx_params {
	int32_t imageWidth;             /* "image_width" */
	int32_t imageHeight;            /* "image_height" */
	int32_t perspectiveType;        /* "perspective_type": 0 three-point, 1 fish-eye, 2 equirectangular,
	                                   3 fish-eye cut (projection_3d.cpp:41-89) */
	int32_t maxRaymarchingSteps;    /* "max_raymarching_steps" */
	int32_t minN;                   /* "minN" */
	int32_t iterThreshMode;         /* "iteration_threshold_mode" */
	int32_t interiorMode;           /* "interior_mode" */
	int32_t constantDEThreshold;    /* "constant_DE_threshold" */
	double camera[3];               /* "camera" */
	double target[3];               /* "target" */
	double topVector[3];            /* "camera_top" */
	double fov;                     /* "fov", degrees */
	double sweetSpotHAngle;         /* "sweet_spot_horizontal_angle", degrees */
	double sweetSpotVAngle;         /* "sweet_spot_vertical_angle", degrees */
	double detailLevel;             /* "detail_level" */
	double DEThresh;                /* "DE_thresh" */
	double DEFactor;                /* "DE_factor" */
	double smoothness;              /* "smoothness" */
	double viewDistanceMax;         /* "view_distance_max" */

	/* material #1 (cMaterial, mandelbulber2/src/material.cpp) */
	double surfaceColor[3];         /* "mat1_surface_color" as 0..1 */
	double shading;                 /* "mat1_shading" */
	double luminosity;              /* "mat1_luminosity" */
	double luminosityColor[3];      /* "mat1_luminosity_color" */
	double fillLightColor[3];       /* "fill_light_color" */

	mbx_light light;

	/* ambient occlusion (shader_ambient_occlusion.cpp, shader_fast_ambient_occlusion.cpp) */
	int32_t ambientOcclusionEnabled; /* "ambient_occlusion_enabled" */
	int32_t ambientOcclusionMode;    /* "ambient_occlusion_mode" */
	int32_t ambientOcclusionQuality; /* "ambient_occlusion_quality" */
	int32_t aoVectorsCount;          /* number of explicit AO directions (replaces the
	                                    lightmap-pruned set of PrepareAOVectors) */
	double ambientOcclusion;         /* "ambient_occlusion" */
	double ambientOcclusionColor[3]; /* "ambient_occlusion_color" */
	double ambientOcclusionFastTune; /* "ambient_occlusion_fast_tune" */
	double aoVectors[MBX_MAX_AO_VECTORS][3];
	double aoColors[MBX_MAX_AO_VECTORS][3];

	/* background (shader_background.cpp) */
	int32_t background3ColorsEnable;  /* "background_3_colors_enable" */
	int32_t reserved1;
	double backgroundColor1[3];       /* "background_color_1" */
	double backgroundColor2[3];       /* "background_color_2" */
	double backgroundColor3[3];       /* "background_color_3" */
	double backgroundBrightness;      /* "background_brightness" */
	double backgroundGamma;           /* "background_gamma" */

	/* Monte Carlo (render_worker.cpp:190-233, shader_global_illumination.cpp) */
	int32_t DOFMonteCarlo;                   /* "DOF_monte_carlo" */
	int32_t DOFSamples;                      /* "DOF_samples" */
	int32_t DOFMonteCarloGlobalIllumination; /* "DOF_MC_global_illumination" */
	int32_t reflectionsMax;                  /* "reflections_max" (GI bounce depth) */
	double monteCarloGIRadianceLimit;        /* "MC_GI_radiance_limit" */

	/* random stream */
	int32_t rngMode;     /* enum mbx_rng_mode */
	uint32_t randomSeed; /* per-frame seed mixed into the per-pixel streams */

	mbx_fractal fractal;

	/* rotation of the fractal node, degrees ("formula_rotation", the node's "rotation",
	 * objects_tree.cpp:292-382): points enter the formula as p_local = R p with
	 * R = Rz(z) Ry(y) Rx(x) (CRotationMatrix::SetRotation2, algebra.cpp:296-302) */
	double formulaRotation[3];

	/* palette surface colouring of material #1 (SurfaceColour, shader_surface_color.cpp:39-108;
	 * the colouring pass of Compute, compute_fractal.cpp:343-436; CalculateColorIndex,
	 * fractal_coloring.cpp:111-343; cColorGradient, color_gradient.cpp:138-360) */
	int32_t useColorsFromPalette;     /* "mat1_use_colors_from_palette" */
	int32_t surfaceGradientEnable;    /* "mat1_surface_gradient_enable" */
	int32_t coloringAlgorithm;        /* "mat1_fractal_coloring_algorithm": 0 standard |z|, 1 z.p,
	                                     2 sphere, 3 cross, 4 line */
	int32_t coloringPreV215;          /* "mat1_fractal_coloring_color_preV215_false" */
	int32_t coloring4D;               /* "mat1_fractal_coloring_color_4D_enabled_false" */
	int32_t surfaceGradientCount;     /* stops of "mat1_surface_color_gradient" as cColorGradient holds
	                                     them: sorted by position, the first colour repeated at 1.0
	                                     (2 .. MBX_MAX_GRADIENT_STOPS) */
	double coloringSpeed;             /* "mat1_coloring_speed" */
	double coloringPaletteOffset;     /* "mat1_coloring_palette_offset" */
	double coloringSphereRadius;      /* "mat1_fractal_coloring_sphere_radius" */
	double coloringLineDirection[3];  /* "mat1_fractal_coloring_line_direction" (x, y, z) */
	double surfaceGradientPosition[MBX_MAX_GRADIENT_STOPS]; /* 0 .. 1 */
	double surfaceGradientColor[MBX_MAX_GRADIENT_STOPS][3]; /* 8-bit channel values, 0 .. 255 */

	/* specular highlights of material #1 (SpecularHighlightCombined / SpecularHighlight,
	 * shader_specular_highlight_combined.cpp:39-68, shader_specular_highlight.cpp:37-73) */
	int32_t specularPlasticEnable;    /* "mat1_specular_plastic_enable" */
	int32_t metallic;                 /* "mat1_metallic" */
	double specular;                  /* "mat1_specular" (plastic intensity) */
	double specularWidth;             /* "mat1_specular_width" */
	double specularMetallic;          /* "mat1_specular_metallic" */
	double specularMetallicWidth;     /* "mat1_specular_metallic_width" */
	double specularMetallicRoughness; /* "mat1_specular_metallic_roughness" */
	double specularColor[3];          /* "mat1_specular_color" */
	double luminosityEmissive;        /* "mat1_luminosity_emissive" (emission seen by GI bounces) */

	/* glow (VolumetricShader, shader_volumetric.cpp:82-90, 196-204) */
	int32_t glowEnabled;              /* "glow_enabled" */
	int32_t reserved2;
	double glowIntensity;             /* "glow_intensity" */
	double glowColor1[3];             /* "glow_color_1" */
	double glowColor2[3];             /* "glow_color_2" */

	/* hybrid sequence (cHybridFractalSequences::CreateSequence, hybrid_fractal_sequences.cpp:290-377):
	 * hybridCount 0 renders `fractal` alone; 1 .. MBX_MAX_FRACTALS renders the hybrid of `fractal`
	 * and hybridFractals[0 .. hybridCount - 2], in that order.  A hybrid of one container is not
	 * the container alone: it takes the hybrid DE finish, bailout and colour index. */
	int32_t hybridCount;              /* fractal containers in the hybrid */
	int32_t repeatFrom;               /* "repeat_from" (1-based) */
	double linearDEOffset;            /* "linear_DE_offset" (linear DE finish of a hybrid) */
	mbx_fractal hybridFractals[MBX_MAX_FRACTALS - 1];

	/* "advanced quality" march limits: CalcDistThresh clamps the threshold to
	 * [detailSizeMin, detailSizeMax] (render_worker.cpp:671-675), RayMarching clamps each step
	 * (render_worker.cpp:817-834) and the delta DE offsets by deltaDERelativeDelta * detail
	 * (calculate_distance.cpp:614-617) */
	int32_t advancedQuality;          /* "advanced_quality" */
	int32_t reserved3;
	double absMaxMarchingStep;        /* "abs_max_marching_step" */
	double absMinMarchingStep;        /* "abs_min_marching_step" */
	double relMaxMarchingStep;        /* "rel_max_marching_step" */
	double relMinMarchingStep;        /* "rel_min_marching_step" */
	double detailSizeMax;             /* "detail_size_max" */
	double detailSizeMin;             /* "detail_size_min" */
	double deltaDERelativeDelta;      /* "deltade_relative_delta" */

	/* image adjustments for mbx_compile_image (not used by the render) */
	mbx_image_adjustments image;
	/* global foldings before every formula iteration of Compute (compute_fractal.cpp:192-202):
	 * BoxFolding and SphericalFolding (legacy_fractal_transforrms.cpp:43-101), read as
	 * sCommonParams::foldings (fractparams.cpp:346-351); either one takes the general kernel */
	int32_t boxFolding;             /* "box_folding" */
	int32_t sphericalFolding;       /* "spherical_folding" */
	double boxFoldingLimit;         /* "box_folding_limit" */
	double boxFoldingValue;         /* "box_folding_value" */
	double sphericalFoldingOuter;   /* "spherical_folding_outer" */
	double sphericalFoldingInner;   /* "spherical_folding_inner" */
	/* normals from the iteration counts of an 11 x 11 x 11 grid of orbits round the hit
	 * (CalculateNormals, shader_calculate_normals.cpp:81-109) instead of the DE's central
	 * differences */
	int32_t slowShading;            /* "slow_shading" */
	/* params->N (fractparams.cpp:189; "N", default 250): the 1 + N - iterations pseudo-distance
	 * of slow shading (shader_calculate_normals.cpp:103), apart from the node's maxiter */
	int32_t N;
	/* adaptive Monte Carlo (render_worker.cpp:436-449): after each sample past the first
	 * DOFMinSamples + 1, the pixel stops once MonteCarloDOFNoiseEstimation (:1871-1901) falls
	 * below DOFMaxNoise * 0.01 */
	int32_t DOFMinSamples;          /* "DOF_min_samples" (default 10) */
	int32_t reserved5;
	double DOFMaxNoise;             /* "DOF_max_noise" (default 1) */
	/* "ao_light_map_rotation" (degrees, a single angle: fractparams.cpp:62-64 reads it as the x
	 * of a vector, SetRotation turns it into a rotation about z) for mbx_prepare_ao_vectors */
	double aoLightMapRotation;
	/* Monte Carlo lens, chromatic aberration and soft shadows, each only with DOFMonteCarlo:
	 * MonteCarloDOF (render_worker.cpp:239-242, 1832-1844) moves the ray's origin on the lens and
	 * aims it at the focal plane; the aberration (:255-264, 374-380) bends the view vector by a
	 * random hue and tints the sample; soft shadows jitter the light vector inside the shadow
	 * cone (shader_aux_shadow.cpp:56-87) */
	int32_t DOFEnabled;                        /* "DOF_enabled" */
	int32_t DOFMonteCarloChromaticAberration;  /* "DOF_MC_CA_enable" */
	double DOFFocus;                           /* "DOF_focus" (default 6) */
	double DOFRadius;                          /* "DOF_radius" (default 10) */
	double DOFMonteCarloCACameraDispersion;    /* "DOF_MC_CA_camera_dispersion" (default 1; a float in
	                                              sParamRender) */
	int32_t monteCarloSoftShadows;             /* "MC_soft_shadows_enable" */
	/* screen-space AO (mbx_render_ssao; cSSAOWorker::doWork, ssao_worker.cpp:63-261): random
	 * angles and radii per pixel instead of the fixed angle set */
	int32_t SSAORandomMode;                    /* "SSAO_random_mode" */

	/* fog of VolumetricShader (shader_volumetric.cpp:70-700) along the primary ray's march steps:
	 * basic fog (opacity step / visibility per step, :228-241, 575-590) and the volumetric
	 * distance fog (DistanceFogOpacity, shader_distance_fog_opacity.cpp:37-52, colours :325-356,
	 * blend :556-571).  Iteration fog, clouds, shadows in fog and a volumetric DE factor other
	 * than 1 are not rendered (the render refuses them). */
	int32_t basicFogEnabled;                   /* "basic_fog_enabled" */
	int32_t basicFogCastShadows;               /* "basic_fog_cast_shadows" */
	double basicFogVisibility;                 /* "basic_fog_visibility" (default 20) */
	double basicFogColor[3];                   /* "basic_fog_color" */
	int32_t volFogEnabled;                     /* "volumetric_fog_enabled" */
	int32_t distanceFogShadows;                /* "distance_fog_shadows" */
	double volFogDensity;                      /* "volumetric_fog_density" (a float in sParamRender) */
	double volFogColour1Distance;              /* "volumetric_fog_colour_1_distance" */
	double volFogColour2Distance;              /* "volumetric_fog_colour_2_distance" */
	double volFogDistanceFactor;               /* "volumetric_fog_distance_factor" */
	double volFogDistanceFromSurface;          /* "volumetric_fog_distance_from_surface" */
	double volFogColour[3][3];                 /* "fog_color_1", "fog_color_2", "fog_color_3" */
	int32_t iterFogEnabled;                    /* "iteration_fog_enable" (not rendered) */
	int32_t cloudsEnabled;                     /* "clouds_enable" (not rendered) */
	double volumetricLightDEFactor;            /* "volumetric_light_DE_Factor" (default 1) */
	/* antialiasing (render_worker.cpp:103-104, 187-202): antialiasingSize^2 samples per pixel on a
	 * regular sub-pixel grid, averaged; with Monte Carlo samples the grid takes the place of the
	 * random sub-pixel jitter and the sample count is DOFSamples * antialiasingSize^2 */
	int32_t antialiasingEnabled;               /* "antialiasing_enabled" */
	int32_t antialiasingSize;                  /* "antialiasing_size" (1 .. 10, default 2) */
} mbx_params;

/* Per-pixel output planes (device pointers; any may be NULL except where noted).
 * Layout: row-major, pixel (x, y) at index y*imageWidth + x. */
typedef struct mbx_outputs {
	float *rgb;      /* 3 floats per pixel: the pre-post-processing image buffer */
	float *depth;    /* ray depth t; 1e20 on a miss (render_worker.cpp:353) */
	int32_t *iters;  /* iteration count of the last DE evaluation at the returned point */
	uint8_t *mask;   /* 1 = surface hit */
	float *normal;   /* 3 floats per pixel, world-space unit normal (0 on a miss) */
	int32_t *steps;  /* raymarching step count per pixel (coarse + binary search) */
	uint8_t *colour; /* 3 bytes per pixel: the surface colour of the hit as the reference's colour
	                    buffer holds it, uchar(c * 255) per sample and the integer mean over Monte
	                    Carlo samples (render_worker.cpp:402-404, 433-435, 470-472); 0 on a miss.
	                    The input of mbx_render_ssao */
	uint16_t *opacity; /* fog opacity, ushort(opacity * 65535) per sample and the integer mean
	                      (render_worker.cpp:400, 432, 469); 0 without fog */
	/* the image's optional channels (cImage optional buffers, render_worker.cpp:355-366, 405-420,
	 * 474-476, 502-516), 3 floats per pixel each, of the last Monte Carlo sample: */
	float *world;    /* the ray's end point (the hit, or where the march stopped on a miss) */
	float *specular; /* the specular highlight of the hit; 0 on a miss */
	float *shadow;   /* light 1's shadow factor (AuxShadow; 1 without cast shadows, 0 with the light
	                    off or facing away); 1 on a miss */
	float *gi;       /* the Monte Carlo GI term, the mean over the samples */
} mbx_outputs;

/* Work statistics of the last launch on a context (cStatistics analogue,
 * mandelbulber2/src/statistics.h:42-75). */
typedef struct mbx_stats {
	uint64_t rays;            /* primary rays (pixels x samples) */
	uint64_t de_evals;        /* CalculateDistance calls */
	uint64_t iterations;      /* formula iterations (totalNumberOfIterations) */
	double kernel_ms;         /* device time of the last render launch */
	uint64_t far_steps;       /* march steps taken without a DE evaluation: beyond the far-field
	                             radius the step is provably the clamped 3 (see MBX_TUNE_FAR_FIELD);
	                             de_evals + far_steps = the reference's CalculateDistance calls */
} mbx_stats;

typedef struct mbx_ctx mbx_ctx;

/* Library version (MBX_ABI_VERSION). */
int mbx_version(void);

/* sizeof(mbx_params) as compiled into the library, for binding-layout checks. */
size_t mbx_sizeof_params(void);

/* Create a context on HIP device `device`.  Replaces the OpenCL engine's
 * CreateKernel4Program / PreAllocateBuffers / CreateCommandQueue sequence
 * (mandelbulber2/src/render_job.cpp:916-927, opencl_engine.cpp:277-375).  `stream` is a
 * hipStream_t (NULL = default stream). */
int mbx_ctx_create(int device, void *stream, mbx_ctx **out);
void mbx_ctx_destroy(mbx_ctx *ctx);

/* Last error message of the context (never NULL).  With a NULL context: why the last
 * mbx_ctx_create on the calling thread failed. */
const char *mbx_last_error(const mbx_ctx *ctx);

/* Validate a parameter block and derive the per-frame constants (camera rotation, light
 * vector, fractal sequence data) on the host, then upload them.  Replaces
 * cOpenClEngineRenderFractal::SetParameters (opencl_engine_render_fractal.cpp:1114-1245)
 * and the sParamRender / cHybridFractalSequences set-up of cRenderJob::Execute
 * (render_job.cpp:399-406, 477). */
int mbx_scene_upload(mbx_ctx *ctx, const mbx_params *params);

/* Derived per-frame constants of the uploaded scene, as the host computed them in FP64:
 * the camera rotation mRot (row-major 3x3, render_worker.cpp:539-564), the CalcFOV value
 * (projection_3d.cpp:140-159), light #1's direction (light.cpp:130-185) and the device DE
 * kernel kind chosen for the fractal (0 Mandelbulb-8 trig-free, 1 Mandelbulb trig, 2 Mandelbox
 * analytic, 3 Mandelbox delta, 4 Kaleidoscopic IFS, 5 Mandelbulb in FP64).  Any output pointer
 * may be NULL. */
int mbx_scene_info(const mbx_ctx *ctx, double *mrot9, double *fov, double *light_dir3,
	int *de_kind);

/* Render the uploaded scene into caller-owned device buffers.  Replaces RenderMulti
 * (opencl_engine_render_fractal.cpp:2089-2425) and, on the CPU side, cRenderer::RenderImage
 * (render_image.cpp:354) / cRenderWorker::doWork (render_worker.cpp:91-530).
 * Asynchronous on the context stream; mbx_get_stats synchronises. */
int mbx_render(mbx_ctx *ctx, const mbx_outputs *out);

/* Point DE query: for n points (xyz doubles, host memory) return the distance and
 * iteration count of CalculateDistance at detail size `detail_size` (host memory).  Mirrors
 * the OpenCL distance engine (mandelbulber2/opencl/engines/distance_engine.cl:42) and
 * cRenderWorker-free CalculateDistance (calculate_distance.cpp:60-64). */
int mbx_distance(mbx_ctx *ctx, const double *points, int n, double detail_size,
	float *dist_out, int32_t *iters_out);

/* PrepareAOVectors (render_worker.cpp:599-639): the directions and colours of multi-ray AO from
 * the lightmap texture ("file_lightmap", which the caller decodes: width x height RGB floats,
 * row-major, each channel as cTexture holds it, the 8-bit value / 256).  NULL stands for the
 * texture cTexture substitutes when the file does not load (5 x 5 white, texture.cpp:150-160).
 * Fills params->aoVectors, aoColors and aoVectorsCount from params->ambientOcclusionQuality and
 * params->aoLightMapRotation.  Host only.  MBX_ERR_UNSUPPORTED when more than
 * MBX_MAX_AO_VECTORS directions would be kept, MBX_ERR_INVALID for bad arguments. */
int mbx_prepare_ao_vectors(mbx_params *params, const float *lightmap, int width, int height);

/* ---- formula plugins --------------------------------------------------------------------
 * The reference's formula contract is cAbstractFractal (formula/definition/abstract_fractal.h:
 * 74-109): FormulaCode(CVector4 &z, const sFractal *fractal, sExtendedAux &aux) plus metadata
 * each formula's constructor sets.  The device holds one <Name>Iteration(z, fractal, aux) per
 * formula of its table (csrc/mbx_generic.h); these calls expose the table's metadata and run one
 * iteration of any of its formulas. */

/* The metadata of one formula: cAbstractFractal's fields (abstract_fractal.h:85-93) as the
 * formula's constructor sets them. */
typedef struct mbx_formula_info {
	int32_t internalID;         /* enumFractalFormula (MBX_FORMULA_*) */
	int32_t DEType;             /* mbx_de_type */
	int32_t DEFunctionType;     /* mbx_de_function */
	int32_t cpixelAddition;     /* mbx_cpixel_addition */
	double defaultBailout;
	int32_t DEAnalyticFunction; /* mbx_analytic_function */
	int32_t coloringFunction;   /* mbx_coloring_function */
	char internalName[64];      /* internalName, e.g. "mandelbox" */
	char nameInComboBox[64];    /* nameInComboBox, e.g. "Mandelbox" */
} mbx_formula_info;

/* Number of formulas the device renders (the rows of mbx_formula_get).  Host only.  The
 * reference's counterpart is newFractalList (formula/definition/all_fractal_list.cpp:39). */
int mbx_formula_count(void);

/* Metadata of formula `index` (0 <= index < mbx_formula_count()), replacing the getters of
 * cAbstractFractal (abstract_fractal.h:96-104).  Host only.  MBX_ERR_INVALID for an index out
 * of range or a NULL info. */
int mbx_formula_get(int index, mbx_formula_info *info);

/* sExtendedAux (src/fractal.h:77-104), field for field. */
typedef struct mbx_extended_aux {
	double c[4];       /* variable c */
	double const_c[4]; /* constant c */
	double old_z[4];   /* hybrid colour 2 */
	double pos_neg;
	double r;
	double DE;
	double DE0;
	double dist;
	double pseudoKleinianDE;
	double actualScale;
	double actualScaleA;
	double color;
	double colorHybrid;
	double temp1000;
	int32_t i;
	int32_t reserved;
} mbx_extended_aux;

/* One FormulaCode call (abstract_fractal.h:82) of fractal->formula per point, on the device of
 * ctx: z (4 doubles per point) and aux as given, the formula's parameters from `fractal` after
 * RecalculateFractalParams (fractal.cpp:653-678); the updated z and aux are written to z_out and
 * aux_out (host memory, n entries; they may alias the inputs).  The formula-level known-answer
 * hook: the reference's own formula sources give the expected outputs.  Weight, c addition and
 * bailout belong to the sequence, not to the formula, and are not applied.  MBX_ERR_UNSUPPORTED
 * for a formula outside the table; needs no uploaded scene. */
int mbx_formula_iterate(mbx_ctx *ctx, const mbx_fractal *fractal, int n, const double *z_in,
	const mbx_extended_aux *aux_in, double *z_out, mbx_extended_aux *aux_out);

/* Render n frames (each its own parameter block and output planes) in one launch.  By default
 * that launch is the grid kernel (one 64-lane workgroup per 64-pixel tile of every frame, launched
 * directly); MBX_TUNE_KERNEL selects one plain launch per frame or the persistent kernel, and
 * MBX_TUNE_GRAPH replays the launch from a captured HIP graph.  The frames may differ in
 * resolution, camera and formula parameters but must share one DE kernel kind (Mandelbulb frames
 * of different powers share the general-power kernel).
 * Replaces the per-frame job loop of cKeyframeAnimation::RenderKeyframes
 * (mandelbulber2/src/animation_keyframes.cpp:1058-1111), which runs one full render (and one
 * OpenCL launch) per frame.  Asynchronous; mbx_get_stats synchronises. */
int mbx_render_batch(mbx_ctx *ctx, const mbx_params *frames, const mbx_outputs *outs, int n);

/* mbx_render_batch without any device allocation: the same kernels and results, with the
 * frame constants, job table, tile queues and work counters in a static device array of
 * about 4.7 MiB per GPU (MBX_NOALLOC_MAX_FRAMES frame blocks), which each GPU the library runs on
 * holds until the process ends.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc); the context is created beforehand.  n must be 1 .. MBX_NOALLOC_MAX_FRAMES
 * (MBX_ERR_INVALID otherwise).  All calls on one GPU share that GPU's array, which keeps the
 * last call's contents: calls that overlap on one GPU (two contexts on different streams)
 * race.  Graph replay (MBX_TUNE_GRAPH) does not apply; the kernel mode does.  Stats of the call
 * are read with mbx_get_stats. */
int mbx_render_batch_noalloc(mbx_ctx *ctx, const mbx_params *frames, const mbx_outputs *outs, int n);
/* Launch tuning keys for mbx_tune. */
#define MBX_TUNE_KERNEL 0    /* 0 = one grid launch over every tile of every frame of a render or batch
                                (default), 1 = one plain launch per frame, 2 = persistent waves
                                draining tile queues */
#define MBX_TUNE_WAVES 1     /* persistent kernel: number of 64-lane waves launched (0 = default) */
#define MBX_TUNE_GRAPH 2     /* modes 0 and 2: 1 = replay a captured HIP graph, 0 = direct launches
                                (default: with frame-sized kernels the graph launch costs ~10 us
                                more per render than launching its two kernels directly) */
#define MBX_TUNE_DELTA_DE 3  /* delta DE: 0 = the reference's finite differences (default),
                                1 = forward-mode tangents (smooth limit of the same quotient) */
#define MBX_TUNE_TILE_WIDTH 4 /* pixels per tile row: 4, 8 (default), 16 or 32; a tile is 64
                                 pixels (one wave), so its height is 64 / width */
#define MBX_TUNE_FAR_FIELD 5  /* 1 (default) = Mandelbulb march steps beyond the far-field radius,
                                 where the first formula iteration bounds the DE above the value at
                                 which the reference's step clamps to 3, take that step without
                                 evaluating the DE (pixel results are identical); 0 = evaluate */
#define MBX_TUNE_PRECISION 6  /* Mandelbulb and Kaleidoscopic IFS march precision: 0 (default) = FP32,
                                 or FP64 position and orbit when FP32 rounding of a coordinate near
                                 the nearest hit exceeds 1/256 of the view's smallest distance
                                 threshold (deep zooms and close-ups; the IFS then takes the general
                                 FP64 kernel); 1 = always FP32; 2 = always FP64.  Mandelbox scenes
                                 always march in FP64 */
#define MBX_TUNE_TIMING 7     /* 1 (default) = bracket each mbx_render / mbx_render_batch with events so
                                 that mbx_get_stats reports its device time; 0 = no events (each event
                                 marker costs the stream ~10 us between launches; kernel_ms reads 0) */
#define MBX_TUNE_SPLIT 8      /* grid mode, frames with shadows, AO, palette, highlights or glow and no
                                 Monte Carlo sampling: a march launch that shades the misses and lists
                                 the hits, then a shading launch over the hit lists (hits shaded in
                                 full waves; hit records of up to 48 bytes per pixel of the batch).
                                 1 (default) = for the Kaleidoscopic IFS kind, where it is faster;
                                 2 = for every kind; 0 = never.  The geometry planes are bit for bit
                                 those of the one-launch render, the shading planes too on the FP64
                                 kinds and to the last bit of FP32 rounding on the FP32 kinds */

/* Set a launch or numerics option (see MBX_TUNE_*); scene options apply from the next
 * mbx_scene_upload / mbx_render_batch.  No reference counterpart: the OpenCL engine's
 * adaptive job size (opencl_engine.cpp:429-450) is the closest analogue. */
int mbx_tune(mbx_ctx *ctx, int key, int value);

/* Current value of a tuning key (the wave count resolves the default). */
int mbx_get_tune(mbx_ctx *ctx, int key, int *value);

/* Project autotuner: time the uploaded scene in the current kernel mode for each tile shape
 * (tile widths 8, 16, 4, 32) and, for the persistent kernel, a set of wave counts (`reps` renders
 * each into `out`), and keep the fastest.  Returns the wave count in *best_waves and the time per
 * render in *best_ms (either may be NULL); mbx_get_tune(MBX_TUNE_TILE_WIDTH) gives the tile width. */
int mbx_autotune(mbx_ctx *ctx, const mbx_outputs *out, int reps, int *best_waves, double *best_ms);

/* Statistics of the last render (synchronises the stream). */
int mbx_get_stats(mbx_ctx *ctx, mbx_stats *stats);

/* ---- differentiable render and adjoint (inverse rendering) --------------------------------
 * No counterpart in the reference, which renders only forward; these make the uploaded scene's
 * image a differentiable function of MBX_DIFF_NPARAMS parameters, in this order: */
#define MBX_DIFF_NPARAMS 10
#define MBX_DIFF_POWER 0      /* fractal.bulb.power */
#define MBX_DIFF_ROTATION 1   /* formulaRotation x, y, z (degrees): indices 1, 2, 3 */
#define MBX_DIFF_CAMERA 4     /* camera x, y, z: indices 4, 5, 6 */
#define MBX_DIFF_TARGET 7     /* target x, y, z: indices 7, 8, 9 */

/* Per-ray checkpoint written by mbx_render_diff and read by mbx_adjoint (device memory, one per
 * pixel): the refined hit depth, so the adjoint never re-marches. */
typedef struct mbx_ray_checkpoint {
	double t;        /* depth along the unit view vector (refined) */
	int32_t found;   /* the march hit the surface */
	int32_t refined; /* the hit was refined to DE(p) = distThresh(p); 0 = not differentiable */
	int32_t iters;   /* formula iterations at the refined point */
	int32_t pad;
} mbx_ray_checkpoint;

/* Differentiable render of the uploaded scene (Mandelbulb, analytic logarithmic DE, c = p; one
 * directional light at an absolute angle, Lambert; no shadows, AO or Monte Carlo; background
 * gamma 1): the usual march finds the hit, which is then refined in FP64 by Newton iterations on
 * DE(camera + t dir) = distThresh(t), shaded with the analytic normal (the DE gradient by
 * forward-mode dual numbers) in FP64, and checkpointed.  Outputs as mbx_render; `checkpoints`
 * (device, imageWidth*imageHeight entries) is required.  MBX_ERR_UNSUPPORTED names the scene
 * setting that is not differentiable here. */
int mbx_render_diff(mbx_ctx *ctx, const mbx_outputs *out, mbx_ray_checkpoint *checkpoints);

/* Adjoint of mbx_render_diff: given dL/drgb (device, 3 floats per pixel) and the checkpoints of
 * the last mbx_render_diff of the same scene, writes dL/dtheta (host, MBX_DIFF_NPARAMS doubles)
 * and, if pixel_grad is not NULL, each pixel's contribution (device, MBX_DIFF_NPARAMS doubles
 * per pixel).  Hits use the implicit derivative of the refined depth and the derivative of the
 * analytic normal (second derivatives of the DE by hyper-dual numbers); misses the derivative of
 * the background along the view vector.  Synchronises the stream. */
int mbx_adjoint(mbx_ctx *ctx, const mbx_ray_checkpoint *checkpoints, const float *dL_drgb, double *grad,
	double *pixel_grad);

/* ---- parameter defaults and .fract scene files ------------------------------------------- */

/* Fill *p with the reference's defaults: the values InitParams, InitFractalParams,
 * InitMaterialParams(1) and InitLightParams(1) register (mandelbulber2/src/initparameters.cpp),
 * with colours as toRGBFloat converts them (c / 65536, color_structures.hpp:106-109).  Host only;
 * needs no context or device.  The multi-ray AO direction set is left empty (aoVectorsCount 0). */
void mbx_params_default(mbx_params *p);

/* Effects a scene file turns on that this library does not render (mbx_fract_info.unsupported).
 * A scene loads with them regardless; the image is the path's image without them. */
#define MBX_FRACT_HYBRID 0x1u          /* hybrid sequence, or a boolean tree of several objects */
#define MBX_FRACT_PRIMITIVES 0x2u      /* primitive objects (plane, sphere, water, ...) */
#define MBX_FRACT_TRANSFORM 0x4u       /* node position, scale or repeat; the limits box */
#define MBX_FRACT_PALETTE 0x8u         /* palette colouring beyond the surface gradient: colour by
                                          numbers, or specular / diffuse / luminosity / roughness /
                                          reflectance / transparency gradients */
#define MBX_FRACT_SPECULAR 0x10u       /* (unused: specular highlights are rendered) */
#define MBX_FRACT_GLOW 0x20u           /* glow with a volumetric DE factor other than 1, or on GI
                                          bounces (MC_global_illumination_volumetric) */
#define MBX_FRACT_FOG 0x40u            /* iteration fog, clouds, shadows in fog, or basic / distance
                                          fog with a volumetric DE factor other than 1 (basic and
                                          distance fog themselves are rendered) */
#define MBX_FRACT_REFLECTIONS 0x80u    /* ray-traced reflections or transparency */
#define MBX_FRACT_LIGHTS 0x100u        /* auxiliary, random, fake or volumetric lights; light #1 not
                                          directional */
#define MBX_FRACT_TEXTURES 0x200u      /* textured background, environment map, material textures */
#define MBX_FRACT_AO_SCREEN_SPACE 0x400u /* (unused: screen-space AO is rendered by mbx_render_ssao) */
#define MBX_FRACT_AO_LIGHTMAP 0x800u   /* multi-ray AO: its directions and colours come from the
                                          lightmap texture, which the reader cannot decode
                                          (aoVectorsCount is left 0): pass the decoded
                                          file_lightmap to mbx_prepare_ao_vectors */
#define MBX_FRACT_POSTPROCESS 0x1000u  /* brightness, contrast, gamma, saturation, HDR, blur, post DOF */
#define MBX_FRACT_MC_ADAPTIVE 0x2000u  /* (unused: the adaptive Monte Carlo stop is rendered) */
#define MBX_FRACT_PERSPECTIVE 0x4000u  /* stereo rendering, or a fish-eye or equirectangular projection
                                          with the Monte Carlo lens, chromatic aberration or fog */
#define MBX_FRACT_FORMULA 0x8000u      /* a formula outside the device's table (the MBX_FORMULA_* ids) */
#define MBX_FRACT_ANIMATION 0x10000u   /* [frames] / [keyframes] sections (not read) */
#define MBX_FRACT_MATERIAL 0x20000u    /* other material effects: iridescence, roughness, subsurface,
                                          perlin noise, luminosity gradient */
#define MBX_FRACT_INTERIOR 0x40000u    /* (unused: interior and iteration-threshold modes are rendered) */
#define MBX_FRACT_DISABLED 0x80000u    /* the fractal node is disabled or absent: nothing to render */
#define MBX_FRACT_ANTIALIASING 0x100000u /* antialiasing with the Monte Carlo lens, chromatic aberration
                                            or fog (antialiasing alone is rendered) */
#define MBX_FRACT_MC_EFFECTS 0x200000u /* (unused: Monte Carlo soft shadows and chromatic aberration are rendered) */
#define MBX_FRACT_NEBULA 0x400000u     /* nebula mode (the OpenCL nebula engine renders the scene) */

/* What mbx_load_fract found besides the parameter block. */
typedef struct mbx_fract_info {
	double fileVersion;    /* "# version x.yy" of the header */
	int32_t format;        /* 0 "all parameters", 1 "only modified parameters" */
	int32_t lines;         /* parameter lines decoded */
	int32_t errors;        /* malformed lines (no space, no ';', or the ';' before the space) */
	int32_t formula;       /* enumFractalFormula id of the rendered fractal node */
	uint32_t unsupported;  /* MBX_FRACT_* */
	int32_t reserved;
	char message[256];     /* the first error ("" if none) */
} mbx_fract_info;

/* Read a Mandelbulber settings file (the text of a .fract file, `len` bytes, UTF-8) into *params:
 * the reference's defaults (mbx_params_default) overwritten by the file's [main_parameters] and
 * [fractal_N] lines, with the version-dependent renames and migrations of cSettings
 * (mandelbulber2/src/settings.cpp:513-977, 1015-1245, 1262-1492, 1494-1754, 2551-2760) that concern
 * these parameters, and the objects tree reduced to its one fractal node (objects_tree.cpp:36-96).
 * Image size, camera, light #1, material of the node, AO, background, Monte Carlo and the
 * formula's parameters are read; everything else is reported in info->unsupported when it changes
 * the image.  Returns MBX_OK, or MBX_ERR_INVALID when the header is not a Mandelbulber settings
 * header, the file is newer than version 2.35, it holds application settings, or more than three
 * lines are malformed (the reference's limit, settings.cpp:888-901); info->message says which.
 * `info` may be NULL.  Host only; needs no context or device. */
int mbx_load_fract(const char *text, size_t len, mbx_params *params, mbx_fract_info *info);

/* cImage::CompileImage + ConvertTo8bitChar (mandelbulber2/src/cimage.cpp:317-397, 460-470): the
 * float image `rgb` (3 floats per pixel, row-major, as mbx_outputs.rgb) of width x height pixels
 * to 16-bit RGB `rgb16` through brightness, contrast about 0.5, the tanh tone map (hdr),
 * saturation about the luma sqrt(0.299 R^2 + 0.587 G^2 + 0.114 B^2), truncation to 16 bits and the
 * reference's 65536-entry gamma table int(powf(i / 65536, 1 / gamma) * 65535), and to 8-bit RGB
 * `rgb8` = rgb16 / 256.  All three are device pointers; rgb16 or rgb8 may be NULL.  Asynchronous
 * on the context stream (mbx_get_stats or a stream sync waits for it).  MBX_ERR_INVALID for a
 * non-positive size or gamma, or both outputs NULL. */
int mbx_compile_image(mbx_ctx *ctx, const mbx_image_adjustments *adj, const float *rgb, int32_t width,
	int32_t height, uint16_t *rgb16, uint8_t *rgb8);

/* The z-buffer image of ImageFileSavePNG::SavePNG (mandelbulber2/src/file_image.cpp:628-742):
 * `depth` (device, 1 float per pixel, as mbx_outputs.depth) to 8- or 16-bit grey in `out` (device,
 * uint8_t or uint16_t per pixel).  The range [minZ, maxZ] is the frame's (its smallest depth and
 * its largest depth below 1e19) or, with constant_range, [min_depth, max_depth] as floats; then
 * z1 = logf(z / minZ) / logf(maxZ / minZ) (logarithmic) or (z - minZ) / (maxZ - minZ), clamped to
 * [0, 1], int(z1 * 254) or int(z1 * 65534); a miss (z > 1e19) is 255 or 65535; invert subtracts
 * from 255 or 65535.  The reference's defaults (zbuffer_logarithmic true, zbuffer_invert false,
 * zbuffer_constant_range false, initparameters.cpp:618-622).  Asynchronous on the context stream. */
int mbx_compile_zbuffer(mbx_ctx *ctx, const float *depth, int32_t width, int32_t height, int32_t bits,
	int32_t logarithmic, int32_t invert, int32_t constant_range, double min_depth, double max_depth, void *out);

/* Screen-space ambient occlusion, the post effect cRenderer runs after the render when
 * ambient_occlusion_mode is 2 (render_image.cpp:498-508; cRenderSSAO::RenderSSAO,
 * render_ssao.cpp:76-200, over the whole image with no progressive step; per pixel
 * cSSAOWorker::doWork, ssao_worker.cpp:63-261).  For each pixel with depth below 1e19 the depth
 * plane is scanned along quality = max(3, ambientOcclusionQuality^2) angles at radii r^2 * W /
 * quality^2 / 2 pixels (r = 1, 2, ... < quality), the steepest rise of the surface towards the
 * camera per angle gives atan(max rise), and the mean of 0.5 - atan / pi over the angles that
 * found a sample is the ambient term; the pixel gains colour / 256 * ambient * ambientOcclusion *
 * (1 - opacity / 65535) * ambientOcclusionColor.  The image's last column is left as it is, as the
 * reference's worker leaves it.  params gives the image size, perspectiveType and fov (CalcFOV,
 * projection_3d.cpp:140-159), the quality, intensity, colour and SSAORandomMode; in random mode
 * each pixel takes its angle offsets and radius step from its own Park-Miller stream, seeded by
 * (pixel index, randomSeed ^ 0x55A0A055), where the reference draws them from its one shared
 * Random() stream.  depth (float), colour (3 x uint8), opacity (uint16, NULL = 0), rgb_in and
 * rgb_out (3 x float; may be the same buffer) are device planes of imageWidth x imageHeight pixels
 * as mbx_render writes them.  FP64 arithmetic in the reference's operation order.  Asynchronous on
 * the context stream.  MBX_ERR_INVALID for a bad size or a NULL plane other than opacity. */
int mbx_render_ssao(mbx_ctx *ctx, const mbx_params *params, const float *depth, const uint8_t *colour,
	const uint16_t *opacity, const float *rgb_in, float *rgb_out);

/* ImageFileSavePNG::SavePNG (mandelbulber2/src/file_image.cpp:481-600): write `pixels` (host
 * memory, row-major, `channels` samples per pixel: 3 = RGB as the colour image, 1 = grey as the
 * z-buffer image; `bits` 8 = uint8_t samples, 16 = uint16_t samples, stored big-endian in the file
 * as png_set_swap makes them) as a non-interlaced PNG file at `path`.  Host only.  Returns MBX_OK,
 * MBX_ERR_INVALID for bad arguments, MBX_ERR_IO when the file cannot be written. */
int mbx_write_png(const char *path, int32_t width, int32_t height, int32_t bits, int32_t channels,
	const void *pixels);

/* ImageFileSaveTIFF::SaveTIFF (mandelbulber2/src/file_image.cpp:1590-1760): write `pixels` (host
 * memory, row-major, `channels` contiguous samples per pixel: 1 grey, 3 RGB, 4 RGB + alpha;
 * `bits` 8 or 16 = uint8_t / uint16_t samples, 32 = float samples) as a baseline little-endian
 * TIFF with the fields the reference sets: strips of 64 rows deflated (compression 32946),
 * contiguous planar configuration, MSB-to-LSB fill order, sample format unsigned or IEEE float;
 * a fourth sample is marked unassociated alpha.  The caller composes the samples as the reference
 * does for each channel (the 8/16-bit planes of mbx_compile_image for the colour image, those
 * planes / 65536 as floats for 32 bits).  Host only.  MBX_OK, MBX_ERR_INVALID, MBX_ERR_IO. */
int mbx_write_tiff(const char *path, int32_t width, int32_t height, int32_t bits, int32_t channels,
	const void *pixels);

/* ImageFileSaveEXR::SaveEXR (file_image.cpp:1249-1530): write `nchannels` channels as a
 * single-part scanline OpenEXR file with ZIPS compression.  Channel c is named names[c] (1 to 31
 * characters, e.g. "R", "G", "B", "A", "Z", "n.X", "s.R" as the reference names them), stored as
 * HALF (pixel_types[c] 1, rounded to nearest even) or FLOAT (2), read from host memory at
 * planes[c][(y * width + x) * strides[c]] (so an interleaved RGB plane is three channels with
 * stride 3); the file lists the channels sorted by name, as the format requires.  linear sets the
 * channels' pLinear flag (the reference's linear_colorspace).  The reference writes the raw float
 * image (not the post-processed one) for R, G, B.  Host only.  MBX_OK, MBX_ERR_INVALID (bad size,
 * type, stride, empty / long / repeated name), MBX_ERR_IO. */
int mbx_write_exr(const char *path, int32_t width, int32_t height, int32_t nchannels, const char *const *names,
	const int32_t *pixel_types, const float *const *planes, const int32_t *strides, int32_t linear);

/* The HALF value mbx_write_exr stores for a float (IEEE binary16, round to nearest even). */
uint16_t mbx_float_to_half(float f);

#ifdef __cplusplus
}
#endif

#endif /* MBX_H_ */

"""ctypes mirror of include/mbx.h (the C ABI of libmbx.so).

Field names follow the reference's sParamRender / sFractal members and .fract keys
(mandelbulber2/src/fractparams.cpp:80-356, mandelbulber2/src/fractal.cpp:41-120); the C header
documents each field.  Keep this file in the same order as the header: the layout is checked
by tests against sizeof(mbx_params) reported by the compiled library.
"""
import ctypes as C

MBX_IFS_VECTOR_COUNT = 9
MBX_MANDELBOX_FOLDS = 2
MBX_MAX_AO_VECTORS = 1024
MBX_MAX_GRADIENT_STOPS = 64
MBX_MAX_FRACTALS = 9
MBX_NOALLOC_MAX_FRAMES = 64  # frames of one mbx_render_batch_noalloc call

MBX_FORMULA_NONE = 0
MBX_FORMULA_MANDELBULB = 2
MBX_FORMULA_MANDELBULB_POWER2 = 3
MBX_FORMULA_HYPERCOMPLEX = 4
MBX_FORMULA_QUATERNION = 5
MBX_FORMULA_MENGER_SPONGE = 7
MBX_FORMULA_MANDELBOX = 8
MBX_FORMULA_KALEIDOSCOPIC_IFS = 10
MBX_FORMULA_XENODREAMBUIE = 11
MBX_FORMULA_BOX_FOLD_BULB_POW2 = 16
MBX_FORMULA_AEXION = 19

MBX_AO_FAST = 0
MBX_AO_MULTIPLE_RAYS = 1
MBX_AO_SCREEN_SPACE = 2

MBX_RNG_PER_PIXEL = 0
MBX_RNG_GLOBAL_SEQUENTIAL = 1

D3 = C.c_double * 3


class MbxFractalMandelbulb(C.Structure):
    _fields_ = [
        ("power", C.c_double),
        ("alphaAngleOffset", C.c_double),
        ("betaAngleOffset", C.c_double),
    ]


class MbxFractalMandelbox(C.Structure):
    _fields_ = [
        ("scale", C.c_double),
        ("foldingLimit", C.c_double),
        ("foldingValue", C.c_double),
        ("foldingSphericalMin", C.c_double),
        ("foldingSphericalFixed", C.c_double),
        ("offset", D3),
        ("rotationMain", D3),
        ("rotation", (D3 * 3) * MBX_MANDELBOX_FOLDS),
        ("colorFactor", D3),
        ("colorFactorSp1", C.c_double),
        ("colorFactorSp2", C.c_double),
        ("colorFactorR", C.c_double),
        ("rotationsEnabled", C.c_int32),
        ("mainRotationEnabled", C.c_int32),
    ]


class MbxFractalIFS(C.Structure):
    _fields_ = [
        ("scale", C.c_double),
        ("offset", D3),
        ("rotation", D3),
        ("edge", D3),
        ("direction", D3 * MBX_IFS_VECTOR_COUNT),
        ("rotations", D3 * MBX_IFS_VECTOR_COUNT),
        ("distance", C.c_double * MBX_IFS_VECTOR_COUNT),
        ("intensity", C.c_double * MBX_IFS_VECTOR_COUNT),
        ("enabled", C.c_int32 * MBX_IFS_VECTOR_COUNT),
        ("absX", C.c_int32),
        ("absY", C.c_int32),
        ("absZ", C.c_int32),
        ("mengerSpongeMode", C.c_int32),
        ("rotationEnabled", C.c_int32),
        ("edgeEnabled", C.c_int32),
    ]


class MbxFractalTransform(C.Structure):
    _fields_ = [
        ("scale3", C.c_double),
        ("aexionCadd", C.c_double),
        ("foldingIntPowFoldFactor", C.c_double),
        ("foldingIntPowZFactor", C.c_double),
        ("analyticDEScale1", C.c_double),
        ("analyticDEOffset1", C.c_double),
        ("analyticDEOffset2", C.c_double),
        ("analyticDEEnabled", C.c_int32),
        ("analyticDEEnabledFalse", C.c_int32),
    ]


class MbxFractal(C.Structure):
    _fields_ = [
        ("formula", C.c_int32),
        ("formulaMaxiter", C.c_int32),
        ("deltaDEMethod", C.c_int32),
        ("deltaDEFunction", C.c_int32),
        ("useDefaultBailout", C.c_int32),
        ("juliaMode", C.c_int32),
        ("dontAddCConstant", C.c_int32),
        ("checkForBailout", C.c_int32),
        ("bailout", C.c_double),
        ("juliaC", D3),
        ("constantFactor", D3),
        ("initialWAxis", C.c_double),
        ("formulaIterations", C.c_int32),
        ("formulaStartIteration", C.c_int32),
        ("formulaStopIteration", C.c_int32),
        ("reserved0", C.c_int32),
        ("formulaWeight", C.c_double),
        ("bulb", MbxFractalMandelbulb),
        ("mandelbox", MbxFractalMandelbox),
        ("IFS", MbxFractalIFS),
        ("transform", MbxFractalTransform),
    ]


class MbxLight(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32),
        ("castShadows", C.c_int32),
        ("penetrating", C.c_int32),
        ("relativePosition", C.c_int32),
        ("rotation", D3),
        ("intensity", C.c_double),
        ("visibility", C.c_double),
        ("size", C.c_double),
        ("softShadowCone", C.c_double),
        ("contourSharpness", C.c_double),
        ("color", D3),
    ]


class MbxImageAdjustments(C.Structure):
    _fields_ = [
        ("brightness", C.c_float),
        ("contrast", C.c_float),
        ("gamma", C.c_float),
        ("saturation", C.c_float),
        ("hdr", C.c_int32),
        ("reserved", C.c_int32),
    ]


class MbxParams(C.Structure):
    _fields_ = [
        ("imageWidth", C.c_int32),
        ("imageHeight", C.c_int32),
        ("perspectiveType", C.c_int32),
        ("maxRaymarchingSteps", C.c_int32),
        ("minN", C.c_int32),
        ("iterThreshMode", C.c_int32),
        ("interiorMode", C.c_int32),
        ("constantDEThreshold", C.c_int32),
        ("camera", D3),
        ("target", D3),
        ("topVector", D3),
        ("fov", C.c_double),
        ("sweetSpotHAngle", C.c_double),
        ("sweetSpotVAngle", C.c_double),
        ("detailLevel", C.c_double),
        ("DEThresh", C.c_double),
        ("DEFactor", C.c_double),
        ("smoothness", C.c_double),
        ("viewDistanceMax", C.c_double),
        ("surfaceColor", D3),
        ("shading", C.c_double),
        ("luminosity", C.c_double),
        ("luminosityColor", D3),
        ("fillLightColor", D3),
        ("light", MbxLight),
        ("ambientOcclusionEnabled", C.c_int32),
        ("ambientOcclusionMode", C.c_int32),
        ("ambientOcclusionQuality", C.c_int32),
        ("aoVectorsCount", C.c_int32),
        ("ambientOcclusion", C.c_double),
        ("ambientOcclusionColor", D3),
        ("ambientOcclusionFastTune", C.c_double),
        ("aoVectors", D3 * MBX_MAX_AO_VECTORS),
        ("aoColors", D3 * MBX_MAX_AO_VECTORS),
        ("background3ColorsEnable", C.c_int32),
        ("reserved1", C.c_int32),
        ("backgroundColor1", D3),
        ("backgroundColor2", D3),
        ("backgroundColor3", D3),
        ("backgroundBrightness", C.c_double),
        ("backgroundGamma", C.c_double),
        ("DOFMonteCarlo", C.c_int32),
        ("DOFSamples", C.c_int32),
        ("DOFMonteCarloGlobalIllumination", C.c_int32),
        ("reflectionsMax", C.c_int32),
        ("monteCarloGIRadianceLimit", C.c_double),
        ("rngMode", C.c_int32),
        ("randomSeed", C.c_uint32),
        ("fractal", MbxFractal),
        ("formulaRotation", C.c_double * 3),
        ("useColorsFromPalette", C.c_int32),
        ("surfaceGradientEnable", C.c_int32),
        ("coloringAlgorithm", C.c_int32),
        ("coloringPreV215", C.c_int32),
        ("coloring4D", C.c_int32),
        ("surfaceGradientCount", C.c_int32),
        ("coloringSpeed", C.c_double),
        ("coloringPaletteOffset", C.c_double),
        ("coloringSphereRadius", C.c_double),
        ("coloringLineDirection", D3),
        ("surfaceGradientPosition", C.c_double * MBX_MAX_GRADIENT_STOPS),
        ("surfaceGradientColor", D3 * MBX_MAX_GRADIENT_STOPS),
        ("specularPlasticEnable", C.c_int32),
        ("metallic", C.c_int32),
        ("specular", C.c_double),
        ("specularWidth", C.c_double),
        ("specularMetallic", C.c_double),
        ("specularMetallicWidth", C.c_double),
        ("specularMetallicRoughness", C.c_double),
        ("specularColor", D3),
        ("luminosityEmissive", C.c_double),
        ("glowEnabled", C.c_int32),
        ("reserved2", C.c_int32),
        ("glowIntensity", C.c_double),
        ("glowColor1", D3),
        ("glowColor2", D3),
        ("hybridCount", C.c_int32),
        ("repeatFrom", C.c_int32),
        ("linearDEOffset", C.c_double),
        ("hybridFractals", MbxFractal * (MBX_MAX_FRACTALS - 1)),
        ("advancedQuality", C.c_int32),
        ("reserved3", C.c_int32),
        ("absMaxMarchingStep", C.c_double),
        ("absMinMarchingStep", C.c_double),
        ("relMaxMarchingStep", C.c_double),
        ("relMinMarchingStep", C.c_double),
        ("detailSizeMax", C.c_double),
        ("detailSizeMin", C.c_double),
        ("deltaDERelativeDelta", C.c_double),
        ("image", MbxImageAdjustments),
        ("boxFolding", C.c_int32),
        ("sphericalFolding", C.c_int32),
        ("boxFoldingLimit", C.c_double),
        ("boxFoldingValue", C.c_double),
        ("sphericalFoldingOuter", C.c_double),
        ("sphericalFoldingInner", C.c_double),
        ("slowShading", C.c_int32),
        ("N", C.c_int32),
        ("DOFMinSamples", C.c_int32),
        ("reserved5", C.c_int32),
        ("DOFMaxNoise", C.c_double),
        ("aoLightMapRotation", C.c_double),
        ("DOFEnabled", C.c_int32),
        ("DOFMonteCarloChromaticAberration", C.c_int32),
        ("DOFFocus", C.c_double),
        ("DOFRadius", C.c_double),
        ("DOFMonteCarloCACameraDispersion", C.c_double),
        ("monteCarloSoftShadows", C.c_int32),
        ("SSAORandomMode", C.c_int32),
        ("basicFogEnabled", C.c_int32),
        ("basicFogCastShadows", C.c_int32),
        ("basicFogVisibility", C.c_double),
        ("basicFogColor", D3),
        ("volFogEnabled", C.c_int32),
        ("distanceFogShadows", C.c_int32),
        ("volFogDensity", C.c_double),
        ("volFogColour1Distance", C.c_double),
        ("volFogColour2Distance", C.c_double),
        ("volFogDistanceFactor", C.c_double),
        ("volFogDistanceFromSurface", C.c_double),
        ("volFogColour", D3 * 3),
        ("iterFogEnabled", C.c_int32),
        ("cloudsEnabled", C.c_int32),
        ("volumetricLightDEFactor", C.c_double),
        ("antialiasingEnabled", C.c_int32),
        ("antialiasingSize", C.c_int32),
    ]


MBX_DIFF_NPARAMS = 10
DIFF_PARAM_NAMES = ("power", "rotation_x", "rotation_y", "rotation_z", "camera_x", "camera_y",
                    "camera_z", "target_x", "target_y", "target_z")


class MbxRayCheckpoint(C.Structure):
    _fields_ = [("t", C.c_double), ("found", C.c_int32), ("refined", C.c_int32),
                ("iters", C.c_int32), ("pad", C.c_int32)]


class MbxOutputs(C.Structure):
    _fields_ = [
        ("rgb", C.c_void_p),
        ("depth", C.c_void_p),
        ("iters", C.c_void_p),
        ("mask", C.c_void_p),
        ("normal", C.c_void_p),
        ("steps", C.c_void_p),
        ("colour", C.c_void_p),
        ("opacity", C.c_void_p),
        ("world", C.c_void_p),
        ("specular", C.c_void_p),
        ("shadow", C.c_void_p),
        ("gi", C.c_void_p),
    ]


class MbxStats(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64),
        ("de_evals", C.c_uint64),
        ("iterations", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("far_steps", C.c_uint64),
    ]


# effects a .fract file turns on that the library does not render (mbx_fract_info.unsupported)
FRACT_FLAGS = {
    "hybrid": 0x1, "primitives": 0x2, "transform": 0x4, "palette": 0x8, "specular": 0x10,
    "glow": 0x20, "fog": 0x40, "reflections": 0x80, "lights": 0x100, "textures": 0x200,
    "ao_screen_space": 0x400, "ao_lightmap": 0x800, "postprocess": 0x1000, "mc_adaptive": 0x2000,
    "perspective": 0x4000, "formula": 0x8000, "animation": 0x10000, "material": 0x20000,
    "interior": 0x40000, "disabled": 0x80000, "antialiasing": 0x100000, "mc_effects": 0x200000,
    "nebula": 0x400000,
}


class MbxFractInfo(C.Structure):
    _fields_ = [
        ("fileVersion", C.c_double),
        ("format", C.c_int32),
        ("lines", C.c_int32),
        ("errors", C.c_int32),
        ("formula", C.c_int32),
        ("unsupported", C.c_uint32),
        ("reserved", C.c_int32),
        ("message", C.c_char * 256),
    ]


class MbxFormulaInfo(C.Structure):
    """mbx_formula_info: cAbstractFractal's metadata (abstract_fractal.h:85-93)."""
    _fields_ = [
        ("internalID", C.c_int32),
        ("DEType", C.c_int32),
        ("DEFunctionType", C.c_int32),
        ("cpixelAddition", C.c_int32),
        ("defaultBailout", C.c_double),
        ("DEAnalyticFunction", C.c_int32),
        ("coloringFunction", C.c_int32),
        ("internalName", C.c_char * 64),
        ("nameInComboBox", C.c_char * 64),
    ]


class MbxExtendedAux(C.Structure):
    """mbx_extended_aux: sExtendedAux (fractal.h:77-104), field for field."""
    _fields_ = [
        ("c", C.c_double * 4),
        ("const_c", C.c_double * 4),
        ("old_z", C.c_double * 4),
        ("pos_neg", C.c_double),
        ("r", C.c_double),
        ("DE", C.c_double),
        ("DE0", C.c_double),
        ("dist", C.c_double),
        ("pseudoKleinianDE", C.c_double),
        ("actualScale", C.c_double),
        ("actualScaleA", C.c_double),
        ("color", C.c_double),
        ("colorHybrid", C.c_double),
        ("temp1000", C.c_double),
        ("i", C.c_int32),
        ("reserved", C.c_int32),
    ]


MBX_CPIXEL_ENABLED_BY_DEFAULT = 0
MBX_CPIXEL_DISABLED_BY_DEFAULT = 1
MBX_CPIXEL_ALREADY_HAS = 2
MBX_COLORING_DEFAULT = 0
MBX_COLORING_ABOX = 1
MBX_COLORING_IFS = 2
MBX_COLORING_AMAZING_SURF = 3
MBX_COLORING_DONUT = 4


def fract_flag_names(mask):
    return [k for k, v in FRACT_FLAGS.items() if mask & v]


def set3(arr, values):
    for i, v in enumerate(values):
        arr[i] = float(v)

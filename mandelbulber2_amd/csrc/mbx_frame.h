// mbx_frame.h - per-frame constants as the kernels read them.
//
// The host derives every value here from an mbx_params block in double precision
// (mbx_host.cpp) and uploads one FrameConst per frame into device memory.  All threads of a
// launch read the same FrameConst, so the loads are wave-uniform and become scalar (SGPR)
// loads: this plays the role of the OpenCL engine's __constant sClInConstants
// (mandelbulber2/opencl/input_data_structures.h:62-65).
#pragma once

#include <stdint.h>

#include "../../include/mbx.h"

namespace mbx
{

// Distance-estimator kinds the kernels are specialised for.
enum DEKind : int32_t
{
	kBulb8 = 0,      // Mandelbulb power 8, analytic log DE, trig-free power (w = 0)
	kBulbTrig = 1,   // Mandelbulb any power, analytic log DE, reference trig form
	kBoxAnalytic = 2,
	kBoxDelta = 3,   // Mandelbox, delta DE (reference finite differences, or tangents)
	kIfsAnalytic = 4,
	kBulbD = 5,      // Mandelbulb any power, analytic DE, FP64 position and orbit (deep zooms)
	kGeneric = 6,    // any formula of the table and hybrid sequences: FP64, 4D orbit, per-iteration
	                 // formula from the sequence
	kNumDEKinds = 7
};

// One container of the general formula path (kGeneric), the `const sFractal *fractal` that each
// formula's <Name>Iteration receives: cHybridFractalSequences::sFractalData
// (hybrid_fractal_sequences.h:30-56) with its sFractal after RecalculateFractalParams
// (fractal.cpp:653-678), in FP64.  An identity rotation is flagged: the reference skips the
// multiply for it (CRotationMatrix's zero flag).
struct FractalParams
{
	int32_t formula, addC, checkBailout, ifsMainRotId;
	int32_t mbMainRotOn, mbRotOn; // Mandelbox main rotation, rotated folds
	int32_t mbRotId[2][3];        // rotated folds: identity (skipped) matrices
	double mbMainRot[9];          // Mandelbox main rotation (SetRotation2)
	double mbRot[2][3][9];        // rotated folds: mandelbox.rot[fold][axis] (SetRotation2)
	double mbRotInv[2][3][9];     // and their transposes (rotinv)
	double weight, bailout;
	double power, alpha, beta; // Mandelbulb, Xenodreambuie (radians)
	double mbScale, mbAbsScale, mbFoldingLimit, mbFoldingValue, mbFR2, mbMR2, mbFactor1;
	double mbOffset[3], mbColor[3], mbSp1, mbSp2;
	double ifsScale, ifsOffset[3], ifsEdge[3];
	double ifsMainRot[9];
	double ifsRot[MBX_IFS_VECTOR_COUNT][9];
	double ifsDirection[MBX_IFS_VECTOR_COUNT][4];
	double ifsDistance[MBX_IFS_VECTOR_COUNT], ifsIntensity[MBX_IFS_VECTOR_COUNT];
	int32_t ifsEnabled[MBX_IFS_VECTOR_COUNT], ifsRotId[MBX_IFS_VECTOR_COUNT];
	int32_t ifsAbs[3], ifsMenger, ifsRotationEnabled, ifsEdgeEnabled;
	int32_t adeEnabled, adeEnabledFalse;
	double scale3, cadd, fipFoldFactor, fipZFactor, adeScale1, adeOffset1, adeOffset2;
};

// formula of each iteration: maxN * 5 entries with maxN = 250 (hybrid_fractal_sequences.cpp:294-298)
constexpr int kGenSeqLength = 1250;

// Random(max) draws 0 .. max inclusive: Monte Carlo lens angles are whole radians 0 .. 65536
// (Random(65536), render_worker.cpp:1837); soft-shadow ball radii are the cube roots of
// Random(10000) / 10000 (shader_aux_shadow.cpp:83), 10001 values
constexpr int kLensAngles = 65536 + 1;
constexpr int kSoftRadii = 10000 + 1;

struct FrameConst
{
	// image and camera (render_worker.cpp:539-564, projection_3d.cpp:77-89)
	int32_t width, height;
	float invWidth, invHeight; // 1/W, 1/H (pixel -> image plane)
	float aspectRatio;
	float fov;                 // 2 tan(fov/2)
	float mRot[9];             // row-major camera rotation
	float camera[3];
	double cameraD[3];         // camera in FP64 for the precision path
	double mRotD[9];
	double fovD, aspectRatioD, invWidthD, invHeightD;
	float threshK;             // distThresh = |camera - p| * threshK   (CalcDistThresh)
	float deltaK;              // delta      = |camera - p| * deltaK    (CalcDelta)
	double threshKD, deltaKD;
	double resolutionD, detailLevelD, DEFactorD;
	int32_t constantDEThreshold;
	float DEThresh;
	double DEThreshD;
	float DEFactor;
	float searchLimit;         // 1 - 0.001 * detailLevel
	float viewDistanceMax;
	float farR2;               // march points with |p|^2 >= farR2 take the clamped step 3 without a
	                           // DE evaluation (+inf: never)
	float smoothness;
	double smoothnessD;
	// interior mode and iteration-threshold mode (general kernel only; calculate_distance.cpp:561-600,
	// render_worker.cpp:653-679, 809-812, 913-933)
	int32_t interiorMode, iterThreshMode;
	int32_t maxRaymarchingSteps;
	int32_t minN;
	// advanced quality (render_worker.cpp:671-675, 817-834): threshold and step limits
	int32_t advancedQuality;
	float absMaxStep, absMinStep, relMaxStep, relMinStep, detailSizeMax, detailSizeMin;
	double absMaxStepD, absMinStepD, relMaxStepD, relMinStepD, detailSizeMaxD, detailSizeMinD;
	double deltaRel; // delta-DE offset per detail size: 1e-5, or deltade_relative_delta with advanced quality

	// fractal sequence (hybrid_fractal_sequences.cpp:181-287, 377-503)
	int32_t deKind;
	int32_t formulaMaxiter;
	int32_t deFunction;        // delta DE finish: 1 linear, 2 logarithmic
	int32_t analyticFunction;  // analytic finish: 1 linear, 2 log, 3 IFS
	int32_t addC;
	int32_t julia;
	float bailout;
	float juliaC[3];           // juliaConstant * constantMultiplier
	float constMul[3];
	float initialW;
	// Mandelbulb
	float power, alpha, beta;  // radians
	// Mandelbox (after RecalculateFractalParams, fractal.cpp:673-675)
	float mbScale, mbAbsScale, mbFoldingLimit, mbFoldingValue, mbFR2, mbMR2, mbFactor1;
	float mbOffset[3];
	// the Mandelbox orbit runs in FP64 (full-rate FMA on gfx950): its constants in FP64
	double mbScaleD, mbAbsScaleD, mbFoldingLimitD, mbFoldingValueD, mbFR2D, mbMR2D, mbFactor1D;
	double mbOffsetD[3];
	double mbMainRotD[9];  // Mandelbox main rotation (SetRotation2 of mandelbox_rotation_main)
	int32_t mbMainRotOn;   // mandelbox_main_rotation_enabled
	int32_t mbPad;
	double juliaCD[3], constMulD[3];
	double bailoutD;
	double powerD, alphaD, betaD; // Mandelbulb in FP64 (radians)
	int32_t deltaTangent; // delta DE by forward-mode tangents instead of the reference's differences
	int32_t mbRotationsEnabled, mbMainRotationEnabled;
	float mbMainRot[9];
	float mbRot[2][3][9];
	float mbRotInv[2][3][9];
	// IFS
	float ifsScale, ifsAbsScale;
	float ifsOffset[3];
	float ifsEdge[3];
	int32_t ifsAbs[3];
	int32_t ifsMenger, ifsRotationEnabled, ifsEdgeEnabled, ifsEnabledCount;
	float ifsMainRot[9];
	float ifsRot[MBX_IFS_VECTOR_COUNT][9];
	float ifsDirection[MBX_IFS_VECTOR_COUNT][3];
	float ifsDistance[MBX_IFS_VECTOR_COUNT];
	float ifsIntensity[MBX_IFS_VECTOR_COUNT];
	int32_t ifsEnabled[MBX_IFS_VECTOR_COUNT];

	// light #1, directional (light.cpp:80-195, shader_light_shading.cpp:40-122)
	int32_t lightEnabled, castShadows, penetrating, softShadows;
	float lightDir[3];
	double lightDirD[3];
	float lightIntensity, lightVisibility, lightSize, lightContourSharpness, softRange;
	float lightColor[3];
	float viewDistanceMaxLight;

	// material #1
	float surfaceColor[3];
	float shading;
	float luminosity[3];       // luminosity * luminosityColor
	float fillLight[3];
	float lumEmissive[3];      // luminosity * luminosityColor * luminosityEmissive (GI bounces)

	// palette colouring (shader_surface_color.cpp:39-108, compute_fractal.cpp:343-436,
	// fractal_coloring.cpp:318-340, color_gradient.cpp:128-229)
	int32_t palette;           // use_colors_from_palette
	int32_t paletteGradient;   // surface_gradient_enable
	int32_t colorAlgorithm;    // 0 standard, 1 z.p, 2 sphere, 3 cross, 4 line
	int32_t colorPreV215, color4D;
	int32_t gradientCount;     // 2 .. MBX_MAX_GRADIENT_STOPS, positions non-decreasing
	double coloringSpeedD, paletteOffsetD; // the material's float values, widened
	double colorSphereRadiusD;
	double colorLineDirD[3];
	float colorSphereRadius;
	float colorLineDir[3];
	double mbColorD[3], mbSp1D, mbSp2D, mbColorRD; // Mandelbox fold colour factors
	float gradientPos[MBX_MAX_GRADIENT_STOPS];
	float gradientColor[MBX_MAX_GRADIENT_STOPS][3]; // 8-bit channel values

	// specular highlights (shader_specular_highlight_combined.cpp:39-68,
	// shader_specular_highlight.cpp:38-73)
	int32_t specPlastic, specMetal;
	float specPlasticExp, specMetalExp; // 30 / width
	float specPlasticK, specMetalK, specRoughness;
	float specColor[3];

	// glow (shader_volumetric.cpp:86-92, 196-204)
	int32_t glow;
	float glowIntensity;
	float glowColor1[3], glowColor2[3];
	// fog along the primary march (shader_volumetric.cpp:228-241, 325-356, 556-590): basic fog
	// and the volumetric distance fog, lit by (1, 1, 1)
	int32_t fog;           // either fog is on: the kernel's fog variant accumulates per march step
	int32_t basicFog, volFog;
	float glowStepOpacity; // a glow blend's opacity, glowIntensity / 512 * DEFactor clamped to 1
	double basicFogVisibilityD;
	float basicFogColor[3];
	float volFogDensity;
	double volFogFactorD, volFogFromSurfaceD, volFogColour1DistD, volFogColour2DistD;
	float volFogColour[3][3];
	// projection (projection_3d.cpp:41-89): 0 three-point, 1 fish-eye, 2 equirectangular, 3 the
	// fish-eye cut at the hemisphere (image-plane radius fishCutD); the equirectangular image halves
	// the distance threshold (threshScaleD = 0.5, render_worker.cpp:669); antialiasing grid size
	// (0 = off, render_worker.cpp:187-199)
	int32_t perspective, aaSize;
	double fishCutD, threshScaleD;

	// ambient occlusion
	int32_t aoEnabled, aoMode, aoQuality, aoCount;
	float aoIntensity;
	float aoColor[3];
	float aoFastTune;
	float aoVectors[MBX_MAX_AO_VECTORS][3];
	double aoVectorsD[MBX_MAX_AO_VECTORS][3]; // the same in FP64 for the FP64 kinds
	float aoColors[MBX_MAX_AO_VECTORS][3];

	// background (shader_background.cpp:39-200)
	int32_t background3Colors;
	float bg1[3], bg2[3], bg3[3];
	float bgBrightness, bgInvGamma;

	// Monte Carlo
	int32_t monteCarlo, samples, globalIllumination, reflectionsMax;
	int32_t DOFMinSamples;  // adaptive stop: past this many samples ...
	// the Monte Carlo lens (MonteCarloDOF), chromatic aberration and soft shadows
	int32_t dofEnabled, chromaticAberration, mcSoftShadows, reservedMc;
	double dofFocusD, dofRadiusD, softRangeD;
	float caCameraDispersion;
	// the context's tables of the random draws' transcendental images: lensSinCos[2k], [2k + 1] =
	// sin k, cos k (k < kLensAngles radians) and softCbrt[k] = pow(k / 10000, 1/3) (k < kSoftRadii),
	// indexed by the draws themselves
	const double *lensSinCos;
	const double *softCbrt;
	double DOFMaxNoise01;   // ... once the noise estimate is below DOF_max_noise * 0.01
	float giRadianceLimit;

	uint32_t randomSeed;
	// rotation of the fractal node: DE points enter the formula as R p (world -> local)
	int32_t nodeRotated;
	float nodeRot[9];
	double nodeRotD[9];

	// general formula path (kGeneric): the sequence (cHybridFractalSequences::sSequence,
	// hybrid_fractal_sequences.cpp:290-503)
	int32_t genSlots, genHybrid, genSeqLength;
	int32_t genDEType;     // MBX_DE_ANALYTIC / MBX_DE_DELTA
	int32_t genDEFunction; // hybrid and delta finish: 1 linear, 2 logarithmic
	int32_t genAnalytic;   // single-formula analytic finish: 1 linear, 2 log, 3 IFS
	int32_t genColoring;   // single-formula colour index: 0 default, 1 Mandelbox, 2 IFS
	int32_t genPad;
	double genInitialW, genLinearDEOffset;
	// global foldings before every formula iteration (compute_fractal.cpp:192-202): general kernel
	int32_t boxFold, sphFold;
	double boxFoldLimitD, boxFoldValueD, sphFR2D, sphMR2D, sphFactor1D;
	// normals from the iteration counts of an 11^3 grid (shader_calculate_normals.cpp:81-109)
	int32_t slowShading;
	int32_t slowShadingN; // params->N of the pseudo-distance 1 + N - iterations
	uint8_t genSeq[kGenSeqLength];
	FractalParams gen[MBX_MAX_FRACTALS];
};

// Per-pixel seed of the Park-Miller stream (identical on host and device).
#if defined(__HIPCC__)
__host__ __device__
#endif
inline uint32_t pixel_seed(uint32_t pixel_index, uint32_t frame_seed)
{
	uint32_t h = pixel_index * 0x9E3779B1u ^ (frame_seed + 0x7F4A7C15u);
	h ^= h >> 16;
	h *= 0x7FEB352Du;
	h ^= h >> 15;
	h *= 0x846CA68Bu;
	h ^= h >> 16;
	return 1u + h % 2147483646u;
}

#if defined(__HIP__) // HIP sources only (the host .cpp files include this header as C++)
// An out-of-line function (ours, or one the compiler chose not to inline) sees every argument as
// possibly different per lane, so it would read the frame constants with vector loads into VGPRs.
// Every lane passes the same frame: taking the address from the first active lane makes it
// provably uniform, and reading it through the constant address space lets the compiler use
// scalar loads.
template <class T>
__device__ __forceinline__ const T &uniform_ref(const T &r)
{
	const uint64_t v = (uint64_t)(uintptr_t)&r;
	const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
	const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
	typedef const __attribute__((address_space(4))) T ConstT;
	ConstT *c = (ConstT *)(uintptr_t)(((uint64_t)hi << 32) | lo);
	return *(const T *)c;
}

#endif

} // namespace mbx

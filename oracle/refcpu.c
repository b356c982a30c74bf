/*
 * refcpu.c - FP64 CPU restatement of Mandelbulber2's distance-estimation raymarch path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU baseline: it is
 * compiled into oracle/liboracle.so and loaded only by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg.  The product library (libmbx.so) never links or calls it.
 *
 * Every function restates the reference C++ in plain C, in double precision (the reference
 * CPU path is FP64 throughout: CVector3/CVector4 are double, mandelbulber2/src/algebra.hpp:58,290),
 * keeping the reference's order of operations so that FP64 results agree bit for bit where
 * the same libm is used.  Each function cites the reference file:line it follows.
 *
 * Scope: one single-formula fractal node (identity transform, scale 1), formulas Mandelbulb,
 * Mandelbox and Kaleidoscopic IFS, analytic and delta DE, the RayMarching coarse loop and
 * binary search, 6-tap normals, ObjectShader with one directional light (Lambert, no
 * specular), hard AuxShadow, multi-ray and fast AO, the 3-colour background with the sun
 * disk, Monte Carlo sub-pixel sampling and the GI bounce loop.
 *
 * Pinning (see tests/): formulas against vectors produced by the reference's own formula
 * sources (oracle/_ref); Park-Miller against the reference's common_math.cpp; Compute against
 * the 2D Multibrot closed form on the z=0 plane and the Mandelbox linear regime; the march
 * against dense FP64 ray sampling; normals against a high-order finite difference of the DE;
 * camera against the closed-form orthonormal frame; AuxShadow and multi-ray AO against dense
 * brute-force walks along their rays; fast AO against its closed form without the tune; GI by
 * depth 0 leaving pixels unchanged and missed bounces adding the background; the delta DE against
 * an independent forward-mode Jacobian on affine orbits; the Mandelbox, IFS and light rotations
 * against quarter turns derived from the reference's rotation definitions (tests/test_oracle_pins.py).
 */
#include <math.h>
#include <stddef.h>

void sincos(double x, double *s, double *c); /* GNU libm */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mbx.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define M_PI_180 0.01745329251994329576923690768489

/* ------------------------------------------------------------------------------------------ */
/* vectors: CVector3 / CVector4 (mandelbulber2/src/algebra.hpp:58-260, 290-460)                */
/* ------------------------------------------------------------------------------------------ */
typedef struct { double x, y, z; } v3;
typedef struct { double x, y, z, w; } v4;

static inline v3 V3(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v4 V4(double x, double y, double z, double w) { v4 r = {x, y, z, w}; return r; }
static inline v3 v3_add(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3_sub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3_mul(v3 a, double s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline double v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double v3_len(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
/* algebra.hpp:188-195 */
static inline v3 v3_cross(v3 a, v3 v)
{
	return V3(a.y * v.z - a.z * v.y, -a.x * v.z + a.z * v.x, a.x * v.y - a.y * v.x);
}
/* algebra.hpp:204-211: multiply by the reciprocal of the length */
static inline v3 v3_normalize(v3 a)
{
	const double norm = 1.0 / v3_len(a);
	return V3(a.x * norm, a.y * norm, a.z * norm);
}
/* algebra.hpp:214-221 */
static inline int v3_notfinite(v3 a) { return !(isfinite(a.x) & isfinite(a.y) & isfinite(a.z)); }
static inline int v3_eq(v3 a, v3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

static inline v4 v4_add(v4 a, v4 b) { return V4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline v4 v4_sub(v4 a, v4 b) { return V4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
static inline v4 v4_mul(v4 a, double s) { return V4(a.x * s, a.y * s, a.z * s, a.w * s); }
static inline double v4_dot(v4 a, v4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
/* algebra.hpp:401: the 4D length includes w */
static inline double v4_len(v4 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w); }
/* algebra.hpp:444-451 */
static inline int v4_notfinite(v4 a)
{
	return !(isfinite(a.x) & isfinite(a.y) & isfinite(a.z) & isfinite(a.w));
}

/* CVector3::RotateAroundVectorByAngle (mandelbulber2/src/algebra.cpp:71-77) */
static v3 v3_rotate_around(v3 v, v3 axis, double angle)
{
	v3 r = v3_mul(v, cos(angle));
	r = v3_add(r, v3_mul(v3_cross(axis, v), sin(angle)));
	r = v3_add(r, v3_mul(v3_mul(axis, v3_dot(axis, v)), 1 - cos(angle)));
	return r;
}

/* ------------------------------------------------------------------------------------------ */
/* CRotationMatrix (mandelbulber2/src/algebra.cpp:178-410)                                     */
/* ------------------------------------------------------------------------------------------ */
typedef struct { double m[9]; int zero; } rotm; /* row-major m11..m33 */

static void rot_null(rotm *r)
{
	memset(r->m, 0, sizeof(r->m));
	r->m[0] = r->m[4] = r->m[8] = 1.0;
	r->zero = 1;
}

/* CMatrix33::operator* (algebra.cpp:153-166): matrix = matrix * rot */
static void rot_mul_right(rotm *r, const double b[9])
{
	const double *a = r->m;
	double c[9];
	c[0] = a[0] * b[0] + a[1] * b[3] + a[2] * b[6];
	c[1] = a[0] * b[1] + a[1] * b[4] + a[2] * b[7];
	c[2] = a[0] * b[2] + a[1] * b[5] + a[2] * b[8];
	c[3] = a[3] * b[0] + a[4] * b[3] + a[5] * b[6];
	c[4] = a[3] * b[1] + a[4] * b[4] + a[5] * b[7];
	c[5] = a[3] * b[2] + a[4] * b[5] + a[5] * b[8];
	c[6] = a[6] * b[0] + a[7] * b[3] + a[8] * b[6];
	c[7] = a[6] * b[1] + a[7] * b[4] + a[8] * b[7];
	c[8] = a[6] * b[2] + a[7] * b[5] + a[8] * b[8];
	memcpy(r->m, c, sizeof(c));
	r->zero = 0;
}

/* algebra.cpp:198-247 */
static void rot_x(rotm *r, double angle)
{
	if (angle != 0.0)
	{
		const double s = sin(angle), c = cos(angle);
		double b[9] = {1.0, 0, 0, 0, c, -s, 0, s, c};
		rot_mul_right(r, b);
	}
}
static void rot_y(rotm *r, double angle)
{
	if (angle != 0.0)
	{
		const double s = sin(angle), c = cos(angle);
		double b[9] = {c, 0, s, 0, 1.0, 0, -s, 0, c};
		rot_mul_right(r, b);
	}
}
static void rot_z(rotm *r, double angle)
{
	if (angle != 0.0)
	{
		const double s = sin(angle), c = cos(angle);
		double b[9] = {c, -s, 0, s, c, 0, 0, 0, 1.0};
		rot_mul_right(r, b);
	}
}
/* algebra.cpp:262-268 SetRotation(CVector3): Z(x), X(y), Y(z) */
static void rot_set1(rotm *r, v3 a) { rot_null(r); rot_z(r, a.x); rot_x(r, a.y); rot_y(r, a.z); }
/* algebra.cpp:297-303 SetRotation2: Z(z), Y(y), X(x) */
static void rot_set2(rotm *r, v3 a) { rot_null(r); rot_z(r, a.z); rot_y(r, a.y); rot_x(r, a.x); }
/* algebra.cpp:305-311 SetRotation3: Z(x), Y(y), X(z) */
static void rot_set3(rotm *r, v3 a) { rot_null(r); rot_z(r, a.x); rot_y(r, a.y); rot_x(r, a.z); }
/* algebra.cpp:398-412 */
static rotm rot_transpose(const rotm *r)
{
	rotm t;
	const double *a = r->m;
	double b[9] = {a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]};
	memcpy(t.m, b, sizeof(b));
	t.zero = 0;
	return t;
}
/* algebra.cpp:320-344, CMatrix33 * CVector3 (algebra.cpp:168-175) */
static v3 rot_apply3(const rotm *r, v3 v)
{
	if (r->zero) return v;
	const double *a = r->m;
	return V3(a[0] * v.x + a[1] * v.y + a[2] * v.z, a[3] * v.x + a[4] * v.y + a[5] * v.z,
		a[6] * v.x + a[7] * v.y + a[8] * v.z);
}
static v4 rot_apply4(const rotm *r, v4 v)
{
	if (r->zero) return v;
	v3 xyz = rot_apply3(r, V3(v.x, v.y, v.z));
	return V4(xyz.x, xyz.y, xyz.z, v.w);
}

/* ------------------------------------------------------------------------------------------ */
/* Park-Miller RNG (mandelbulber2/src/common_math.cpp:47-72)                                   */
/* ------------------------------------------------------------------------------------------ */
int mbxo_random_int(uint32_t *seed)
{
	uint32_t lo = 16807u * (*seed & 0xFFFFu);
	uint32_t hi = 16807u * (*seed >> 16);
	lo += (hi & 0x7FFFu) << 16;
	hi >>= 15;
	lo += hi;
	if (lo > 2147483647u) lo -= 2147483647u;
	*seed = lo;
	return (int)lo;
}

int mbxo_random(uint32_t *seed, int max) { return mbxo_random_int(seed) % (max + 1); }

/* Per-pixel stream seed shared with the GPU kernels: an integer mix of (pixel, frame seed)
 * folded into Park-Miller's state range [1, 2^31-2].  Project-defined (the reference shares one
 * racy global seed between threads, common_math.cpp:47). */
uint32_t mbxo_pixel_seed(uint32_t pixel_index, uint32_t frame_seed)
{
	uint32_t h = pixel_index * 0x9E3779B1u ^ (frame_seed + 0x7F4A7C15u);
	h ^= h >> 16;
	h *= 0x7FEB352Du;
	h ^= h >> 15;
	h *= 0x846CA68Bu;
	h ^= h >> 16;
	return 1u + h % 2147483646u;
}

/* ------------------------------------------------------------------------------------------ */
/* scene: sParamRender + cHybridFractalSequences::sSequence subset, derived on the host        */
/* ------------------------------------------------------------------------------------------ */
/* One fractal container of the sequence: cHybridFractalSequences::sFractalData
 * (hybrid_fractal_sequences.h:30-56) with its sFractal after RecalculateFractalParams
 * (fractal.cpp:653-678). */
typedef struct
{
	int formula;
	int addC, checkForBailout, iterations, startIteration, stopIteration;
	double weight, bailout;
	double bulbPower, bulbAlpha, bulbBeta;
	double mbScale, mbFoldingLimit, mbFoldingValue, mbFR2, mbMR2, mbFactor1, mbSp1, mbSp2;
	v4 mbOffset;
	v3 mbColorFactor;
	int mbRotationsEnabled, mbMainRotationEnabled;
	rotm mbMainRot, mbRot[2][3], mbRotInv[2][3];
	double ifsScale;
	v4 ifsOffset;
	v3 ifsEdge;
	rotm ifsMainRot, ifsRot[MBX_IFS_VECTOR_COUNT];
	v4 ifsDirection[MBX_IFS_VECTOR_COUNT];
	double ifsDistance[MBX_IFS_VECTOR_COUNT], ifsIntensity[MBX_IFS_VECTOR_COUNT];
	int ifsEnabled[MBX_IFS_VECTOR_COUNT];
	int ifsAbsX, ifsAbsY, ifsAbsZ, ifsMenger, ifsRotationEnabled, ifsEdgeEnabled;
	double scale3, aexionCadd, fipFoldFactor, fipZFactor, adeScale1, adeOffset1, adeOffset2;
	int adeEnabled, adeEnabledFalse;
} slot_t;

/* sequence length: maxN * 5 with maxN = 250 (hybrid_fractal_sequences.cpp:294-298) */
#define SEQ_LENGTH 1250

typedef struct
{
	const mbx_params *p;
	/* camera (render_worker.cpp:539-564, fractparams.cpp:145, render_job.cpp:477) */
	rotm mRot;
	v3 camera;
	double fov;        /* CalcFOV: 2 tan(deg/360 pi) */
	double resolution; /* 1 / imageHeight */
	double aspectRatio;
	int persp; /* enumPerspectiveType: 0 three-point, 1 fish-eye, 2 equirectangular, 3 fish-eye cut */
	/* sequence (cHybridFractalSequences::sSequence, hybrid_fractal_sequences.h:58-103) */
	int formula; /* the first container's formula */
	int formulaMaxiter;
	int DEType;
	int DEFunctionType;
	int DEAnalyticFunction;
	int coloringFunction;
	int isHybrid;
	int juliaEnabled;
	v3 juliaConstant;
	v3 constantMultiplier;
	double initialWAxis;
	double linearDEOffset;
	int nSlots;
	int seqLength;
	unsigned char seq[SEQ_LENGTH];
	slot_t slot[MBX_MAX_FRACTALS];
	/* light #1 (light.cpp:51-195) */
	v3 lightDirection;
	double lightIntensity, lightVisibility, lightSize, lightSoftShadowCone, lightContourSharpness;
	float lightColor[3];
	/* AO */
	int aoCount;
	/* fractal node rotation: world -> local (objects_tree.cpp:292-382, calculate_distance.cpp:296) */
	int nodeRotated;
	rotm nodeRot;
} scene_t;

/* cCameraTarget::SetCameraTargetTop (mandelbulber2/src/camera_target.cpp:86-126) */
typedef struct { v3 forward, top, right; double yaw, pitch, roll; } camtarget_t;

static double correct_angle(double angle) { return fmod(angle + 3 * M_PI, 2 * M_PI) - M_PI; }

static camtarget_t camera_target_top(v3 camera, v3 target, v3 top)
{
	camtarget_t c;
	v3 f = v3_eq(camera, target) ? V3(0.0, 1.0, 0.0) : v3_sub(target, camera);
	f = v3_normalize(f);
	c.forward = f;
	double yaw = atan2(f.y, f.x) - 0.5 * M_PI;                /* GetAlpha, algebra.hpp:212 */
	double pitch = atan2(f.z, sqrt(f.x * f.x + f.y * f.y)); /* GetBeta, algebra.hpp:213 */
	v3 t = v3_normalize(top);
	t = v3_rotate_around(t, V3(0.0, 0.0, 1.0), -yaw);
	t = v3_rotate_around(t, V3(1.0, 0.0, 0.0), -pitch);
	double roll = -atan2(t.z, t.x) + 0.5 * M_PI;
	yaw = correct_angle(yaw);
	pitch = correct_angle(pitch);
	roll = correct_angle(roll);
	v3 tv = V3(0.0, 0.0, 1.0);
	tv = v3_rotate_around(tv, V3(0.0, 1.0, 0.0), roll);
	tv = v3_rotate_around(tv, V3(1.0, 0.0, 0.0), pitch);
	tv = v3_rotate_around(tv, V3(0.0, 0.0, 1.0), yaw);
	c.top = tv;
	c.right = v3_cross(f, tv);
	c.yaw = yaw;
	c.pitch = pitch;
	c.roll = roll;
	return c;
}

static v3 v3_from(const double a[3]) { return V3(a[0], a[1], a[2]); }

/* cAbstractFractal metadata of the formulas on the path (the constructors of
 * formula/definition/fractal_<name>.cpp) */
typedef struct
{
	int id, DEType, DEFunctionType, cpixel, DEAnalyticFunction, coloringFunction;
	double defaultBailout;
} formula_meta_t;
enum { CPIXEL_ENABLED = 0, CPIXEL_DISABLED = 1, CPIXEL_ALREADY_HAS = 2 };
enum { COLORING_DEFAULT = 0, COLORING_ABOX = 1, COLORING_IFS = 2 };
enum { DE_FUNCTION_WITHOUT = 99, DE_FUNCTION_CUSTOM = 5, DE_FUNCTION_COUNT = 6 };

static const formula_meta_t kFormulas[] = {
	{MBX_FORMULA_MANDELBULB, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LOGARITHMIC, CPIXEL_ENABLED, MBX_ANALYTIC_LOGARITHMIC, COLORING_DEFAULT, 10.0},
	{MBX_FORMULA_MANDELBULB_POWER2, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LOGARITHMIC, CPIXEL_ENABLED, MBX_ANALYTIC_LOGARITHMIC, COLORING_DEFAULT, 10.0},
	{MBX_FORMULA_HYPERCOMPLEX, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LOGARITHMIC, CPIXEL_ENABLED, MBX_ANALYTIC_LOGARITHMIC, COLORING_DEFAULT, 10.0},
	{MBX_FORMULA_QUATERNION, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LOGARITHMIC, CPIXEL_ENABLED, MBX_ANALYTIC_LOGARITHMIC, COLORING_DEFAULT, 10.0},
	{MBX_FORMULA_MENGER_SPONGE, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LINEAR, CPIXEL_DISABLED, MBX_ANALYTIC_IFS, COLORING_IFS, 10.0},
	{MBX_FORMULA_MANDELBOX, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LINEAR, CPIXEL_ENABLED, MBX_ANALYTIC_LINEAR, COLORING_ABOX, 100.0},
	{MBX_FORMULA_KALEIDOSCOPIC_IFS, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LINEAR, CPIXEL_DISABLED, MBX_ANALYTIC_IFS, COLORING_IFS, 100.0},
	{MBX_FORMULA_XENODREAMBUIE, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LOGARITHMIC, CPIXEL_ENABLED, MBX_ANALYTIC_LOGARITHMIC, COLORING_DEFAULT, 10.0},
	{MBX_FORMULA_BOX_FOLD_BULB_POW2, MBX_DE_DELTA, MBX_DE_FUNCTION_LINEAR, CPIXEL_ENABLED, MBX_ANALYTIC_LINEAR, COLORING_DEFAULT, 10000.0},
	{MBX_FORMULA_AEXION, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LOGARITHMIC, CPIXEL_ALREADY_HAS, MBX_ANALYTIC_LOGARITHMIC, COLORING_DEFAULT, 10000.0},
};

static const formula_meta_t *formula_meta(int id)
{
	for (size_t k = 0; k < sizeof(kFormulas) / sizeof(kFormulas[0]); k++)
		if (kFormulas[k].id == id) return &kFormulas[k];
	return NULL;
}

/* sFractal of one container after RecalculateFractalParams (fractal.cpp:41-120, 653-678) */
static void slot_init(slot_t *sl, const mbx_fractal *f)
{
	sl->formula = f->formula;
	sl->iterations = f->formulaIterations;
	sl->startIteration = f->formulaStartIteration;
	sl->stopIteration = f->formulaStopIteration;
	sl->weight = f->formulaWeight;
	sl->checkForBailout = f->checkForBailout;
	sl->bulbPower = f->bulb.power;
	sl->bulbAlpha = f->bulb.alphaAngleOffset * M_PI_180;
	sl->bulbBeta = f->bulb.betaAngleOffset * M_PI_180;
	const mbx_fractal_mandelbox *mb = &f->mandelbox;
	sl->mbScale = mb->scale;
	sl->mbFoldingLimit = mb->foldingLimit;
	sl->mbFoldingValue = mb->foldingValue;
	sl->mbFR2 = mb->foldingSphericalFixed * mb->foldingSphericalFixed;
	sl->mbMR2 = mb->foldingSphericalMin * mb->foldingSphericalMin;
	sl->mbFactor1 = sl->mbFR2 / sl->mbMR2;
	sl->mbOffset = V4(mb->offset[0], mb->offset[1], mb->offset[2], 0.0);
	sl->mbColorFactor = v3_from(mb->colorFactor);
	sl->mbSp1 = mb->colorFactorSp1;
	sl->mbSp2 = mb->colorFactorSp2;
	sl->mbRotationsEnabled = mb->rotationsEnabled;
	sl->mbMainRotationEnabled = mb->mainRotationEnabled;
	rot_set2(&sl->mbMainRot, v3_mul(v3_from(mb->rotationMain), M_PI_180));
	for (int fold = 0; fold < MBX_MANDELBOX_FOLDS; fold++)
		for (int axis = 0; axis < 3; axis++)
		{
			rot_set2(&sl->mbRot[fold][axis], v3_mul(v3_from(mb->rotation[fold][axis]), M_PI_180));
			sl->mbRotInv[fold][axis] = rot_transpose(&sl->mbRot[fold][axis]);
		}
	const mbx_fractal_ifs *ifs = &f->IFS;
	sl->ifsScale = ifs->scale;
	sl->ifsOffset = V4(ifs->offset[0], ifs->offset[1], ifs->offset[2], 0.0);
	sl->ifsEdge = v3_from(ifs->edge);
	rot_set3(&sl->ifsMainRot, v3_mul(v3_from(ifs->rotation), M_PI_180));
	for (int i = 0; i < MBX_IFS_VECTOR_COUNT; i++)
	{
		rot_set3(&sl->ifsRot[i], v3_mul(v3_from(ifs->rotations[i]), M_PI_180));
		/* CVector4::Normalize includes w (= 0) */
		v4 d = V4(ifs->direction[i][0], ifs->direction[i][1], ifs->direction[i][2], 0.0);
		double norm = 1.0 / v4_len(d);
		sl->ifsDirection[i] = V4(d.x * norm, d.y * norm, d.z * norm, d.w * norm);
		sl->ifsDistance[i] = ifs->distance[i];
		sl->ifsIntensity[i] = ifs->intensity[i];
		sl->ifsEnabled[i] = ifs->enabled[i];
	}
	sl->ifsAbsX = ifs->absX;
	sl->ifsAbsY = ifs->absY;
	sl->ifsAbsZ = ifs->absZ;
	sl->ifsMenger = ifs->mengerSpongeMode;
	sl->ifsRotationEnabled = ifs->rotationEnabled;
	sl->ifsEdgeEnabled = ifs->edgeEnabled;
	const mbx_fractal_transform *t = &f->transform;
	sl->scale3 = t->scale3;
	sl->aexionCadd = t->aexionCadd;
	sl->fipFoldFactor = t->foldingIntPowFoldFactor;
	sl->fipZFactor = t->foldingIntPowZFactor;
	sl->adeScale1 = t->analyticDEScale1;
	sl->adeOffset1 = t->analyticDEOffset1;
	sl->adeOffset2 = t->analyticDEOffset2;
	sl->adeEnabled = t->analyticDEEnabled;
	sl->adeEnabledFalse = t->analyticDEEnabledFalse;
}

/* cHybridFractalSequences::CollectSequenceData and CreateSequence for the rendered node
 * (hybrid_fractal_sequences.cpp:181-287, 290-503): one container (hybridCount 0), or the hybrid
 * of p->fractal and p->hybridFractals[0 .. hybridCount - 2] */
static int sequence_init(scene_t *s, const mbx_params *p)
{
	const mbx_fractal *node = &p->fractal;
	const int isHybrid = p->hybridCount >= 1;
	const int n = isHybrid ? p->hybridCount : 1;
	if (n > MBX_MAX_FRACTALS) return MBX_ERR_INVALID;
	const mbx_fractal *c[MBX_MAX_FRACTALS];
	c[0] = &p->fractal;
	for (int k = 1; k < n; k++) c[k] = &p->hybridFractals[k - 1];
	s->isHybrid = isHybrid;
	s->nSlots = n;
	s->formula = c[0]->formula;
	s->formulaMaxiter = node->formulaMaxiter;
	s->juliaEnabled = node->juliaMode;
	s->juliaConstant = v3_from(node->juliaC);
	s->constantMultiplier = v3_from(node->constantFactor);
	s->initialWAxis = node->initialWAxis;
	s->linearDEOffset = p->linearDEOffset;

	/* CollectSequenceData */
	double maxBailout = 0.0;
	for (int k = 0; k < n; k++)
	{
		slot_t *sl = &s->slot[k];
		slot_init(sl, c[k]);
		if (c[k]->formula == MBX_FORMULA_NONE) continue;
		const formula_meta_t *m = formula_meta(c[k]->formula);
		if (!m) return MBX_ERR_UNSUPPORTED;
		if (!isHybrid) sl->checkForBailout = 1;
		int addc;
		if (m->cpixel == CPIXEL_ALREADY_HAS)
			addc = 0;
		else
		{
			addc = !c[k]->dontAddCConstant;
			if (m->cpixel == CPIXEL_DISABLED) addc = !addc;
		}
		sl->addC = addc;
		if (node->useDefaultBailout)
		{
			if (isHybrid)
				maxBailout = fmax(maxBailout, m->defaultBailout);
			else
				sl->bailout = m->defaultBailout;
		}
		else
			sl->bailout = node->bailout;
	}
	if (isHybrid && node->useDefaultBailout)
		for (int k = 0; k < n; k++) s->slot[k].bailout = maxBailout;

	/* CreateSequence: the formula of each iteration */
	const int repeatFromIndex = (int)fmax(0.0, fmin((double)(p->repeatFrom - 1), (double)(n - 1)));
	int fractalNo = 0, counter = 0, rapidEnd = 0, last = 0;
	s->seqLength = SEQ_LENGTH;
	for (int i = 0; i < SEQ_LENGTH; i++)
	{
		counter++;
		int search = 0;
		while ((c[fractalNo]->formula == MBX_FORMULA_NONE || i < s->slot[fractalNo].startIteration ||
					 i > s->slot[fractalNo].stopIteration) && search < n)
		{
			fractalNo++;
			if (fractalNo >= n) fractalNo = repeatFromIndex;
			search++;
		}
		if (search >= n)
		{
			rapidEnd = 1;
			break;
		}
		s->seq[i] = (unsigned char)fractalNo;
		last = i;
		if (counter >= s->slot[fractalNo].iterations)
		{
			counter = 0;
			fractalNo++;
			if (fractalNo >= n) fractalNo = repeatFromIndex;
		}
	}
	if (rapidEnd) s->seqLength = last + 1;

	/* DE type and function */
	int forceDelta = node->deltaDEMethod == 1, forceAnalytic = node->deltaDEMethod == 2;
	if (isHybrid)
	{
		s->DEType = MBX_DE_ANALYTIC;
		if (node->deltaDEFunction == 0)
		{
			int count[DE_FUNCTION_COUNT + 1];
			for (int i = 1; i <= DE_FUNCTION_COUNT; i++) count[i] = 0;
			for (int k = 0; k < n; k++)
			{
				if (c[k]->formula == MBX_FORMULA_NONE) continue; /* the none formula: no DE function */
				const formula_meta_t *m = formula_meta(c[k]->formula);
				count[m->DEFunctionType] += s->slot[k].iterations;
				if (!forceDelta && !forceAnalytic && m->DEType == MBX_DE_DELTA)
				{
					s->DEType = MBX_DE_DELTA;
					forceDelta = 1;
				}
			}
			if (count[DE_FUNCTION_CUSTOM] > 0)
				s->DEFunctionType = DE_FUNCTION_CUSTOM;
			else
			{
				int maxCount = -1;
				for (int i = 1; i <= DE_FUNCTION_COUNT; i++)
					if (count[i] > maxCount)
					{
						maxCount = count[i];
						s->DEFunctionType = i;
					}
			}
		}
		else
		{
			s->DEFunctionType = node->deltaDEFunction;
			for (int k = 0; k < n; k++)
			{
				const formula_meta_t *m = formula_meta(c[k]->formula);
				if (m && m->DEType == MBX_DE_DELTA)
				{
					s->DEType = MBX_DE_DELTA;
					break;
				}
			}
		}
		if (forceDelta) s->DEType = MBX_DE_DELTA;
		if (forceAnalytic) s->DEType = MBX_DE_ANALYTIC;
		s->coloringFunction = COLORING_DEFAULT;
	}
	else
	{
		const formula_meta_t *m = formula_meta(c[0]->formula);
		if (!m) return MBX_ERR_UNSUPPORTED;
		s->DEType = m->DEType;
		s->DEFunctionType = m->DEFunctionType;
		s->DEAnalyticFunction = m->DEAnalyticFunction;
		s->coloringFunction = m->coloringFunction;
		if (forceDelta) s->DEType = MBX_DE_DELTA;
		if (forceAnalytic) s->DEType = MBX_DE_ANALYTIC;
		if (node->deltaDEFunction != 0)
		{
			s->DEFunctionType = node->deltaDEFunction;
			s->DEAnalyticFunction = node->deltaDEFunction == MBX_DE_FUNCTION_LOGARITHMIC ? MBX_ANALYTIC_LOGARITHMIC
																							  : MBX_ANALYTIC_LINEAR;
		}
	}
	if (s->DEFunctionType != MBX_DE_FUNCTION_LINEAR && s->DEFunctionType != MBX_DE_FUNCTION_LOGARITHMIC)
		return MBX_ERR_UNSUPPORTED; /* pseudo-Kleinian, Jos-Kleinian, custom and max-axis finishes */
	return MBX_OK;
}

static inline int seq_slot(const scene_t *s, int i) { return i < s->seqLength ? s->seq[i] : s->seq[0]; }

/* Derive the per-frame constants.  Returns 0 on success, <0 on an unsupported setting. */
static int scene_init(scene_t *s, const mbx_params *p)
{
	int rc;
	memset(s, 0, sizeof(*s));
	s->p = p;
	if (p->imageWidth <= 0 || p->imageHeight <= 0) return MBX_ERR_INVALID;
	if (p->perspectiveType < 0 || p->perspectiveType > 3) return MBX_ERR_INVALID;
	if (p->antialiasingEnabled && (p->antialiasingSize < 1 || p->antialiasingSize > 10)) return MBX_ERR_INVALID;
	/* the lens of the other projections (render_worker.cpp:1845-1866) is not restated */
	if (p->perspectiveType != 0 && p->DOFMonteCarlo && p->DOFEnabled) return MBX_ERR_UNSUPPORTED;

	/* PrepareMainVectors (render_worker.cpp:539-564) */
	camtarget_t ct = camera_target_top(v3_from(p->camera), v3_from(p->target), v3_from(p->topVector));
	rot_null(&s->mRot);
	rot_z(&s->mRot, ct.yaw);
	rot_x(&s->mRot, ct.pitch);
	rot_y(&s->mRot, ct.roll);
	rot_z(&s->mRot, -(p->sweetSpotHAngle / 180.0 * M_PI));
	rot_x(&s->mRot, p->sweetSpotVAngle / 180.0 * M_PI);
	s->camera = v3_from(p->camera);
	s->persp = p->perspectiveType;
	/* CalcFOV (projection_3d.cpp:140-159) */
	s->fov = s->persp == 0 ? 2.0 * tan(p->fov / 360.0 * M_PI) : p->fov / 180.0 * M_PI;
	s->resolution = 1.0 / p->imageHeight; /* render_job.cpp:477 */
	/* the equirectangular image spans 2 x 1 of the image plane (render_worker.cpp:98-100) */
	s->aspectRatio = s->persp == 2 ? 2.0 : (double)p->imageWidth / p->imageHeight;

	rc = sequence_init(s, p);
	if (rc != MBX_OK) return rc;

	/* cLight::setParameters for a directional light (light.cpp:80-195) */
	const mbx_light *L = &p->light;
	/* light.cpp:90: rotation = deg / 180.8 * M_PI (the reference divides by 180.8);
	 * light.cpp:92-95: a directional light flips the second angle */
	const v3 rdeg = v3_from(L->rotation);
	v3 rotation = V3(rdeg.x / 180.8 * M_PI, rdeg.y / 180.8 * M_PI, rdeg.z / 180.8 * M_PI);
	rotation = V3(rotation.x * 1.0, rotation.y * -1.0, rotation.z * 1.0);
	if (L->relativePosition)
	{
		/* light.cpp:130-148 (use_target_point false) */
		v3 d = v3_mul(ct.forward, -1.0);
		d = v3_rotate_around(d, ct.forward, rotation.z);
		d = v3_rotate_around(d, ct.right, rotation.y);
		d = v3_rotate_around(d, ct.top, rotation.x);
		s->lightDirection = d;
	}
	else
	{
		/* light.cpp:181-185 */
		rotm lr;
		rot_set1(&lr, rotation);
		s->lightDirection = rot_apply3(&lr, V3(0.0, -1.0, 0.0));
	}
	s->lightIntensity = L->intensity;
	s->lightVisibility = L->visibility;
	s->lightSize = L->size;
	s->lightSoftShadowCone = L->softShadowCone / 180.0 * M_PI;
	s->lightContourSharpness = L->contourSharpness;
	for (int k = 0; k < 3; k++) s->lightColor[k] = (float)L->color[k];

	s->aoCount = p->aoVectorsCount;
	if (p->ambientOcclusionEnabled && p->ambientOcclusionMode == MBX_AO_MULTIPLE_RAYS)
	{
		if (s->aoCount < 1 || s->aoCount > MBX_MAX_AO_VECTORS) return MBX_ERR_INVALID;
	}

	/* The node's local-to-world matrix is R^T (rotation part of objects_tree.cpp:295-309 with
	 * R = SetRotation2(rotation deg * pi/180)); its inverse, the world-to-local matrix applied
	 * to every DE point (calculate_distance.cpp:296), is R. */
	s->nodeRotated = p->formulaRotation[0] != 0.0 || p->formulaRotation[1] != 0.0 ||
		p->formulaRotation[2] != 0.0;
	rot_set2(&s->nodeRot, V3(p->formulaRotation[0] * (M_PI / 180.0), p->formulaRotation[1] * (M_PI / 180.0),
							 p->formulaRotation[2] * (M_PI / 180.0)));
	return MBX_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* formulas (FormulaCode)                                                                      */
/* ------------------------------------------------------------------------------------------ */
typedef struct
{
	v4 c, const_c, old_z;
	double pos_neg, r, DE, DE0, dist, pseudoKleinianDE, actualScale, actualScaleA, color,
		colorHybrid, temp1000;
	int i;
} aux_t; /* sExtendedAux (mandelbulber2/src/fractal.h:77-104) */

/* cFractalMandelbulb::FormulaCode (formula/definition/fractal_mandelbulb.cpp:29-43) */
static void mandelbulb_iteration(const slot_t *s, v4 *z, aux_t *aux)
{
	const double th0 = asin(z->z / aux->r) + s->bulbBeta;
	const double ph0 = atan2(z->y, z->x) + s->bulbAlpha;
	double rp = pow(aux->r, s->bulbPower - 1.0);
	const double th = th0 * s->bulbPower;
	const double ph = ph0 * s->bulbPower;
	const double cth = cos(th);
	aux->DE = (rp * aux->DE) * s->bulbPower + 1.0;
	rp *= aux->r;
	z->x = cth * cos(ph) * rp;
	z->y = cth * sin(ph) * rp;
	z->z = sin(th) * rp;
}

static inline double sign_d(double v) { return (double)((0.0 < v) - (v < 0.0)); }

/* cFractalMandelbox::FormulaCode (formula/definition/fractal_mandelbox.cpp:31-107) */
static void mandelbox_iteration(const slot_t *s, v4 *zp, aux_t *aux)
{
	v4 z = *zp;
	if (s->mbRotationsEnabled)
	{
		const double colP[3] = {s->mbColorFactor.x, s->mbColorFactor.y, s->mbColorFactor.z};
		for (int dim = 0; dim < 3; dim++)
		{
			v4 zRot = rot_apply4(&s->mbRot[0][dim], z);
			double *zr = (dim == 0) ? &zRot.x : ((dim == 1) ? &zRot.y : &zRot.z);
			if (*zr > s->mbFoldingLimit)
			{
				*zr = s->mbFoldingValue - *zr;
				z = rot_apply4(&s->mbRotInv[0][dim], zRot);
				aux->color += colP[dim];
			}
			else
			{
				zRot = rot_apply4(&s->mbRot[1][dim], z);
				zr = (dim == 0) ? &zRot.x : ((dim == 1) ? &zRot.y : &zRot.z);
				if (*zr < -s->mbFoldingLimit)
				{
					*zr = -s->mbFoldingValue - *zr;
					z = rot_apply4(&s->mbRotInv[1][dim], zRot);
					aux->color += colP[dim];
				}
			}
		}
	}
	else
	{
		if (fabs(z.x) > s->mbFoldingLimit)
		{
			z.x = sign_d(z.x) * s->mbFoldingValue - z.x;
			aux->color += s->mbColorFactor.x;
		}
		if (fabs(z.y) > s->mbFoldingLimit)
		{
			z.y = sign_d(z.y) * s->mbFoldingValue - z.y;
			aux->color += s->mbColorFactor.y;
		}
		if (fabs(z.z) > s->mbFoldingLimit)
		{
			z.z = sign_d(z.z) * s->mbFoldingValue - z.z;
			aux->color += s->mbColorFactor.z;
		}
	}
	const double r2 = v4_dot(z, z);
	z = v4_add(z, s->mbOffset);
	if (r2 < s->mbMR2)
	{
		z = v4_mul(z, s->mbFactor1);
		aux->DE *= s->mbFactor1;
		aux->color += s->mbSp1;
	}
	else if (r2 < s->mbFR2)
	{
		const double tglad_factor2 = s->mbFR2 / r2;
		z = v4_mul(z, tglad_factor2);
		aux->DE *= tglad_factor2;
		aux->color += s->mbSp2;
	}
	z = v4_sub(z, s->mbOffset);
	if (s->mbMainRotationEnabled) z = rot_apply4(&s->mbMainRot, z);
	z = v4_mul(z, s->mbScale);
	aux->DE = aux->DE * fabs(s->mbScale) + 1.0;
	*zp = z;
}

/* cFractalKaleidoscopicIfs::FormulaCode (formula/definition/fractal_kaleidoscopic_ifs.cpp:30-74) */
static void kifs_iteration(const slot_t *s, v4 *zp, aux_t *aux)
{
	v4 z = *zp;
	if (s->ifsAbsX) z.x = fabs(z.x);
	if (s->ifsAbsY) z.y = fabs(z.y);
	if (s->ifsAbsZ) z.z = fabs(z.z);
	for (int i = 0; i < MBX_IFS_VECTOR_COUNT; i++)
	{
		if (s->ifsEnabled[i])
		{
			z = rot_apply4(&s->ifsRot[i], z);
			double length = v4_dot(z, s->ifsDirection[i]);
			if (length < s->ifsDistance[i])
			{
				z = v4_sub(z, v4_mul(s->ifsDirection[i],
												2.0 * (length - s->ifsDistance[i]) * s->ifsIntensity[i]));
			}
		}
	}
	if (s->ifsRotationEnabled)
		z = v4_add(rot_apply4(&s->ifsMainRot, v4_sub(z, s->ifsOffset)), s->ifsOffset);
	if (s->ifsEdgeEnabled)
	{
		if (s->ifsEdge.x > 0.0) z.x = s->ifsEdge.x - fabs(s->ifsEdge.x - z.x);
		if (s->ifsEdge.y > 0.0) z.y = s->ifsEdge.y - fabs(s->ifsEdge.y - z.y);
		if (s->ifsEdge.z > 0.0) z.z = s->ifsEdge.z - fabs(s->ifsEdge.z - z.z);
	}
	z = v4_mul(z, s->ifsScale);
	if (s->ifsMenger)
	{
		z.x -= s->ifsOffset.x * (s->ifsScale - 1.0);
		z.y -= s->ifsOffset.y * (s->ifsScale - 1.0);
		if (z.z > 0.5 * s->ifsOffset.z * (s->ifsScale - 1.0)) z.z -= s->ifsOffset.z * (s->ifsScale - 1.0);
	}
	else
	{
		z = v4_sub(z, v4_mul(s->ifsOffset, s->ifsScale - 1.0));
	}
	aux->DE *= fabs(s->ifsScale);
	*zp = z;
}

/* cFractalMandelbulbPower2::FormulaCode (fractal_mandelbulb_power2.cpp:29-44) */
static void mandelbulb_power2_iteration(v4 *z, aux_t *aux)
{
	aux->DE = aux->DE * 2.0 * aux->r;
	const double x2 = z->x * z->x, y2 = z->y * z->y, z2 = z->z * z->z;
	const double temp = 1.0 - z2 / (x2 + y2);
	const double newx = (x2 - y2) * temp;
	const double newy = 2.0 * z->x * z->y * temp;
	const double newz = -2.0 * z->z * sqrt(x2 + y2);
	z->x = newx;
	z->y = newy;
	z->z = newz;
}

/* cFractalHypercomplex::FormulaCode (fractal_hypercomplex.cpp:31-44) */
static void hypercomplex_iteration(v4 *z, aux_t *aux)
{
	aux->DE = aux->DE * 2.0 * aux->r;
	const double newx = z->x * z->x - z->y * z->y - z->z * z->z - z->w * z->w;
	const double newy = 2.0 * z->x * z->y - 2.0 * z->w * z->z;
	const double newz = 2.0 * z->x * z->z - 2.0 * z->y * z->w;
	const double neww = 2.0 * z->x * z->w - 2.0 * z->y * z->z;
	*z = V4(newx, newy, newz, neww);
}

/* cFractalQuaternion::FormulaCode (fractal_quaternion.cpp:29-40) */
static void quaternion_iteration(v4 *z, aux_t *aux)
{
	aux->DE = aux->DE * 2.0 * aux->r;
	const double newx = z->x * z->x - z->y * z->y - z->z * z->z;
	const double newy = 2.0 * z->x * z->y;
	const double newz = 2.0 * z->x * z->z;
	z->x = newx;
	z->y = newy;
	z->z = newz;
}

/* cFractalMengerSponge::FormulaCode (fractal_menger_sponge.cpp:30-71) with the fold colour
 * (fold_color_aux_color_enabled_false) off */
static void menger_sponge_iteration(const slot_t *s, v4 *zp, aux_t *aux)
{
	v4 z = V4(fabs(zp->x), fabs(zp->y), fabs(zp->z), fabs(zp->w));
	double temp;
	if (z.x < z.y)
	{
		temp = z.y;
		z.y = z.x;
		z.x = temp;
	}
	if (z.x < z.z)
	{
		temp = z.z;
		z.z = z.x;
		z.x = temp;
	}
	if (z.y < z.z)
	{
		temp = z.z;
		z.z = z.y;
		z.y = temp;
	}
	z = v4_mul(z, s->scale3);
	z.x -= 2.0;
	z.y -= 2.0;
	if (z.z > 1.0) z.z -= 2.0;
	aux->DE *= s->scale3;
	*zp = z;
}

/* cFractalXenodreambuie::FormulaCode (fractal_xenodreambuie.cpp:30-47) */
static void xenodreambuie_iteration(const slot_t *s, v4 *z, aux_t *aux)
{
	double rp = pow(aux->r, s->bulbPower - 1.0);
	aux->DE = rp * aux->DE * s->bulbPower + 1.0;
	rp *= aux->r;
	const double th = (atan2(z->y, z->x) + s->bulbBeta) * s->bulbPower;
	double ph = acos(z->z / aux->r) + s->bulbAlpha;
	if (fabs(ph) > 0.5 * M_PI) ph = sign_d(ph) * M_PI - ph;
	ph *= s->bulbPower;
	const double sph = sin(ph);
	z->x = rp * cos(th) * sph;
	z->y = rp * sin(th) * sph;
	z->z = rp * cos(ph);
}

/* cFractalBoxFoldBulbPow2::FormulaCode (fractal_box_fold_bulb_pow2.cpp:28-77) */
static void box_fold_bulb_pow2_iteration(const slot_t *s, v4 *zp, aux_t *aux)
{
	v4 z = *zp;
	const double ff = s->fipFoldFactor;
	if (fabs(z.x) > ff) z.x = sign_d(z.x) * ff * 2.0 - z.x;
	if (fabs(z.y) > ff) z.y = sign_d(z.y) * ff * 2.0 - z.y;
	if (fabs(z.z) > ff) z.z = sign_d(z.z) * ff * 2.0 - z.z;
	const double fR2_2 = 1.0, mR2_2 = 0.25;
	const double r2_2 = v4_dot(z, z);
	const double tglad_factor1_2 = fR2_2 / mR2_2;
	if (r2_2 < mR2_2)
	{
		z = v4_mul(z, tglad_factor1_2);
		aux->DE *= tglad_factor1_2;
	}
	else if (r2_2 < fR2_2)
	{
		const double tglad_factor2_2 = fR2_2 / r2_2;
		z = v4_mul(z, tglad_factor2_2);
		aux->DE *= tglad_factor2_2;
	}
	z = v4_mul(z, 2.0);
	const double x2 = z.x * z.x, y2 = z.y * z.y, z2 = z.z * z.z;
	const double temp = 1.0 - z2 / (x2 + y2);
	v4 zt;
	zt.x = (x2 - y2) * temp;
	zt.y = 2.0 * z.x * z.y * temp;
	zt.z = -2.0 * z.z * sqrt(x2 + y2);
	zt.w = z.w;
	z = zt;
	z.z *= s->fipZFactor;
	if (s->adeEnabledFalse)
		aux->DE = (aux->DE + 1.0) * 5.0 * aux->r * s->adeScale1 *
				sqrt(s->fipZFactor * s->fipZFactor + 2.0 + s->adeOffset2) + s->adeOffset1;
	*zp = z;
}

/* cFractalAexion::FormulaCode (fractal_aexion.cpp:29-60): at the first iteration it replaces
 * the constant with a folded copy of the point */
static void aexion_iteration(const slot_t *s, v4 *zp, aux_t *aux)
{
	v4 z = *zp;
	v4 temp = z;
	double t;
	if (aux->i == 0)
	{
		t = s->aexionCadd;
		temp.x = z.x + z.y + z.z;
		temp.y = -z.x - z.y + z.z;
		temp.z = -z.x + z.y - z.z;
		temp.w = z.x - z.y - z.z;
		z = V4(fabs(temp.x) + t, fabs(temp.y) + t, fabs(temp.z) + t, fabs(temp.w) + t);
		aux->const_c = z;
	}
	t = 2.0 * z.w * z.z;
	temp.x = z.x * z.x - z.y * z.y;
	temp.y = t - temp.x;
	temp.x += t;
	t = 2.0 * z.x * z.y;
	temp.z = z.z * z.z - z.w * z.w;
	temp.w = t - temp.z;
	temp.z += t;
	z = v4_add(temp, aux->const_c);
	if (s->adeEnabled)
	{
		double de1 = 1.1 * aux->r;
		de1 = de1 + (v4_len(z) / aux->r - de1) * s->adeScale1;
		aux->DE = de1 * 2.0 * aux->DE + s->adeOffset1;
	}
	*zp = z;
}

static void formula_code(const slot_t *s, v4 *z, aux_t *aux)
{
	switch (s->formula)
	{
		case MBX_FORMULA_MANDELBULB: mandelbulb_iteration(s, z, aux); break;
		case MBX_FORMULA_MANDELBULB_POWER2: mandelbulb_power2_iteration(z, aux); break;
		case MBX_FORMULA_HYPERCOMPLEX: hypercomplex_iteration(z, aux); break;
		case MBX_FORMULA_QUATERNION: quaternion_iteration(z, aux); break;
		case MBX_FORMULA_MENGER_SPONGE: menger_sponge_iteration(s, z, aux); break;
		case MBX_FORMULA_MANDELBOX: mandelbox_iteration(s, z, aux); break;
		case MBX_FORMULA_KALEIDOSCOPIC_IFS: kifs_iteration(s, z, aux); break;
		case MBX_FORMULA_XENODREAMBUIE: xenodreambuie_iteration(s, z, aux); break;
		case MBX_FORMULA_BOX_FOLD_BULB_POW2: box_fold_bulb_pow2_iteration(s, z, aux); break;
		case MBX_FORMULA_AEXION: aexion_iteration(s, z, aux); break;
		default: break;
	}
}

/* SmoothCVector (common_math.cpp:132-163): the weighted blend of a hybrid step with the
 * length interpolated separately from the direction */
static v4 smooth_cvector(v4 v1, v4 v2, double k)
{
	if (k <= 0.0) return v1;
	if (k >= 1.0) return v2;
	const double nk = 1.0 - k;
	const double lenInterp = v4_len(v1) * nk + v4_len(v2) * k;
	const v4 vt = v4_add(v4_mul(v1, nk), v4_mul(v2, k));
	const double lt = v4_len(vt);
	if (lt > 0.0) return v4_mul(V4(vt.x / lt, vt.y / lt, vt.z / lt, vt.w / lt), lenInterp);
	return v1;
}

/* z + c for the iteration's container (compute_fractal.cpp:232-281): the Julia constant, or
 * the constant of the point (aux.const_c, which keeps its w) */
static v4 add_constant(const scene_t *s, v4 z, const aux_t *aux)
{
	if (s->juliaEnabled)
		return v4_add(z, V4(s->juliaConstant.x * s->constantMultiplier.x, s->juliaConstant.y * s->constantMultiplier.y,
							 s->juliaConstant.z * s->constantMultiplier.z, 0.0));
	/* CVector4 * CVector3 keeps w (algebra.hpp:340-343) */
	return v4_add(z, V4(aux->const_c.x * s->constantMultiplier.x, aux->const_c.y * s->constantMultiplier.y,
						 aux->const_c.z * s->constantMultiplier.z, aux->const_c.w));
}

/* BoxFolding and SphericalFolding (legacy_fractal_transforrms.cpp:43-101), applied before every
 * formula iteration when enabled (compute_fractal.cpp:192-202), each followed by aux.r = |z| */
static void global_foldings(const mbx_params *p, v4 *zp, aux_t *aux)
{
	v4 z = *zp;
	if (p->boxFolding)
	{
		const double L = p->boxFoldingLimit, V = p->boxFoldingValue;
		if (z.x > L) { z.x = V - z.x; aux->color *= 0.9; }
		else if (z.x < -L) { z.x = -V - z.x; aux->color *= 0.9; }
		if (z.y > L) { z.y = V - z.y; aux->color *= 0.9; }
		else if (z.y < -L) { z.y = -V - z.y; aux->color *= 0.9; }
		if (z.z > L) { z.z = V - z.z; aux->color *= 0.9; }
		else if (z.z < -L) { z.z = -V - z.z; aux->color *= 0.9; }
		aux->r = v4_len(z);
	}
	if (p->sphericalFolding)
	{
		const double fR2 = p->sphericalFoldingOuter * p->sphericalFoldingOuter;
		const double mR2 = p->sphericalFoldingInner * p->sphericalFoldingInner;
		const double r2 = aux->r * aux->r;
		const double factor1 = fR2 / mR2;
		if (r2 < mR2)
		{
			z = v4_mul(z, factor1);
			aux->DE *= factor1;
			aux->color *= 0.9;
		}
		else if (r2 < fR2)
		{
			const double factor2 = fR2 / r2;
			z = v4_mul(z, factor2);
			aux->DE *= factor2;
			aux->color *= 0.9;
		}
		aux->r = v4_len(z);
	}
	*zp = z;
}

/* ------------------------------------------------------------------------------------------ */
/* Compute<Mode> (mandelbulber2/src/compute_fractal.cpp:67-654): one container or a hybrid    */
/* ------------------------------------------------------------------------------------------ */
enum { MODE_NORMAL = 0, MODE_DELTA1 = 1 };

typedef struct { v3 z; double distance; int iters; int maxiter; } fout_t;

static void compute(const scene_t *s, int mode, v3 point, int forcedMaxiter, int maxiterMultiplier,
	fout_t *out)
{
	v4 z = V4(point.x, point.y, point.z, 0.0);
	z.w = s->initialWAxis;
	out->maxiter = 1;
	aux_t aux;
	memset(&aux, 0, sizeof(aux));
	aux.c = z;
	aux.const_c = z;
	aux.old_z = z;
	aux.pos_neg = 1.0;
	aux.r = v4_len(z);
	aux.DE = 1.0;
	aux.DE0 = 0.0;
	aux.dist = 1000.0;
	aux.pseudoKleinianDE = 1.0;
	aux.actualScale = s->slot[s->seq[0]].mbScale;
	aux.actualScaleA = 0.0;
	aux.color = 1.0;
	aux.colorHybrid = 0.0;
	aux.temp1000 = 1000.0;

	v4 lastGoodZ = V4(0, 0, 0, 0), lastZ = V4(0, 0, 0, 0);
	int maxN = s->formulaMaxiter * maxiterMultiplier;
	if (forcedMaxiter >= 0) maxN = forcedMaxiter;
	int i;
	for (i = 0; i < maxN; i++)
	{
		lastGoodZ = lastZ;
		lastZ = z;
		const slot_t *f = &s->slot[seq_slot(s, i)];
		global_foldings(s->p, &z, &aux);
		const v4 tempZ = z;
		const double tempAuxDE = aux.DE, tempAuxColor = aux.color;
		aux.i = i;
		if (f->weight > 0.0) formula_code(f, &z, &aux);
		if (f->addC) z = add_constant(s, z, &aux);
		if (s->isHybrid && f->weight < 1.0)
		{
			const double k = f->weight, kn = 1.0 - k;
			z = smooth_cvector(tempZ, z, k);
			aux.DE = aux.DE * k + tempAuxDE * kn;
			aux.color = aux.color * k + tempAuxColor * kn;
		}
		aux.r = v4_len(z);
		if (v4_notfinite(z))
		{
			z = lastZ;
			aux.r = v4_len(z);
			out->maxiter = 1;
			break;
		}
		if (f->checkForBailout && aux.r > f->bailout)
		{
			out->maxiter = 0;
			break;
		}
		if (v4_notfinite(z))
		{
			z = lastGoodZ;
			break;
		}
	}

	if (mode == MODE_NORMAL && s->isHybrid)
	{
		/* hybrid finish by the sequence's DE function (compute_fractal.cpp:502-515) */
		if (aux.DE > 0.0)
		{
			if (s->DEFunctionType == MBX_DE_FUNCTION_LINEAR)
				out->distance = (aux.r - s->linearDEOffset) / aux.DE;
			else
				out->distance = (aux.r > 1.0) ? 0.5 * aux.r * log(aux.r) / aux.DE : 0.0;
		}
		else
			out->distance = aux.r;
	}
	else if (mode == MODE_NORMAL)
	{
		if (aux.DE > 0.0)
		{
			switch (s->DEAnalyticFunction)
			{
				case MBX_ANALYTIC_LOGARITHMIC:
					out->distance = (aux.r > 1.0) ? 0.5 * aux.r * log(aux.r) / aux.DE : 0.0;
					break;
				case MBX_ANALYTIC_LINEAR: out->distance = aux.r / aux.DE; break;
				case MBX_ANALYTIC_IFS: out->distance = (aux.r - 2.0) / aux.DE; break;
				default: out->distance = aux.r; break;
			}
		}
		else
			out->distance = aux.r;
	}
	else
		out->distance = 0.0;
	out->iters = i + 1;
	out->z = V3(z.x, z.y, z.z);
}

/* Compute<calcModeColouring> (compute_fractal.cpp:67-654 with the colouring branch :343-436)
 * followed by CalculateColorIndex for a single formula (fractal_coloring.cpp:318-340).  point is
 * the fractal-local point (sShaderInputData::GetFractalPoint); SurfaceColour calls it with
 * maxiterMultiplier 4 (shader_surface_color.cpp:51-55).  Colouring algorithm numbers follow
 * enumFractalColoring (fractal_coloring.hpp:44-52). */
static double compute_colour_index(const scene_t *s, v3 point)
{
	const mbx_params *p = s->p;
	v4 z = V4(point.x, point.y, point.z, 0.0);
	z.w = s->initialWAxis;
	const v4 startPoint = V4(point.x, point.y, point.z, z.w); /* CVector4(in.point, initialWAxisColor) */
	aux_t aux;
	memset(&aux, 0, sizeof(aux));
	aux.c = z;
	aux.const_c = z;
	aux.old_z = z;
	aux.pos_neg = 1.0;
	aux.r = v4_len(z);
	aux.DE = 1.0;
	aux.dist = 1000.0;
	aux.pseudoKleinianDE = 1.0;
	aux.actualScale = s->slot[s->seq[0]].mbScale;
	aux.color = 1.0;
	aux.temp1000 = 1000.0;
	double colorMin = 1000.0;
	const int alg = p->coloringAlgorithm;
	const v3 lineDir = v3_from(p->coloringLineDirection);
	v4 lastZ = V4(0, 0, 0, 0);
	const int maxN = s->formulaMaxiter * 4;
	for (int i = 0; i < maxN; i++)
	{
		lastZ = z;
		const slot_t *f = &s->slot[seq_slot(s, i)];
		global_foldings(p, &z, &aux);
		const v4 tempZ = z;
		const double tempAuxDE = aux.DE, tempAuxColor = aux.color;
		aux.i = i;
		if (f->weight > 0.0) formula_code(f, &z, &aux);
		if (f->addC) z = add_constant(s, z, &aux);
		if (s->isHybrid && f->weight < 1.0)
		{
			const double k = f->weight, kn = 1.0 - k;
			z = smooth_cvector(tempZ, z, k);
			aux.DE = aux.DE * k + tempAuxDE * kn;
			aux.color = aux.color * k + tempAuxColor * kn;
		}
		aux.r = v4_len(z);
		if (v4_notfinite(z))
		{
			z = lastZ;
			aux.r = v4_len(z);
			break;
		}
		v4 colorZ = z;
		if (!p->coloring4D) colorZ.w = 0.0;
		double len = 0.0;
		switch (alg)
		{
			case 0: len = v4_len(colorZ); break;
			case 1: len = fabs(v4_dot(colorZ, startPoint)); break;
			case 2: len = fabs(v4_len(v4_sub(colorZ, startPoint)) - p->coloringSphereRadius); break;
			case 3:
			{
				const double ax = fabs(colorZ.x), ay = fabs(colorZ.y), az = fabs(colorZ.z);
				len = (ax < ay) ? (ax < az ? ax : az) : (ay < az ? ay : az); /* dMin */
				if (p->coloring4D && fabs(colorZ.w) < len) len = fabs(colorZ.w); /* std::min */
				break;
			}
			case 4: len = fabs(colorZ.x * lineDir.x + colorZ.y * lineDir.y + colorZ.z * lineDir.z); break;
			default: len = aux.r; break;
		}
		const double change = v4_len(v4_sub(z, lastZ)) / aux.r;
		if (!f->checkForBailout) continue; /* the colouring test sits in the bailout block */
		const int isBox = f->formula == MBX_FORMULA_MANDELBOX;
		if (!p->coloringPreV215)
		{
			if (!isBox)
			{
				if (len < colorMin) colorMin = len;
				if (aux.r > f->bailout) break;
			}
			else if (alg == 0)
			{
				if (aux.r > 1e15 || change < 1e-15) break;
			}
			else
			{
				if (len < colorMin) colorMin = len;
				if (aux.r > f->bailout || change < 1e-15) break;
			}
		}
		else
		{
			if (!isBox || alg != 0)
			{
				if (len < colorMin) colorMin = len;
				if (aux.r > 1e15 || change < 1e-15) break;
			}
			else if (aux.r > 1e15 || change < 1e-15)
				break;
		}
	}
	if (s->isHybrid)
	{
		/* hybrid colour index (fractal_coloring.cpp:292-316) */
		colorMin = fmin(100.0, colorMin);
		const double r2 = fmin(aux.r / fabs(aux.DE), 20.0);
		return colorMin * 1000.0 + aux.color * 100.0 + r2 * 5000.0;
	}
	switch (s->coloringFunction)
	{
		case COLORING_ABOX:
			return aux.color * 100.0 + aux.r * p->fractal.mandelbox.colorFactorR / 1e13 +
				((alg != 0) ? colorMin * 1000.0 : 0.0);
		case COLORING_IFS: return colorMin * 1000.0;
		default: return colorMin * 5000.0; /* coloringFunctionDefault */
	}
}

/* ------------------------------------------------------------------------------------------ */
/* CalculateDistance for one fractal node (calculate_distance.cpp:225-516 -> 533-763)          */
/* ------------------------------------------------------------------------------------------ */
typedef struct { double distance; int iters; int totalIters; } dout_t;

static double calculate_distance(const scene_t *s, v3 point, double detailSize, int normalCalc,
	dout_t *out)
{
	const mbx_params *p = s->p;
	double distance = 0;
	const int maxiterMultiplier = (normalCalc && p->iterThreshMode) ? 5 : 1;
	fout_t fo;
	out->totalIters = 0;
	if (s->nodeRotated) point = rot_apply3(&s->nodeRot, point);
	if (s->DEType == MBX_DE_ANALYTIC)
	{
		compute(s, MODE_NORMAL, point, -1, maxiterMultiplier, &fo);
		distance = fo.distance;
		int maxiter = fo.maxiter;
		out->iters = fo.iters;
		out->totalIters += fo.iters;
		if (!p->iterThreshMode) maxiter = 0;
		if (normalCalc) maxiter = 0;
		if (maxiter) distance = 0.0;
		if (fo.iters < p->minN && distance < detailSize) distance = detailSize;
		if (p->interiorMode && !normalCalc)
		{
			if (distance < 0.5 * detailSize || fo.maxiter) distance = detailSize;
		}
		else if (p->interiorMode && normalCalc)
		{
			if (distance < 0.9 * detailSize) distance = detailSize - distance;
		}
		if (p->iterThreshMode && !normalCalc && !fo.maxiter)
		{
			if (distance < detailSize) distance = detailSize * 1.01;
		}
		if (distance < 0.0) distance = 0.0;
		if (distance > 10.0) distance = 10.0;
	}
	else
	{
		const double deltaDE = fmax(v3_len(point) * 1e-14,
			(p->advancedQuality ? p->deltaDERelativeDelta : 1e-5) * detailSize);
		compute(s, MODE_DELTA1, point, -1, maxiterMultiplier, &fo);
		const double r = v3_len(fo.z);
		v3 zFromIters = fo.z;
		(void)zFromIters;
		int outMaxiter = fo.maxiter;
		int maxiter = fo.maxiter;
		out->iters = fo.iters;
		out->totalIters += fo.iters;
		if (!p->iterThreshMode) maxiter = 0; /* limits disabled */
		const int forced = fo.iters;
		compute(s, MODE_DELTA1, v3_add(point, V3(deltaDE, 0.0, 0.0)), forced, maxiterMultiplier, &fo);
		double r2 = v3_len(fo.z);
		const double dr1 = fabs(r2 - r) / deltaDE;
		out->totalIters += fo.iters;
		compute(s, MODE_DELTA1, v3_add(point, V3(0.0, deltaDE, 0.0)), forced, maxiterMultiplier, &fo);
		r2 = v3_len(fo.z);
		const double dr2 = fabs(r2 - r) / deltaDE;
		out->totalIters += fo.iters;
		compute(s, MODE_DELTA1, v3_add(point, V3(0.0, 0.0, deltaDE)), forced, maxiterMultiplier, &fo);
		r2 = v3_len(fo.z);
		const double dr3 = fabs(r2 - r) / deltaDE;
		out->totalIters += fo.iters;
		const double dr = sqrt(dr1 * dr1 + dr2 * dr2 + dr3 * dr3);
		if (dr > 0)
		{
			if (s->DEFunctionType == MBX_DE_FUNCTION_LINEAR)
				distance = 0.5 * r / dr;
			else if (s->DEFunctionType == MBX_DE_FUNCTION_LOGARITHMIC)
				distance = 0.5 * r * log(r) / dr;
		}
		else
			distance = r;
		if (!p->iterThreshMode) outMaxiter = 0;
		if (normalCalc) outMaxiter = 0;
		if (outMaxiter) distance = 0.0;
		if (fo.iters < p->minN && distance < detailSize) distance = detailSize;
		if (p->interiorMode && !normalCalc)
		{
			if (distance < 0.5 * detailSize || maxiter) distance = detailSize;
		}
		else if (p->interiorMode && normalCalc)
		{
			if (distance < 0.9 * detailSize) distance = detailSize - distance;
		}
		if (p->iterThreshMode && !normalCalc && !maxiter)
		{
			if (distance < detailSize) distance = detailSize * 1.01;
		}
		if (distance < 0.0) distance = 0.0;
		if (distance > 10.0) distance = 10.0;
	}
	/* single node: boolean-add merge with 1e20 and absScale 1 (calculate_distance.cpp:167-204) */
	if (!(distance < 1e20)) distance = 1e20;
	out->distance = distance;
	return distance;
}

/* ------------------------------------------------------------------------------------------ */
/* per-ray context and counters                                                                */
/* ------------------------------------------------------------------------------------------ */
typedef struct
{
	const scene_t *s;
	uint32_t *seed;
	uint64_t de_evals;
	uint64_t iterations;
} ray_ctx;

static double de(ray_ctx *rc, v3 point, double detailSize, int normalCalc, int *iters)
{
	dout_t o;
	double d = calculate_distance(rc->s, point, detailSize, normalCalc, &o);
	rc->de_evals++;
	rc->iterations += (uint64_t)o.totalIters;
	if (iters) *iters = o.iters;
	return d;
}

/* CalcDistThresh (render_worker.cpp:653-679) */
static double calc_dist_thresh(const scene_t *s, v3 point)
{
	const mbx_params *p = s->p;
	double distThresh;
	if (p->iterThreshMode)
		distThresh = v3_len(v3_sub(s->camera, point)) * s->resolution * s->fov;
	else if (p->constantDEThreshold)
		distThresh = p->DEThresh;
	else
		distThresh = v3_len(v3_sub(s->camera, point)) * s->resolution * s->fov / p->detailLevel;
	if (s->persp == 2) distThresh *= 0.5; /* render_worker.cpp:669 */
	if (p->advancedQuality)
	{
		if (distThresh > p->detailSizeMax) distThresh = p->detailSizeMax;
		if (distThresh < p->detailSizeMin) distThresh = p->detailSizeMin;
	}
	distThresh /= 1.0; /* reduceDetail = 1 outside flight recording (render_job.cpp:1173-1184) */
	return distThresh;
}

/* CalcDelta (render_worker.cpp:684-690) */
static double calc_delta(const scene_t *s, v3 point)
{
	double delta = v3_len(v3_sub(s->camera, point)) * s->resolution * s->fov;
	if (s->persp == 2) delta *= 0.5f; /* render_worker.cpp:688 */
	return delta;
}

/* ------------------------------------------------------------------------------------------ */
/* RayMarching (render_worker.cpp:708-951)                                                     */
/* ------------------------------------------------------------------------------------------ */
typedef struct
{
	int found;
	double depth, distThresh, lastDist;
	v3 point;
	int steps;
	int iters;
	/* sRayMarchingInOut: buffCount and stepBuff[i].step / .point / .distance / .iters /
	 * .distThresh (render_worker.cpp:725-838); the arrays may be NULL when the caller needs no
	 * step buffer */
	int buffCount;
	double *stepStep;
	v3 *stepPoint;
	double *stepDist;
	int *stepIters;
	double *stepThresh;
} march_t;

static void ray_marching(ray_ctx *rc, v3 start, v3 direction, double minScan, double maxScan,
	int binaryEnable, march_t *out)
{
	const scene_t *s = rc->s;
	const mbx_params *p = s->p;
	v3 point = V3(1e30, 1e30, 1e30);
	int found = 0;
	double scan = minScan;
	double dist = 0;
	const double search_accuracy = 0.001 * p->detailLevel;
	const double search_limit = 1.0 - search_accuracy;
	int counter = 0;
	double step = 0.0;
	double distThresh = 0;
	v3 lastPoint;
	int deadComputationFound = 0;
	int iters = 0;
	const double dirLength = v3_len(direction);
	out->buffCount = 0;

	for (int i = 0; i < p->maxRaymarchingSteps; i++)
	{
		lastPoint = point;
		counter++;
		point = v3_add(start, v3_mul(direction, scan));
		if (v3_eq(point, lastPoint) || v3_notfinite(point))
		{
			point = lastPoint;
			found = 1;
			deadComputationFound = 1;
			break;
		}
		distThresh = calc_dist_thresh(s, point);
		dist = de(rc, point, distThresh, 0, &iters);
		/* the step data the volumetric shader reads (render_worker.cpp:788-790) */
		if (out->stepDist)
		{
			out->stepDist[i] = dist;
			out->stepIters[i] = iters;
			out->stepThresh[i] = distThresh;
		}
		if (dist < distThresh)
		{
			found = 1;
			break;
		}
		if (out->stepStep) out->stepStep[i] = step;
		if (p->interiorMode)
			step = (dist - 0.8 * distThresh) * p->DEFactor * (1.0 - mbxo_random(rc->seed, 1000) / 10000.0);
		else
			step = (dist - 0.5 * distThresh) * p->DEFactor * (1.0 - mbxo_random(rc->seed, 1000) / 10000.0);
		if (p->advancedQuality)
		{
			/* render_worker.cpp:819-830 */
			if (step > p->absMaxMarchingStep) step = p->absMaxMarchingStep;
			if (step < p->absMinMarchingStep) step = p->absMinMarchingStep;
			if (distThresh > p->absMinMarchingStep)
			{
				if (step > p->relMaxMarchingStep * distThresh) step = p->relMaxMarchingStep * distThresh;
			}
			if (step < p->relMinMarchingStep * distThresh) step = p->relMinMarchingStep * distThresh;
		}
		else if (step > 3.0)
			step = 3.0;
		if (out->stepPoint) out->stepPoint[i] = point;
		out->buffCount = i + 1;
		scan += step / dirLength;
		if (scan > maxScan) break;
	}

	point = v3_add(start, v3_mul(direction, scan));

	if (found && binaryEnable && !deadComputationFound)
	{
		step *= 0.5;
		for (int i = 0; i < 30; i++)
		{
			counter++;
			if (dist < distThresh && dist > distThresh * search_limit)
				break;
			else
			{
				if (dist > distThresh)
				{
					scan += step;
					point = v3_add(start, v3_mul(direction, scan));
				}
				else if (dist < distThresh * search_limit)
				{
					scan -= step;
					point = v3_add(start, v3_mul(direction, scan));
				}
			}
			distThresh = calc_dist_thresh(s, point);
			dist = de(rc, point, distThresh, 0, &iters);
			step *= 0.5;
		}
	}
	if (p->iterThreshMode)
	{
		scan -= distThresh;
		point = v3_add(start, v3_mul(direction, scan));
		if (found) dist = de(rc, point, distThresh, 0, &iters);
	}
	out->found = found;
	out->lastDist = dist;
	out->depth = scan;
	out->distThresh = distThresh;
	out->point = point;
	out->steps = counter;
	out->iters = iters;
}

/* ------------------------------------------------------------------------------------------ */
/* shaders                                                                                     */
/* ------------------------------------------------------------------------------------------ */
typedef struct { float R, G, B, A; } rgba;
static rgba RGBA(float r, float g, float b, float a) { rgba c = {r, g, b, a}; return c; }

typedef struct
{
	v3 point, normal, viewVector;
	double distThresh, delta, depth;
} shin_t; /* sShaderInputData subset (render_worker.hpp:57-88) */

/* CalculateNormals (shader_calculate_normals.cpp:40-130): central differences of the DE, or with
 * slow_shading the iteration-count gradient over an 11 x 11 x 11 grid of orbits round the point
 * (:81-109; Compute on the grid points themselves, without the node's transform) */
static v3 calculate_normals(ray_ctx *rc, const shin_t *in)
{
	const mbx_params *p = rc->s->p;
	v3 normal;
	if (!p->slowShading)
	{
		double delta = in->distThresh * p->smoothness;
		if (p->interiorMode) delta = in->distThresh * 0.2 * p->smoothness;
		double sx1 = de(rc, v3_add(in->point, V3(delta, 0.0, 0.0)), in->distThresh, 1, NULL);
		double sx2 = de(rc, v3_sub(in->point, V3(delta, 0.0, 0.0)), in->distThresh, 1, NULL);
		double sy1 = de(rc, v3_add(in->point, V3(0.0, delta, 0.0)), in->distThresh, 1, NULL);
		double sy2 = de(rc, v3_sub(in->point, V3(0.0, delta, 0.0)), in->distThresh, 1, NULL);
		double sz1 = de(rc, v3_add(in->point, V3(0.0, 0.0, delta)), in->distThresh, 1, NULL);
		double sz2 = de(rc, v3_sub(in->point, V3(0.0, 0.0, delta)), in->distThresh, 1, NULL);
		normal = V3(sx1 - sx2, sy1 - sy2, sz1 - sz2);
	}
	else
	{
		double delta = in->delta * p->smoothness;
		if (p->interiorMode) delta = in->distThresh * 0.2 * p->smoothness;
		normal = V3(0.0, 0.0, 0.0);
		v3 q;
		for (q.x = -1.0; q.x <= 1.0; q.x += 0.2)
			for (q.y = -1.0; q.y <= 1.0; q.y += 0.2)
				for (q.z = -1.0; q.z <= 1.0; q.z += 0.2)
				{
					const v3 point3 = v3_add(in->point, v3_mul(q, delta));
					fout_t fo;
					compute(rc->s, MODE_NORMAL, point3, -1, 1, &fo);
					const double pseudoDistance = 1 + p->N - fo.iters;
					rc->iterations += (uint64_t)fo.iters;
					normal = v3_add(normal, v3_mul(q, pseudoDistance));
				}
	}
	if (fabs(normal.x) < 1e-50 && fabs(normal.y) < 1e-50 && fabs(normal.z) < 1e-50)
		normal = V3(1.0, 0.0, 0.0);
	else
		normal = v3_normalize(normal);
	if (v3_notfinite(normal)) normal = V3(1.0, 0.0, 0.0);
	return normal;
}

/* Monte Carlo soft shadows (shader_aux_shadow.cpp:78-87): a random point of the ball of radius
 * softRange added to the light vector, uniform in volume by the cube root of the radius */
static v3 mc_soft_light(uint32_t *seed, v3 lightVector, double softRange)
{
	v3 randomVector;
	randomVector.x = mbxo_random(seed, 10000) / 5000.0 - 1.0;
	randomVector.y = mbxo_random(seed, 10000) / 5000.0 - 1.0;
	randomVector.z = mbxo_random(seed, 10000) / 5000.0 - 1.0;
	const double randomSphereRadius = pow(mbxo_random(seed, 10000) / 10000.0, 1.0 / 3.0);
	const v3 randomSphere = v3_mul(randomVector, softRange * randomSphereRadius / v3_len(randomVector));
	return v3_normalize(v3_add(lightVector, randomSphere));
}

/* AuxShadow, hard variant with one directional light (shader_aux_shadow.cpp:39-323) */
static float aux_shadow(ray_ctx *rc, const shin_t *in, double distance, v3 lightVector)
{
	const scene_t *s = rc->s;
	const mbx_params *p = s->p;
	double totalOpacity = 0.0;
	double shadowTemp = 1.0;
	const double DEFactor = p->DEFactor;
	double start = in->distThresh;
	if (p->interiorMode) start = in->distThresh * DEFactor;
	const double softRange = tan(s->lightSoftShadowCone);
	const int mcSoft = p->DOFMonteCarlo && p->monteCarloSoftShadows;
	const int bSoft = !p->iterThreshMode && !p->interiorMode && softRange > 0.0 && !mcSoft;
	if (mcSoft) lightVector = mc_soft_light(rc->seed, lightVector, softRange);
	double maxSoft = 0.0;
	int count = 0;
	double step = in->distThresh;
	for (double i = start; i < distance; i += step)
	{
		v3 point2 = v3_add(in->point, v3_mul(lightVector, i));
		const float dist_thresh = (float)in->distThresh;
		double dist = de(rc, point2, in->distThresh, 0, NULL);
		if (bSoft)
		{
			double angle = (dist - dist_thresh) / i;
			if (angle < 0) angle = 0;
			if (dist < dist_thresh) angle = 0;
			double softShadow = 1.0 - angle / softRange;
			if (p->light.penetrating) softShadow *= (distance - i) / distance;
			if (softShadow < 0) softShadow = 0;
			if (softShadow > maxSoft) maxSoft = softShadow;
		}
		shadowTemp = 1.0 - totalOpacity;
		if (dist < dist_thresh || shadowTemp < 0.0)
		{
			if (p->light.penetrating)
			{
				shadowTemp -= (distance - i) / distance;
				if (shadowTemp < 0.0) shadowTemp = 0.0;
			}
			else
				shadowTemp = 0.0;
			break;
		}
		step = fmin(dist, 1e6) * DEFactor;
		step = fmax(step, 1e-15);
		count++;
		if (count > p->maxRaymarchingSteps) break;
	}
	return bSoft ? (float)(1.0f - maxSoft) : (float)shadowTemp;
}

/* AmbientOcclusion, multiple rays (shader_ambient_occlusion.cpp:41-122) */
static rgba ambient_occlusion(ray_ctx *rc, const shin_t *in)
{
	const scene_t *s = rc->s;
	const mbx_params *p = s->p;
	rgba AO = RGBA(0, 0, 0, 1.0f);
	const double start_dist = in->distThresh;
	const double end_dist = in->delta / s->resolution;
	int start = 0, end = s->aoCount - 1;
	if (p->DOFMonteCarlo)
	{
		int randomSample = mbxo_random(rc->seed, s->aoCount - 1);
		start = randomSample;
		end = randomSample;
	}
	for (int i = start; i <= end; i++)
	{
		v3 v = v3_from(p->aoVectors[i]);
		double dist;
		double opacity = 0.0;
		double shadowTemp = 1.0;
		for (double r = start_dist; r < end_dist; r += dist * 2.0)
		{
			v3 point2 = v3_add(in->point, v3_mul(v, r));
			dist = de(rc, point2, in->distThresh, 0, NULL);
			const float dist_thresh = (float)in->distThresh;
			shadowTemp -= opacity * (end_dist - r) / end_dist;
			if (dist < dist_thresh || shadowTemp < 0.0)
			{
				shadowTemp -= (end_dist - r) / end_dist;
				if (shadowTemp < 0.0) shadowTemp = 0.0;
				break;
			}
		}
		const double intense = shadowTemp;
		AO.R += intense * (float)p->aoColors[i][0];
		AO.G += intense * (float)p->aoColors[i][1];
		AO.B += intense * (float)p->aoColors[i][2];
	}
	if (!p->DOFMonteCarlo)
	{
		AO.R /= s->aoCount;
		AO.G /= s->aoCount;
		AO.B /= s->aoCount;
	}
	return AO;
}

/* FastAmbientOcclusion (shader_fast_ambient_occlusion.cpp:40-66) */
static rgba fast_ambient_occlusion(ray_ctx *rc, const shin_t *in)
{
	const mbx_params *p = rc->s->p;
	const double delta = in->distThresh;
	float aoTemp = 0;
	const double quality = p->ambientOcclusionQuality;
	double lastDist = 1e20;
	for (int i = 1; i < quality * quality; i++)
	{
		double scan = i * i * delta;
		v3 pointTemp = v3_add(in->point, v3_mul(in->normal, scan));
		double dist = de(rc, pointTemp, in->distThresh, 0, NULL);
		if (dist > lastDist * 2) dist = lastDist * 2.0;
		lastDist = dist;
		aoTemp += 1.0 / pow(2.0, i) * (scan - p->ambientOcclusionFastTune * dist) / in->distThresh;
	}
	float ao = 1.0f - 0.2f * aoTemp;
	if (ao < 0) ao = 0;
	return RGBA(ao, ao, ao, 1.0f);
}

/* SpecularHighlight (shader_specular_highlight.cpp:38-73) without diffusion texture or diffuse
 * gradient (diffuse = 1, texDiffuse = 1).  The material's widths and roughness are floats
 * (material.h:76-80).  The roughness factor draws from the random stream (Random(1000)). */
static rgba specular_highlight(ray_ctx *rc, const shin_t *in, v3 lightVector, float specularWidth,
	float roughness)
{
	const mbx_params *p = rc->s->p;
	v3 half = v3_normalize(v3_sub(lightVector, in->viewVector));
	float shade2 = (float)v3_dot(in->normal, half);
	if (shade2 < 0.0f) shade2 = 0.0f;
	const float diffuse = 1.0f;
	shade2 = (float)pow(shade2, 30.0f / specularWidth / diffuse) / diffuse;
	if (roughness > 0.0f) shade2 = (float)(shade2 * (1.0 + mbxo_random(rc->seed, 1000) / 1000.0f * roughness));
	if (shade2 > 15.0f) shade2 = 15.0f;
	rgba spec;
	spec.R = shade2 * (float)p->specularColor[0] * (1.0f * 0.5f + 0.5f);
	spec.G = shade2 * (float)p->specularColor[1] * (1.0f * 0.5f + 0.5f);
	spec.B = shade2 * (float)p->specularColor[2] * (1.0f * 0.5f + 0.5f);
	spec.A = 0.0f;
	return spec;
}

/* SpecularHighlightCombined (shader_specular_highlight_combined.cpp:39-68): plastic highlight,
 * then the metallic one tinted by the surface colour */
static rgba specular_combined(ray_ctx *rc, const shin_t *in, v3 lightVector, rgba surfaceColor)
{
	const mbx_params *p = rc->s->p;
	rgba plastic = RGBA(0, 0, 0, 0), metal = RGBA(0, 0, 0, 0);
	if (p->specularPlasticEnable)
	{
		plastic = specular_highlight(rc, in, lightVector, (float)p->specularWidth, 0.0f);
		const float k = (float)p->specular;
		plastic.R *= k;
		plastic.G *= k;
		plastic.B *= k;
	}
	if (p->metallic)
	{
		metal = specular_highlight(rc, in, lightVector, (float)p->specularMetallicWidth,
			(float)p->specularMetallicRoughness);
		const float k = (float)p->specularMetallic;
		metal.R *= k * surfaceColor.R;
		metal.G *= k * surfaceColor.G;
		metal.B *= k * surfaceColor.B;
	}
	return RGBA(plastic.R + metal.R, plastic.G + metal.G, plastic.B + metal.B, 0.0f);
}

/* LightShading for one directional light (shader_light_shading.cpp:40-122) with its specular
 * highlight, cLight::CalculateLightVector (light.cpp:277-302) and CalculateCone (light.cpp:214:
 * 1 and white for a directional light), summed by AuxLightsShader (shader_aux_light.cpp:38-63) */
static rgba light_shading(ray_ctx *rc, const shin_t *in, rgba surfaceColor, rgba *specOut, float *shadowOut)
{
	const scene_t *s = rc->s;
	const mbx_params *p = s->p;
	*specOut = RGBA(0, 0, 0, 0);
	/* outShadow keeps RayRecursion's default (0, 0, 0, 1) when no light is enabled
	 * (shader_aux_light.cpp:47-55) */
	*shadowOut = 0.0f;
	if (!p->light.enabled) return RGBA(0, 0, 0, 0);
	const v3 lightVector = s->lightDirection;
	const double distance = p->light.penetrating ? in->delta / s->resolution : p->viewDistanceMax;
	float intensity = (float)s->lightIntensity;
	float shade = (float)v3_dot(in->normal, lightVector);
	if (shade < 0) shade = 0;
	const float matShading = (float)p->shading;
	shade = 1.0f - matShading + shade * matShading;
	shade = shade * intensity;
	if (shade > 500.0f) shade = 500.0f;
	rgba specular = specular_combined(rc, in, lightVector, surfaceColor);
	specular.R *= intensity;
	specular.G *= intensity;
	specular.B *= intensity;
	const double specularMax = (specular.R > specular.G) ? (specular.R > specular.B ? specular.R : specular.B)
		: (specular.G > specular.B ? specular.G : specular.B); /* dMax */
	float auxShadow = 1.0f;
	if (p->light.castShadows)
	{
		if (shade > 0.001f || specularMax > 0.001f)
		{
			auxShadow = aux_shadow(rc, in, distance, lightVector);
			specular.R *= auxShadow;
			specular.G *= auxShadow;
			specular.B *= auxShadow;
		}
		else
		{
			auxShadow = 0.0f;
			specular = RGBA(0, 0, 0, 0);
		}
	}
	rgba out;
	out.R = shade * s->lightColor[0] * auxShadow;
	out.G = shade * s->lightColor[1] * auxShadow;
	out.B = shade * s->lightColor[2] * auxShadow;
	out.A = 0;
	specOut->R = specular.R * s->lightColor[0];
	specOut->G = specular.G * s->lightColor[1];
	specOut->B = specular.B * s->lightColor[2];
	specOut->A = 0;
	*shadowOut = auxShadow; /* *outShadow = auxShadow (shader_light_shading.cpp:119) */
	return out;
}

/* cColorGradient::GetColorFloat without smoothing (color_gradient.cpp:128-136, 180-229): the stop
 * below the position by PaletteIterator, then linear interpolation of the 8-bit channels / 256.
 * The last stop and a zero-width interval return the 8-bit values unscaled, as the reference does. */
static rgba gradient_colour(const mbx_params *p, float position)
{
	const int n = p->surfaceGradientCount;
	int idx = 0;
	while (idx < n - 1 && position > (float)p->surfaceGradientPosition[idx + 1]) idx++;
	rgba c;
	c.A = 1.0f;
	if (idx == n - 1)
	{
		c.R = (float)p->surfaceGradientColor[idx - 1][0];
		c.G = (float)p->surfaceGradientColor[idx - 1][1];
		c.B = (float)p->surfaceGradientColor[idx - 1][2];
		return c;
	}
	const float c1[3] = {(float)p->surfaceGradientColor[idx][0], (float)p->surfaceGradientColor[idx][1],
		(float)p->surfaceGradientColor[idx][2]};
	const float c2[3] = {(float)p->surfaceGradientColor[idx + 1][0], (float)p->surfaceGradientColor[idx + 1][1],
		(float)p->surfaceGradientColor[idx + 1][2]};
	const float pos1 = (float)p->surfaceGradientPosition[idx], pos2 = (float)p->surfaceGradientPosition[idx + 1];
	if (pos2 - pos1 > 0.0f)
	{
		const float delta = (position - pos1) / (pos2 - pos1);
		const float nDelta = 1.0f - delta;
		c.R = (c1[0] * nDelta + c2[0] * delta) / 256.0f;
		c.G = (c1[1] * nDelta + c2[1] * delta) / 256.0f;
		c.B = (c1[2] * nDelta + c2[2] * delta) / 256.0f;
	}
	else
	{
		c.R = c1[0];
		c.G = c1[1];
		c.B = c1[2];
	}
	return c;
}

/* SurfaceColour for the fractal object (shader_surface_color.cpp:39-108): the palette colour
 * from the colour index of the fractal-local point, or the material colour */
static rgba surface_colour(const scene_t *s, v3 point)
{
	const mbx_params *p = s->p;
	rgba colour = RGBA((float)p->surfaceColor[0], (float)p->surfaceColor[1], (float)p->surfaceColor[2], 1.0f);
	if (!p->useColorsFromPalette) return colour;
	v3 local = s->nodeRotated ? rot_apply3(&s->nodeRot, point) : point;
	const double colorIndex = compute_colour_index(s, local);
	const double nrCol = fmod(fabs(colorIndex), 248.0 * 256.0);
	const float speed = (float)p->coloringSpeed, offset = (float)p->coloringPaletteOffset;
	const double colorPosition = fmod(nrCol / 256.0 / 10.0 * speed + offset, 1.0);
	if (p->surfaceGradientEnable) colour = gradient_colour(p, (float)colorPosition);
	return colour;
}

/* ObjectShader (shader_object.cpp:40-229) with textures, env mapping, fake lights, perlin noise,
 * iridescence and the non-surface gradients off.  Returns the shaded colour; the surface colour,
 * the specular highlight and the emissive luminosity go to the out arguments. */
static rgba object_shader(ray_ctx *rc, const shin_t *in, rgba *surfaceColour, rgba *specularOut,
	float luminosityEmissiveOut[3], float *shadowOut)
{
	const scene_t *s = rc->s;
	const mbx_params *p = s->p;
	rgba colour = surface_colour(s, in->point);
	*surfaceColour = colour;
	rgba ambient = RGBA(0, 0, 0, 0);
	if (p->ambientOcclusionEnabled)
	{
		if (p->ambientOcclusionMode == MBX_AO_FAST)
			ambient = fast_ambient_occlusion(rc, in);
		else if (p->ambientOcclusionMode == MBX_AO_MULTIPLE_RAYS)
			ambient = ambient_occlusion(rc, in);
	}
	const float aoI = (float)p->ambientOcclusion;
	rgba ambient2;
	ambient2.R = ambient.R * aoI * (float)p->ambientOcclusionColor[0];
	ambient2.G = ambient.G * aoI * (float)p->ambientOcclusionColor[1];
	ambient2.B = ambient.B * aoI * (float)p->ambientOcclusionColor[2];
	rgba auxLightsSpecular;
	rgba auxLights = light_shading(rc, in, colour, &auxLightsSpecular, shadowOut);
	rgba luminosity;
	const float lum = (float)p->luminosity;
	luminosity.R = lum * (float)p->luminosityColor[0];
	luminosity.G = lum * (float)p->luminosityColor[1];
	luminosity.B = lum * (float)p->luminosityColor[2];
	const float emissive = (float)p->luminosityEmissive;
	luminosityEmissiveOut[0] = luminosity.R * emissive;
	luminosityEmissiveOut[1] = luminosity.G * emissive;
	luminosityEmissiveOut[2] = luminosity.B * emissive;
	rgba output;
	output.R = ((float)p->fillLightColor[0] + ambient2.R) * colour.R;
	output.G = ((float)p->fillLightColor[1] + ambient2.G) * colour.G;
	output.B = ((float)p->fillLightColor[2] + ambient2.B) * colour.B;
	output.R += auxLights.R * colour.R;
	output.G += auxLights.G * colour.G;
	output.B += auxLights.B * colour.B;
	output.R += luminosity.R;
	output.G += luminosity.G;
	output.B += luminosity.B;
	output.A = 1.0f;
	*specularOut = RGBA(auxLightsSpecular.R, auxLightsSpecular.G, auxLightsSpecular.B, output.A);
	return output;
}

/* BackgroundShader, 3-colour gradient + directional-light disk (shader_background.cpp:39-200) */
static rgba background_shader(const scene_t *s, v3 viewVector)
{
	const mbx_params *p = s->p;
	float pix[3];
	if (p->background3ColorsEnable)
	{
		v3 vv = v3_normalize(V3(0.0, 0.0, 1.0));
		v3 viewVectorNorm = v3_normalize(viewVector);
		float grad = (float)(v3_dot(viewVectorNorm, vv) + 1.0f);
		for (int k = 0; k < 3; k++)
		{
			const float c1 = (float)p->backgroundColor1[k], c2 = (float)p->backgroundColor2[k],
									c3 = (float)p->backgroundColor3[k];
			if (grad < 1.0f)
			{
				float gradN = 1.0f - grad;
				pix[k] = c3 * gradN + c2 * grad;
			}
			else
			{
				float g = grad - 1.0f;
				float gradN = 1.0f - g;
				pix[k] = c2 * gradN + c1 * g;
			}
			pix[k] = (float)(pix[k] * p->backgroundBrightness);
		}
	}
	else
	{
		for (int k = 0; k < 3; k++)
			pix[k] = (float)((float)p->backgroundColor1[k] * p->backgroundBrightness);
	}
	rgba out;
	const float invGamma = (float)(1.0f / p->backgroundGamma);
	out.R = powf(pix[0], invGamma);
	out.G = powf(pix[1], invGamma);
	out.B = powf(pix[2], invGamma);
	out.A = 0.0f;
	if (p->light.enabled)
	{
		v3 viewVectorNorm = v3_normalize(viewVector);
		double intensity = -(v3_dot(viewVectorNorm, s->lightDirection) - 1.0) * 360.0 / s->lightSize;
		intensity = 1.0 / (1.0 + pow(intensity, 6.0 * s->lightContourSharpness)) * s->lightVisibility
								* s->lightIntensity;
		out.R = (float)(out.R + intensity * s->lightColor[0]);
		out.G = (float)(out.G + intensity * s->lightColor[1]);
		out.B = (float)(out.B + intensity * s->lightColor[2]);
	}
	return out;
}

/* GlobalIlumination (shader_global_illumination.cpp:41-209), volumetric mode off */
static void global_illumination(ray_ctx *rc, const shin_t *input, rgba objectColor, float outRGB[3])
{
	const scene_t *s = rc->s;
	const mbx_params *p = s->p;
	float out[3] = {0, 0, 0};
	shin_t ic = *input;
	rgba objectColorTemp = objectColor;
	rgba resultShader = RGBA(0, 0, 0, 0);
	rgba newColor = objectColor;
	double totalOpacity = 0.0;
	int finished = 0;
	const int maxDepth = p->reflectionsMax;
	for (int rayDepth = 0; rayDepth < maxDepth; rayDepth++)
	{
		v3 reflectedDirection = ic.normal;
		double randomX = (mbxo_random(rc->seed, 20000) - 10000) / 10000.0;
		double randomY = (mbxo_random(rc->seed, 20000) - 10000) / 10000.0;
		double randomZ = (mbxo_random(rc->seed, 20000) - 10000) / 10000.0;
		v3 randomVector = V3(randomX * 1.2, randomY * 1.2, randomZ * 1.2);
		v3 randomizedDirection = v3_normalize(v3_add(reflectedDirection, randomVector));
		ic.viewVector = randomizedDirection;
		double dist = 0.0;
		int found = 0;
		int stepCount = 0;
		for (double scan = ic.distThresh; scan < p->viewDistanceMax; scan += dist * p->DEFactor)
		{
			if (stepCount >= p->maxRaymarchingSteps) break;
			v3 point = v3_add(ic.point, v3_mul(randomizedDirection, scan));
			double distThresh = calc_dist_thresh(s, point);
			ic.distThresh = distThresh;
			ic.delta = distThresh;
			dist = de(rc, point, distThresh, 0, NULL);
			if (dist < distThresh)
			{
				if (scan < distThresh * 2.0)
				{
					outRGB[0] = out[0];
					outRGB[1] = out[1];
					outRGB[2] = out[2];
					return;
				}
				ic.point = point;
				ic.depth = scan;
				ic.normal = calculate_normals(rc, &ic);
				found = 1;
				break;
			}
			stepCount++;
		}
		if (found)
		{
			rgba oc, specular;
			float emissive[3];
			float shadowDummy;
			rgba objectShader = object_shader(rc, &ic, &oc, &specular, emissive, &shadowDummy);
			newColor = oc;
			resultShader.R = objectShader.R + specular.R;
			resultShader.G = objectShader.G + specular.G;
			resultShader.B = objectShader.B + specular.B;
			out[0] += (objectShader.R + specular.R + emissive[0]) * objectColorTemp.R;
			out[1] += (objectShader.G + specular.G + emissive[1]) * objectColorTemp.G;
			out[2] += (objectShader.B + specular.B + emissive[2]) * objectColorTemp.B;
		}
		else
		{
			resultShader = background_shader(s, ic.viewVector);
			finished = 1;
		}
		const double influence = fmin(fmax(1.0 - totalOpacity, 0.0), 1.0);
		const float lim = (float)p->monteCarloGIRadianceLimit;
		resultShader.R = fminf(fmaxf(resultShader.R, 0.0f), lim);
		resultShader.G = fminf(fmaxf(resultShader.G, 0.0f), lim);
		resultShader.B = fminf(fmaxf(resultShader.B, 0.0f), lim);
		out[0] = (float)(out[0] + resultShader.R * objectColorTemp.R * influence);
		out[1] = (float)(out[1] + resultShader.G * objectColorTemp.G * influence);
		out[2] = (float)(out[2] + resultShader.B * objectColorTemp.B * influence);
		totalOpacity += 0.0;
		objectColorTemp = newColor;
		if (finished || totalOpacity > 1.0) break;
	}
	outRGB[0] = out[0];
	outRGB[1] = out[1];
	outRGB[2] = out[2];
}

/* ------------------------------------------------------------------------------------------ */
/* one primary ray: RayRecursion with reflections off (render_worker.cpp:969-1819)            */
/* ------------------------------------------------------------------------------------------ */
typedef struct
{
	float rgb[3];
	double depth;
	int found;
	int iters;
	int steps;
	v3 normal;
	unsigned colour[3]; /* uchar(objectColour * 255) (render_worker.cpp:402-404) */
	float opacity;      /* recursionOut.fogOpacity (render_worker.cpp:354, 1788) */
	/* optional channels (render_worker.cpp:355-366): the march's end point, the specular
	 * highlight, light 1's outShadow and the GI term of the sample */
	v3 world;
	float spec[3];
	float shadow;
	float gi[3];
} ray_result;

/* The glow part of VolumetricShader (shader_volumetric.cpp:70-204) with every other volumetric
 * effect off and volumetric_light_DE_Factor 1: stepCount - 1 blends towards the glow colour,
 * walking the step buffer from the end, each blend taking the steps skipped before it. */
/* DistanceFogOpacity (shader_distance_fog_opacity.cpp:37-52): the float and double
 * intermediates of the reference */
static double distance_fog_opacity(double step, double distance, double fromSurface, double factor,
	float density, double *distanceShifted)
{
	*distanceShifted = fabs(distance - fromSurface) + 0.1 * fromSurface;
	const float densityTemp =
		(float)(step * factor / (*distanceShifted * *distanceShifted + factor * factor));
	double o = 0.3f * density * densityTemp / (1.0f + density * densityTemp) - 0.001;
	if (o < 0.0) o = 0.0;
	return o;
}

/* VolumetricShader (shader_volumetric.cpp:70-700) with the step buffer of the primary march,
 * walked from its last entry back to entry 1: glow (:82-92, 196-204), the volumetric distance
 * fog (:325-356, 556-571) and the basic fog (:228-241, 575-590), lit by (1, 1, 1) as without
 * shadows in fog.  Iteration fog, clouds, volumetric and visible lights, fake lights and the
 * recalculated steps of a volumetric DE factor other than 1 are outside the path.  `opacity`
 * receives opacityOut.R: the fogs' total opacity, or 0 when the walk has no entry. */
static void volumetric_shader(const scene_t *s, const march_t *m, float rgb[3], float *opacity)
{
	const mbx_params *p = s->p;
	const int stepCount = m->buffCount;
	*opacity = 0.0f;
	if (!p->glowEnabled && !p->basicFogEnabled && !p->volFogEnabled) return;
	float out[3] = {rgb[0], rgb[1], rgb[2]};
	float totalOpacity = 0.0f;
	const float glow = stepCount * (float)p->glowIntensity / 512.0f * (float)p->DEFactor;
	float glowN = 1.0f - glow;
	if (glowN < 0.0f) glowN = 0.0f;
	const float g[3] = {(float)p->glowColor1[0] * glowN + (float)p->glowColor2[0] * glow,
		(float)p->glowColor1[1] * glowN + (float)p->glowColor2[1] * glow,
		(float)p->glowColor1[2] * glowN + (float)p->glowColor2[2] * glow};
	const float fogVolDensity = (float)p->volFogDensity;
	float fc[3][3], basicColour[3];
	for (int k = 0; k < 3; k++)
	{
		basicColour[k] = (float)p->basicFogColor[k];
		for (int j = 0; j < 3; j++) fc[j][k] = (float)p->volFogColour[j][k];
	}
	double totalStep = 0.0;
	for (int index = stepCount - 1; index > 0; index--)
	{
		double step = m->stepStep[index];
		const double distance = m->stepDist[index];
		totalStep += step;
		if (totalStep < 1e-10 * calc_dist_thresh(s, m->stepPoint[index])) continue;
		step = totalStep;
		totalStep = 0.0;
		if (p->glowEnabled && stepCount > 0)
		{
			float glowOpacity = (float)(glow / stepCount * p->volumetricLightDEFactor);
			if (glowOpacity > 1.0f) glowOpacity = 1.0f;
			for (int k = 0; k < 3; k++) out[k] = glowOpacity * g[k] + (1.0f - glowOpacity) * out[k];
		}
		float basicFogOpacity = 0.0f;
		if (p->basicFogEnabled) basicFogOpacity = (float)(step / p->basicFogVisibility);
		double distFogOpacity = 0.0;
		float distFogColor[3] = {0.0f, 0.0f, 0.0f};
		if (fogVolDensity > 0.0f && p->volFogEnabled)
		{
			double shifted;
			distFogOpacity = distance_fog_opacity(step, distance, p->volFogDistanceFromSurface,
				p->volFogDistanceFactor, fogVolDensity, &shifted);
			float k = (float)(shifted / p->volFogColour1Distance);
			if (k > 1) k = 1.0f;
			float kn = 1.0f - k;
			float temp[3];
			for (int c = 0; c < 3; c++) temp[c] = fc[0][c] * kn + fc[1][c] * k;
			float k2 = (float)(shifted / p->volFogColour2Distance * k);
			if (k2 > 1) k2 = 1.0;
			kn = 1.0f - k2;
			for (int c = 0; c < 3; c++) distFogColor[c] = temp[c] * kn + fc[2][c] * k2;
		}
		if (p->volFogEnabled && distFogOpacity > 0.0)
		{
			const float light = 1.0f, ao = 0.0f;
			if (distFogOpacity > 1.0) distFogOpacity = 1.0;
			for (int c = 0; c < 3; c++)
				out[c] = (float)(distFogOpacity * distFogColor[c] * (light + ao) + (1.0f - distFogOpacity) * out[c]);
			totalOpacity = (float)(distFogOpacity + (1.0f - distFogOpacity) * totalOpacity);
		}
		if (p->basicFogEnabled && basicFogOpacity > 0.0)
		{
			const float light = 1.0f, ao = 0.0f;
			if (basicFogOpacity > 1.0f) basicFogOpacity = 1.0f;
			for (int c = 0; c < 3; c++)
				out[c] = basicFogOpacity * basicColour[c] * (light + ao) + (1.0f - basicFogOpacity) * out[c];
			totalOpacity = basicFogOpacity + (1.0f - basicFogOpacity) * totalOpacity;
		}
		if (totalOpacity > 1.0f) totalOpacity = 1.0f;
		*opacity = totalOpacity;
	}
	rgb[0] = out[0];
	rgb[1] = out[1];
	rgb[2] = out[2];
}

static void trace_primary(ray_ctx *rc, v3 start, v3 viewVector, ray_result *res)
{
	const scene_t *s = rc->s;
	const mbx_params *p = s->p;
	v3 direction = v3_normalize(viewVector);
	march_t m = {0};
	if (p->glowEnabled || p->basicFogEnabled || p->volFogEnabled)
	{
		const size_t n = (size_t)(p->maxRaymarchingSteps > 0 ? p->maxRaymarchingSteps : 0) + 2;
		m.stepStep = (double *)calloc(n, sizeof(double));
		m.stepPoint = (v3 *)calloc(n, sizeof(v3));
		m.stepDist = (double *)calloc(n, sizeof(double));
		m.stepIters = (int *)calloc(n, sizeof(int));
		m.stepThresh = (double *)calloc(n, sizeof(double));
	}
	ray_marching(rc, start, direction, 0.0, p->viewDistanceMax, 1, &m);
	shin_t in;
	in.point = m.point;
	in.distThresh = m.distThresh;
	in.delta = calc_delta(s, m.point);
	in.viewVector = direction;
	in.depth = m.depth;
	in.normal = V3(0, 0, 0);
	res->found = m.found;
	res->iters = m.iters;
	res->steps = m.steps;
	res->world = m.point;
	/* a miss: no specular, outShadow (1, 1, 1), no GI (render_worker.cpp:1606-1612) */
	res->spec[0] = res->spec[1] = res->spec[2] = 0.0f;
	res->gi[0] = res->gi[1] = res->gi[2] = 0.0f;
	res->shadow = 1.0f;
	if (m.found)
	{
		in.normal = calculate_normals(rc, &in);
		rgba objectColour, specular;
		float emissive[3];
		rgba objectShader = object_shader(rc, &in, &objectColour, &specular, emissive, &res->shadow);
		if (p->DOFMonteCarlo && p->DOFMonteCarloGlobalIllumination)
		{
			float gi[3];
			global_illumination(rc, &in, objectColour, gi);
			objectShader.R += gi[0];
			objectShader.G += gi[1];
			objectShader.B += gi[2];
			for (int k = 0; k < 3; k++) res->gi[k] = gi[k];
		}
		res->spec[0] = specular.R;
		res->spec[1] = specular.G;
		res->spec[2] = specular.B;
		/* resultShader = objectShader + specular, clamped at 0 (render_worker.cpp:1458-1602) */
		float R = objectShader.R + specular.R, G = objectShader.G + specular.G, B = objectShader.B + specular.B;
		if (R < 0.0f) R = 0.0f;
		if (G < 0.0f) G = 0.0f;
		if (B < 0.0f) B = 0.0f;
		res->rgb[0] = R;
		res->rgb[1] = G;
		res->rgb[2] = B;
		res->depth = m.depth;
		res->normal = in.normal;
		/* the colour buffer: uchar(objectColour * 255), a float product converted to an integer
		 * whose low byte is kept */
		res->colour[0] = (unsigned char)(int)(objectColour.R * 255);
		res->colour[1] = (unsigned char)(int)(objectColour.G * 255);
		res->colour[2] = (unsigned char)(int)(objectColour.B * 255);
	}
	else
	{
		rgba bg = background_shader(s, direction);
		res->rgb[0] = bg.R;
		res->rgb[1] = bg.G;
		res->rgb[2] = bg.B;
		res->depth = 1e20;
		res->normal = V3(0, 0, 0);
		/* a miss keeps the colour the sample started with, (0, 0, 0) (render_worker.cpp:312) */
		res->colour[0] = res->colour[1] = res->colour[2] = 0;
	}
	/* VolumetricShader on the result, outside any object (render_worker.cpp:1775-1779) */
	volumetric_shader(s, &m, res->rgb, &res->opacity);
	free(m.stepStep);
	free(m.stepPoint);
	free(m.stepDist);
	free(m.stepIters);
	free(m.stepThresh);
}

/* CalculateViewVector (projection_3d.cpp:41-89): three-point, fish-eye (and its cut form) and
 * equirectangular projections of the image point, then the camera rotation */
static v3 view_vector(const scene_t *s, double ix, double iy)
{
	v3 v;
	if (s->persp == 1 || s->persp == 3)
	{
		/* fish-eye: the angle from the axis grows with the distance from the image centre */
		const double r = sqrt(ix * ix + iy * iy);
		if (r == 0.0)
			v = V3(0.0, 1.0, 0.0);
		else
			v = V3(ix / r * sin(r * s->fov), cos(r * s->fov), iy / r * sin(r * s->fov));
	}
	else if (s->persp == 2)
	{
		/* equirectangular: longitude and latitude linear in the image point */
		const double vx = ix * 0.5, vy = iy * 0.5;
		v = V3(sin(s->fov * vx) * cos(s->fov * vy), cos(s->fov * vx) * cos(s->fov * vy), sin(s->fov * vy));
	}
	else
		v = V3(ix * s->fov, 1.0, iy * s->fov);
	v = v3_normalize(v);
	return rot_apply3(&s->mRot, v);
}

/* MonteCarloDOFNoiseEstimation (render_worker.cpp:1871-1901): the standard error of the pixel's
 * samples so far, relative to its mean brightness when that exceeds 1; sRGBFloat channels are
 * float, the estimate double */
static double mc_noise(const float pixel[3], int repeat, const float sum[3], float stdDevSum[3])
{
	for (int c = 0; c < 3; c++)
	{
		const float avg = sum[c] / (float)(repeat + 1);
		const float d = (pixel[c] - avg) * (pixel[c] - avg);
		stdDevSum[c] += d;
	}
	double total = sqrt((double)(stdDevSum[0] + stdDevSum[1] + stdDevSum[2]) / (repeat + 1.0));
	total /= sqrt((double)(repeat + 1));
	const double brightness = (double)(sum[0] + sum[1] + sum[2]) / (repeat + 1.0);
	if (brightness > 1.0) total /= brightness;
	return total;
}

/* PrepareAOVectors (render_worker.cpp:599-639): latitude rows b in steps of 1/quality, longitude
 * a in steps of (2/quality)/cos(b); direction (cos(a+b) cos b, sin(a+b) cos b, sin b) turned by
 * mRotAmbientOcclusionLightMapRotation (SetRotation of (angle, 0, 0): about z,
 * fractparams.cpp:62-64); colour from lightmap pixel X = int((a+b)/2pi W + 8.5 W) % W,
 * Y = int(b/pi H + 8.5 H) % H; directions whose colour is at most 0.001 in every channel are
 * dropped; none left: one zero direction with no colour.  tex NULL: cTexture's 5 x 5 white
 * substitute (texture.cpp:150-160). */
int mbxo_prepare_ao_vectors(mbx_params *p, const float *tex, int w, int h)
{
	static const float white[3] = {1.0f, 1.0f, 1.0f};
	const int W = tex ? w : 5, H = tex ? h : 5;
	rotm r;
	v3 angles = V3(p->aoLightMapRotation / 180.0 * M_PI, 0.0, 0.0);
	rot_set1(&r, angles);
	int n = 0;
	const double quality = p->ambientOcclusionQuality;
	/* sin and cos of b, and of a + b, each from one sincos call: GCC compiles the reference's
	 * loop so (-O2 and the release -O3 -ffast-math alike), and glibc's sincos may differ from
	 * sin / cos in the last bit */
	for (double b = -0.49 * M_PI; b < 0.49 * M_PI; b += 1.0 / quality)
	{
		double sb, cb;
		sincos(b, &sb, &cb);
		for (double a = 0.0; a < 2.0 * M_PI; a += ((2.0 / quality) / cb))
		{
			double sab, cab;
			sincos(a + b, &sab, &cab);
			v3 d = V3(cab * cb, sab * cb, sb);
			d = rot_apply3(&r, d);
			const int X = (int)((a + b) / (2.0 * M_PI) * W + W * 8.5) % W;
			const int Y = (int)(b / (M_PI)*H + H * 8.5) % H;
			const float *c = tex ? tex + 3 * ((size_t)X + (size_t)Y * W) : white;
			if (!(c[0] > 0.001 || c[1] > 0.001 || c[2] > 0.001)) continue;
			if (n >= MBX_MAX_AO_VECTORS) return MBX_ERR_UNSUPPORTED;
			p->aoVectors[n][0] = d.x;
			p->aoVectors[n][1] = d.y;
			p->aoVectors[n][2] = d.z;
			p->aoColors[n][0] = c[0];
			p->aoColors[n][1] = c[1];
			p->aoColors[n][2] = c[2];
			n++;
		}
	}
	if (n == 0)
	{
		memset(p->aoVectors[0], 0, sizeof(p->aoVectors[0]));
		memset(p->aoColors[0], 0, sizeof(p->aoColors[0]));
		n = 1;
	}
	p->aoVectorsCount = n;
	return 0;
}

/* the noise estimate after each of n samples (3 floats each) of one pixel, accumulated as the
 * pixel loop does (test probe) */
void mbxo_mc_noise(const float *pixels, int n, double *noiseOut)
{
	float acc[3] = {0, 0, 0}, stdDevSum[3] = {0, 0, 0};
	for (int k = 0; k < n; k++)
	{
		for (int c = 0; c < 3; c++) acc[c] += pixels[3 * k + c];
		noiseOut[k] = mc_noise(pixels + 3 * k, k, acc, stdDevSum);
	}
}

/* MonteCarloDOF, three-point perspective (render_worker.cpp:1832-1844): a random point of the
 * lens (radius 0.0015 DOF_radius DOF_focus sqrt(u), angle an integer number of radians) turned
 * by the camera rotation becomes the ray's origin, and the view vector loses 1/focus of it so
 * that the ray still meets the undisturbed one at the focal distance */
static void monte_carlo_dof(ray_ctx *rc, v3 *startRay, v3 *viewVector)
{
	const scene_t *s = rc->s;
	const mbx_params *p = s->p;
	const double randR = 0.0015 * p->DOFRadius * p->DOFFocus * sqrt(mbxo_random(rc->seed, 65536) / 65536.0);
	const double randAngle = mbxo_random(rc->seed, 65536);
	double sa, ca; /* one sincos call, as GCC compiles the reference's sin / cos pair */
	sincos(randAngle, &sa, &ca);
	const v3 randVector = V3(randR * sa, 0.0, randR * ca);
	const v3 randVectorRot = rot_apply3(&s->mRot, randVector);
	*viewVector = v3_sub(*viewVector, v3_mul(randVectorRot, 1.0 / p->DOFFocus));
	*startRay = v3_add(s->camera, randVectorRot);
}

/* Hsv2rgb (hsv2rgb.cpp:40-60) in float */
static void hsv2rgb(float hue, float sat, float val, float rgb[3])
{
	const float h = hue / 60.0f;
	const int i = (int)floorf(hue / 60.0f);
	const float f = h - i;
	const float pp = val * (1 - sat);
	const float q = val * (1 - (sat * f));
	const float t = val * (1 - (sat * (1 - f)));
	float r = 0, g = 0, b = 0;
	switch (i)
	{
		case 0: r = val, g = t, b = pp; break;
		case 1: r = q, g = val, b = pp; break;
		case 2: r = pp, g = val, b = t; break;
		case 3: r = pp, g = q, b = val; break;
		case 4: r = t, g = pp, b = val; break;
		case 5: r = val, g = pp, b = q; break;
		default: break;
	}
	rgb[0] = r;
	rgb[1] = g;
	rgb[2] = b;
}

void mbxo_hsv2rgb(float hue, float sat, float val, float *rgb) { hsv2rgb(hue, sat, val, rgb); }

/* the ray of one sample of a pixel at image-plane point (ox, oy) (render_worker.cpp:226-264):
 * Monte Carlo anti-aliasing jitter, the lens, chromatic aberration; tint is the sample's colour
 * factor (1 without chromatic aberration) */
static void sample_ray(ray_ctx *rc, double ox, double oy, int repeat, v3 *startRay, v3 *viewVector, float tint[3])
{
	const scene_t *s = rc->s;
	const mbx_params *p = s->p;
	const int W = p->imageWidth, H = p->imageHeight;
	double ix = ox, iy = oy;
	if (p->antialiasingEnabled)
	{
		/* the sample's place on the sub-pixel grid (render_worker.cpp:187-199): repeat / size
		 * steps x, repeat % size steps y, and with Monte Carlo samples the x step runs on past the
		 * pixel as the reference's does */
		const int aa = p->antialiasingSize;
		const int xStep = repeat / aa, yStep = repeat % aa;
		ix = ox + (double)xStep / aa / W * s->aspectRatio;
		iy = oy + (double)yStep / aa / H;
	}
	else if (p->DOFMonteCarlo)
	{
		ix = ox + ((double)mbxo_random(rc->seed, 1000) / 1000.0 - 0.5) / W * s->aspectRatio;
		iy = oy + ((double)mbxo_random(rc->seed, 1000) / 1000.0 - 0.5) / H;
	}
	*viewVector = view_vector(s, ix, iy);
	*startRay = s->camera;
	if (p->DOFMonteCarlo && p->DOFEnabled) monte_carlo_dof(rc, startRay, viewVector);
	/* chromatic aberration (render_worker.cpp:255-264): a hue per sample bends the view
	 * vector and tints the sample (:374-380) */
	tint[0] = tint[1] = tint[2] = 1.0f;
	if (p->DOFMonteCarlo && p->DOFMonteCarloChromaticAberration)
	{
		const float actualHue = (float)(mbxo_random(rc->seed, 3600) / 10.0);
		hsv2rgb(fmodf(360.0f + actualHue - 60.0f, 360.0f), 1.0f, 2.0f, tint);
		const v3 randVector = V3(0.0, (double)(actualHue / 20000.0f * (float)p->DOFMonteCarloCACameraDispersion), 0.0);
		*viewVector = v3_normalize(v3_sub(*viewVector, rot_apply3(&s->mRot, randVector)));
	}
}

/* the ray of sample `repeat` of pixel (xs, ys) from the random state *seed: origin, view vector
 * (before the march normalises it) and tint; *seed advances past the sample's draws */
int mbxo_sample_ray_rep(const mbx_params *p, int xs, int ys, int repeat, uint32_t *seed, double *start3, double *view3,
	float *tint3)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	ray_ctx rc = {&s, seed, 0, 0};
	const double ox = ((1.0 / p->imageWidth) * xs + (-0.5)) * s.aspectRatio;
	const double oy = (-1.0 / p->imageHeight) * ys + 0.5;
	v3 st, vv;
	sample_ray(&rc, ox, oy, repeat, &st, &vv, tint3);
	start3[0] = st.x, start3[1] = st.y, start3[2] = st.z;
	view3[0] = vv.x, view3[1] = vv.y, view3[2] = vv.z;
	return 0;
}

int mbxo_sample_ray(const mbx_params *p, int xs, int ys, uint32_t *seed, double *start3, double *view3, float *tint3)
{
	return mbxo_sample_ray_rep(p, xs, ys, 0, seed, start3, view3, tint3);
}

/* n successive Monte Carlo soft-shadow light vectors around the unit vector light3 for a cone of
 * tangent softRange, from the random state *seed */
void mbxo_soft_light(const double *light3, double softRange, uint32_t *seed, int n, double *out3)
{
	for (int i = 0; i < n; i++)
	{
		const v3 l = mc_soft_light(seed, V3(light3[0], light3[1], light3[2]), softRange);
		out3[3 * i] = l.x, out3[3 * i + 1] = l.y, out3[3 * i + 2] = l.z;
	}
}

/* pixel loop body of cRenderWorker::doWork (render_worker.cpp:155-525) for one pixel */
static void render_pixel(ray_ctx *rc, int xs, int ys, ray_result *out)
{
	const scene_t *s = rc->s;
	const mbx_params *p = s->p;
	const int W = p->imageWidth, H = p->imageHeight;
	/* screenRegion (0,0,W,H) -> imageRegion (-0.5,0.5,0.5,-0.5) (render_job.cpp:308-318,
	 * region.hpp:66-71) */
	double ix = (1.0 / W) * xs + (-0.5);
	double iy = (-1.0 / H) * ys + 0.5;
	ix *= s->aspectRatio;
	const double ox = ix, oy = iy;
	int repeats = p->DOFMonteCarlo ? p->DOFSamples : 1;
	if (repeats < 1) repeats = 1;
	if (p->antialiasingEnabled) repeats *= p->antialiasingSize * p->antialiasingSize;
	/* the fish-eye cut renders nothing outside the hemisphere (render_worker.cpp:165-170): no
	 * march, a black sample of depth 1e20 and fog opacity 1 */
	const int hemisphereCut = s->persp == 3 && sqrt(ox * ox + oy * oy) > M_PI * 0.5f / s->fov;
	float acc[3] = {0, 0, 0};
	float stdDevSum[3] = {0, 0, 0};
	unsigned colourSum[3] = {0, 0, 0}, opacitySum = 0;
	float giSum[3] = {0, 0, 0};
	ray_result r;
	memset(out, 0, sizeof(*out));
	for (int repeat = 0; repeat < repeats; repeat++)
	{
		v3 viewVector, startRay;
		float tint[3];
		sample_ray(rc, ox, oy, repeat, &startRay, &viewVector, tint);
		if (hemisphereCut)
		{
			memset(&r, 0, sizeof(r));
			r.depth = 1e20;
			r.opacity = 1.0f;
		}
		else
			trace_primary(rc, startRay, viewVector, &r);
		r.rgb[0] *= tint[0];
		r.rgb[1] *= tint[1];
		r.rgb[2] *= tint[2];
		acc[0] += r.rgb[0];
		acc[1] += r.rgb[1];
		acc[2] += r.rgb[2];
		if (repeat == 0)
		{
			out->depth = r.depth;
			out->found = r.found;
			out->iters = r.iters;
			out->normal = r.normal;
		}
		out->steps += r.steps;
		/* depth of the last sample is what the reference writes (render_worker.cpp:346-348) */
		out->depth = r.depth;
		/* colour and opacity buffers: per-sample bytes and ushorts summed (render_worker.cpp:
		 * 400-404, 432-435) */
		for (int k = 0; k < 3; k++) colourSum[k] += r.colour[k];
		opacitySum += (unsigned short)(r.opacity * 65535.0);
		/* the optional channels: the last sample's, GI summed (render_worker.cpp:355-366) */
		out->world = r.world;
		for (int k = 0; k < 3; k++)
		{
			out->spec[k] = r.spec[k];
			giSum[k] += r.gi[k];
		}
		out->shadow = r.shadow;
		/* adaptive Monte Carlo: stop once the noise estimate is low (render_worker.cpp:436-449) */
		if (p->DOFMonteCarlo)
		{
			const double noise = mc_noise(r.rgb, repeat, acc, stdDevSum);
			if (repeat > p->DOFMinSamples && noise < p->DOFMaxNoise * 0.01)
			{
				repeats = repeat + 1;
				break;
			}
		}
	}
	if (p->DOFMonteCarlo || p->antialiasingEnabled)
	{
		out->rgb[0] = acc[0] / repeats;
		out->rgb[1] = acc[1] / repeats;
		out->rgb[2] = acc[2] / repeats;
	}
	else
	{
		out->rgb[0] = acc[0];
		out->rgb[1] = acc[1];
		out->rgb[2] = acc[2];
	}
	/* integer means over the samples (render_worker.cpp:469-472), the GI mean (:474-476) */
	for (int k = 0; k < 3; k++) out->colour[k] = colourSum[k] / (unsigned)repeats;
	for (int k = 0; k < 3; k++)
		out->gi[k] = (p->DOFMonteCarlo || p->antialiasingEnabled) ? giSum[k] / repeats : giSum[k];
	out->opacity = (float)(opacitySum / (unsigned)repeats);
}

/* ------------------------------------------------------------------------------------------ */
/* exported entry points (ctypes)                                                              */
/* ------------------------------------------------------------------------------------------ */

/* Render rows [y0, y1) of the image.  Any output pointer may be NULL.  Row-interleaved
 * dynamic scheduling over `threads` OpenMP threads stands in for cScheduler's line hand-out
 * (mandelbulber2/src/scheduler.hpp:51-88).  Counters (rays, DE evals, iterations) are
 * accumulated into stats[0..2] when stats is non-NULL. */
int mbxo_render(const mbx_params *p, int y0, int y1, int threads, float *rgb, float *depth,
	int32_t *iters, uint8_t *mask, float *normal, int32_t *steps, uint8_t *colour, uint16_t *opacity,
	float *world, float *specular, float *shadow, float *gi, uint64_t *stats)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	const int W = p->imageWidth;
	if (y0 < 0) y0 = 0;
	if (y1 > p->imageHeight) y1 = p->imageHeight;
	uint64_t n_de = 0, n_it = 0, n_rays = 0;
	if (p->rngMode == MBX_RNG_GLOBAL_SEQUENTIAL)
	{
		uint32_t gseed = 1;
		ray_ctx rc = {&s, &gseed, 0, 0};
		for (int y = y0; y < y1; y++)
			for (int x = 0; x < W; x++)
			{
				ray_result r;
				render_pixel(&rc, x, y, &r);
				const size_t i = (size_t)y * W + x;
				if (rgb) { rgb[3 * i] = r.rgb[0]; rgb[3 * i + 1] = r.rgb[1]; rgb[3 * i + 2] = r.rgb[2]; }
				if (depth) depth[i] = (float)r.depth;
				if (iters) iters[i] = r.iters;
				if (mask) mask[i] = (uint8_t)r.found;
				if (normal) { normal[3 * i] = (float)r.normal.x; normal[3 * i + 1] = (float)r.normal.y; normal[3 * i + 2] = (float)r.normal.z; }
				if (steps) steps[i] = r.steps;
				if (colour) { colour[3 * i] = (uint8_t)r.colour[0]; colour[3 * i + 1] = (uint8_t)r.colour[1]; colour[3 * i + 2] = (uint8_t)r.colour[2]; }
				if (opacity) opacity[i] = (uint16_t)r.opacity;
				if (world) { world[3 * i] = (float)r.world.x; world[3 * i + 1] = (float)r.world.y; world[3 * i + 2] = (float)r.world.z; }
				if (specular) { specular[3 * i] = r.spec[0]; specular[3 * i + 1] = r.spec[1]; specular[3 * i + 2] = r.spec[2]; }
				if (shadow) { shadow[3 * i] = shadow[3 * i + 1] = shadow[3 * i + 2] = r.shadow; }
				if (gi) { gi[3 * i] = r.gi[0]; gi[3 * i + 1] = r.gi[1]; gi[3 * i + 2] = r.gi[2]; }
				n_rays += (uint64_t)(p->DOFMonteCarlo ? p->DOFSamples : 1);
			}
		n_de = rc.de_evals;
		n_it = rc.iterations;
	}
	else
	{
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1) \
	reduction(+ : n_de, n_it, n_rays)
		for (int y = y0; y < y1; y++)
		{
			for (int x = 0; x < W; x++)
			{
				const uint32_t pix = (uint32_t)y * (uint32_t)W + (uint32_t)x;
				uint32_t seed = mbxo_pixel_seed(pix, p->randomSeed);
				ray_ctx rc = {&s, &seed, 0, 0};
				ray_result r;
				render_pixel(&rc, x, y, &r);
				const size_t i = (size_t)y * W + x;
				if (rgb) { rgb[3 * i] = r.rgb[0]; rgb[3 * i + 1] = r.rgb[1]; rgb[3 * i + 2] = r.rgb[2]; }
				if (depth) depth[i] = (float)r.depth;
				if (iters) iters[i] = r.iters;
				if (mask) mask[i] = (uint8_t)r.found;
				if (normal) { normal[3 * i] = (float)r.normal.x; normal[3 * i + 1] = (float)r.normal.y; normal[3 * i + 2] = (float)r.normal.z; }
				if (steps) steps[i] = r.steps;
				if (colour) { colour[3 * i] = (uint8_t)r.colour[0]; colour[3 * i + 1] = (uint8_t)r.colour[1]; colour[3 * i + 2] = (uint8_t)r.colour[2]; }
				if (opacity) opacity[i] = (uint16_t)r.opacity;
				if (world) { world[3 * i] = (float)r.world.x; world[3 * i + 1] = (float)r.world.y; world[3 * i + 2] = (float)r.world.z; }
				if (specular) { specular[3 * i] = r.spec[0]; specular[3 * i + 1] = r.spec[1]; specular[3 * i + 2] = r.spec[2]; }
				if (shadow) { shadow[3 * i] = shadow[3 * i + 1] = shadow[3 * i + 2] = r.shadow; }
				if (gi) { gi[3 * i] = r.gi[0]; gi[3 * i + 1] = r.gi[1]; gi[3 * i + 2] = r.gi[2]; }
				n_de += rc.de_evals;
				n_it += rc.iterations;
				n_rays += (uint64_t)(p->DOFMonteCarlo ? p->DOFSamples : 1);
			}
		}
	}
	if (stats)
	{
		stats[0] += n_rays;
		stats[1] += n_de;
		stats[2] += n_it;
	}
	return MBX_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* screen-space AO: cRenderSSAO::RenderSSAO (render_ssao.cpp:76-200) over the whole image with   */
/* no progressive step, per pixel cSSAOWorker::doWork (ssao_worker.cpp:63-261)                  */
/* ------------------------------------------------------------------------------------------ */

/* The pixel's projected point and a sample's (ssao_worker.cpp:125-150, 192-218), in the
 * reference's operation order; `sample` selects the sample branch, whose fish-eye and
 * equirectangular forms carry the factor pi and whose fish-eye test is on the radius loop's r. */
static void ssao_point(int persp, double fov, double aspectRatio, int width, int height, double px, double py,
	double z, int sample, double r, double *ox, double *oy)
{
	double x2, y2;
	if (persp == 1 || persp == 3)
	{
		if (sample)
		{
			x2 = M_PI * (px / width - 0.5) * aspectRatio;
			y2 = M_PI * (py / height - 0.5);
			const double r2 = sqrt(x2 * x2 + y2 * y2);
			if (r != 0.0)
			{
				x2 = x2 / r2 * sin(r2 * fov) * z;
				y2 = y2 / r2 * sin(r2 * fov) * z;
			}
		}
		else
		{
			x2 = (px / width - 0.5) * aspectRatio;
			y2 = (py / height - 0.5);
			const double rr = sqrt(x2 * x2 + y2 * y2);
			if (rr != 0.0)
			{
				x2 = x2 / rr * sin(rr * fov) * z;
				y2 = y2 / rr * sin(rr * fov) * z;
			}
		}
	}
	else if (persp == 2)
	{
		x2 = M_PI * (px / width - 0.5) * aspectRatio;
		y2 = M_PI * (py / height - 0.5);
		x2 = sin(fov * x2) * cos(fov * y2) * z;
		y2 = sin(fov * y2) * z;
	}
	else if (sample)
	{
		x2 = (px / width - 0.5) * aspectRatio;
		y2 = py / height - 0.5;
		x2 = x2 * (z * fov);
		y2 = y2 * (z * fov);
	}
	else
	{
		x2 = (px / width - 0.5) * aspectRatio;
		y2 = py / height - 0.5;
		x2 = x2 * z * fov;
		y2 = y2 * z * fov;
	}
	*ox = x2;
	*oy = y2;
}

/* rgbOut = rgbIn plus the SSAO term, for the planes of one rendered frame (host memory).
 * opacity may be NULL (0).  In random mode each pixel draws from its own stream seeded by
 * (pixel index, randomSeed ^ 0x55A0A055). */
int mbxo_ssao(const mbx_params *p, const float *depth, const uint8_t *colour, const uint16_t *opacity,
	const float *rgbIn, float *rgbOut, int threads)
{
	const int width = p->imageWidth, height = p->imageHeight;
	if (width <= 0 || height <= 0 || p->ambientOcclusionQuality < 1 || p->ambientOcclusionQuality > 10 ||
		p->perspectiveType < 0 || p->perspectiveType > 3)
		return MBX_ERR_INVALID;
	int quality = (int)(p->ambientOcclusionQuality * p->ambientOcclusionQuality * 1.0);
	if (quality < 3) quality = 3;
	double *cosine = (double *)malloc(sizeof(double) * quality);
	double *sine = (double *)malloc(sizeof(double) * quality);
	for (int i = 0; i < quality; i++)
	{
		sine[i] = sin((double)i / quality * 2.0 * M_PI);
		cosine[i] = cos((double)i / quality * 2.0 * M_PI);
	}
	const int persp = p->perspectiveType;
	const double scale_factor = (double)width / (quality * quality) / 2.0;
	double aspectRatio = (double)width / height;
	if (persp == 2) aspectRatio = 2.0;
	const double fov = persp == 0 ? 2.0 * tan(p->fov / 360.0 * M_PI) : p->fov / 180.0 * M_PI;
	const float intensity = (float)p->ambientOcclusion;
	const float aoColor[3] = {(float)p->ambientOcclusionColor[0], (float)p->ambientOcclusionColor[1],
		(float)p->ambientOcclusionColor[2]};
	const int startX = 0, endX = width, startLine = 0, endLine = height;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
	for (int y = 0; y < height; y++)
	{
		for (int x = 0; x < width; x++)
		{
			const size_t i = (size_t)y * width + x;
			const double z = (double)depth[i];
			const float op = opacity ? opacity[i] / 65535.0f : 0.0f;
			float total_ambient = 0.0f;
			if (z < 1e19)
			{
				double x2, y2;
				ssao_point(persp, fov, aspectRatio, width, height, (double)(x - startX), (double)(y - startLine), z, 0,
					0.0, &x2, &y2);
				float ambient = 0.0f;
				const double angleStep = M_PI * 2.0 / (double)quality;
				const int maxRandom = 62831 / quality;
				double rRandom = 1.0;
				uint32_t seed = mbxo_pixel_seed((uint32_t)i, p->randomSeed ^ 0x55A0A055u);
				if (p->SSAORandomMode) rRandom = 0.5 + mbxo_random(&seed, 65536) / 65536.0;
				int rayCount = 0;
				for (int angleIndex = 0; angleIndex < quality; angleIndex++)
				{
					double ca, sa;
					double angle = angleIndex;
					if (p->SSAORandomMode)
					{
						angle = angleStep * angleIndex + mbxo_random(&seed, maxRandom) / 10000.0;
						ca = cos(angle);
						sa = sin(angle);
					}
					else
					{
						ca = cosine[(int)angle];
						sa = sine[(int)angle];
					}
					double max_diff = -1e50;
					int wasRay = 0;
					for (double r = 1.0; r < quality; r += rRandom)
					{
						const double rr = r * r * scale_factor;
						const double xx = x + rr * ca;
						const double yy = y + rr * sa;
						if ((int)xx == x && (int)yy == y) continue;
						if (xx < startX || xx > endX - 1 || yy < startLine || yy > endLine - 1) continue;
						const double z2 = (double)depth[(size_t)(int)yy * width + (int)xx];
						wasRay = 1;
						double xx2, yy2;
						ssao_point(persp, fov, aspectRatio, width, height, xx - startX, yy - startLine, z2, 1, r, &xx2,
							&yy2);
						const double dx = xx2 - x2;
						const double dy = yy2 - y2;
						const double dz = z2 - z;
						const double dr = sqrt(dx * dx + dy * dy);
						const double diff = -dz / dr;
						if (diff > max_diff) max_diff = diff;
					}
					const double max_angle = atan(max_diff);
					if (wasRay)
					{
						ambient += -max_angle / M_PI + 0.5;
						rayCount++;
					}
				}
				total_ambient = ambient / rayCount;
				if (total_ambient < 0) total_ambient = 0;
			}
			float px[3] = {rgbIn[3 * i], rgbIn[3 * i + 1], rgbIn[3 * i + 2]};
			if (!(x >= endX - 1))
			{
				const float shadeFactor = 1.0f / 256.0f * total_ambient * intensity * (1.0f - op);
				for (int k = 0; k < 3; k++) px[k] = px[k] + colour[3 * i + k] * shadeFactor * aoColor[k];
			}
			rgbOut[3 * i] = px[0];
			rgbOut[3 * i + 1] = px[1];
			rgbOut[3 * i + 2] = px[2];
		}
	}
	free(cosine);
	free(sine);
	return MBX_OK;
}

/* Point DE query: CalculateDistance at each point with the given detail size. */
int mbxo_distance(const mbx_params *p, const double *pts, int n, double detailSize, int normalCalc,
	double *dist, int32_t *iters)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	uint32_t seed = 1;
	ray_ctx rc = {&s, &seed, 0, 0};
	for (int k = 0; k < n; k++)
	{
		int it = 0;
		dist[k] = de(&rc, V3(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]), detailSize, normalCalc, &it);
		if (iters) iters[k] = it;
	}
	return MBX_OK;
}

/* Compute<calcModeNormal> at each point: distance, iters, maxiter flag and final z. */
int mbxo_compute(const mbx_params *p, const double *pts, int n, double *dist, int32_t *iters,
	int32_t *maxiter, double *zout)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	for (int k = 0; k < n; k++)
	{
		fout_t fo;
		compute(&s, MODE_NORMAL, V3(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]), -1, 1, &fo);
		if (dist) dist[k] = fo.distance;
		if (iters) iters[k] = fo.iters;
		if (maxiter) maxiter[k] = fo.maxiter;
		if (zout) { zout[3 * k] = fo.z.x; zout[3 * k + 1] = fo.z.y; zout[3 * k + 2] = fo.z.z; }
	}
	return MBX_OK;
}

/* One FormulaCode call: z4 in/out (4 doubles per item); aux in/out as (r, DE, color). */
int mbxo_formula_step(const mbx_params *p, int n, double *z4, double *aux3)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	for (int k = 0; k < n; k++)
	{
		v4 z = V4(z4[4 * k], z4[4 * k + 1], z4[4 * k + 2], z4[4 * k + 3]);
		aux_t aux;
		memset(&aux, 0, sizeof(aux));
		aux.r = aux3[3 * k];
		aux.DE = aux3[3 * k + 1];
		aux.color = aux3[3 * k + 2];
		formula_code(&s.slot[0], &z, &aux);
		z4[4 * k] = z.x;
		z4[4 * k + 1] = z.y;
		z4[4 * k + 2] = z.z;
		z4[4 * k + 3] = z.w;
		aux3[3 * k] = aux.r;
		aux3[3 * k + 1] = aux.DE;
		aux3[3 * k + 2] = aux.color;
	}
	return MBX_OK;
}

/* One FormulaCode step of container `slot` (0 = p->fractal) with the iteration index and the
 * constant: z4 and constc4 (xyzw) and aux3 (r, DE, color) are updated in place. */
int mbxo_formula_step_ex(const mbx_params *p, int slot, int n, double *z4, double *aux3,
	const int32_t *iter, double *constc4)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	if (slot < 0 || slot >= s.nSlots) return MBX_ERR_INVALID;
	for (int k = 0; k < n; k++)
	{
		v4 z = V4(z4[4 * k], z4[4 * k + 1], z4[4 * k + 2], z4[4 * k + 3]);
		aux_t aux;
		memset(&aux, 0, sizeof(aux));
		aux.r = aux3[3 * k];
		aux.DE = aux3[3 * k + 1];
		aux.color = aux3[3 * k + 2];
		aux.i = iter ? iter[k] : 0;
		if (constc4) aux.const_c = V4(constc4[4 * k], constc4[4 * k + 1], constc4[4 * k + 2], constc4[4 * k + 3]);
		formula_code(&s.slot[slot], &z, &aux);
		z4[4 * k] = z.x;
		z4[4 * k + 1] = z.y;
		z4[4 * k + 2] = z.z;
		z4[4 * k + 3] = z.w;
		aux3[3 * k] = aux.r;
		aux3[3 * k + 1] = aux.DE;
		aux3[3 * k + 2] = aux.color;
		if (constc4)
		{
			constc4[4 * k] = aux.const_c.x;
			constc4[4 * k + 1] = aux.const_c.y;
			constc4[4 * k + 2] = aux.const_c.z;
			constc4[4 * k + 3] = aux.const_c.w;
		}
	}
	return MBX_OK;
}

/* The global foldings alone (BoxFolding then SphericalFolding as the params enable them) on z4
 * and aux3 (r, DE, color), in place. */
int mbxo_global_foldings(const mbx_params *p, int n, double *z4, double *aux3)
{
	for (int k = 0; k < n; k++)
	{
		v4 z = V4(z4[4 * k], z4[4 * k + 1], z4[4 * k + 2], z4[4 * k + 3]);
		aux_t aux;
		memset(&aux, 0, sizeof(aux));
		aux.r = aux3[3 * k];
		aux.DE = aux3[3 * k + 1];
		aux.color = aux3[3 * k + 2];
		global_foldings(p, &z, &aux);
		z4[4 * k] = z.x;
		z4[4 * k + 1] = z.y;
		z4[4 * k + 2] = z.z;
		z4[4 * k + 3] = z.w;
		aux3[3 * k] = aux.r;
		aux3[3 * k + 1] = aux.DE;
		aux3[3 * k + 2] = aux.color;
	}
	return MBX_OK;
}

/* Camera/light derivation for host-code parity: mRot (9), fov, lightDirection (3). */
int mbxo_frame_constants(const mbx_params *p, double *mrot9, double *fov, double *light3)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	rotm m = s.mRot;
	if (m.zero) rot_null(&m);
	memcpy(mrot9, m.m, 9 * sizeof(double));
	*fov = s.fov;
	light3[0] = s.lightDirection.x;
	light3[1] = s.lightDirection.y;
	light3[2] = s.lightDirection.z;
	return MBX_OK;
}

/* Primary ray of one pixel: view direction (normalised) and the march result. */
int mbxo_march_pixel(const mbx_params *p, int x, int y, double *dir3, double *depth, double *thr,
	int32_t *found, int32_t *iters, double *point3)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	const int W = p->imageWidth, H = p->imageHeight;
	double ix = (1.0 / W) * x + (-0.5);
	double iy = (-1.0 / H) * y + 0.5;
	ix *= s.aspectRatio;
	v3 d = v3_normalize(view_vector(&s, ix, iy));
	uint32_t seed = mbxo_pixel_seed((uint32_t)y * (uint32_t)W + (uint32_t)x, p->randomSeed);
	ray_ctx rc = {&s, &seed, 0, 0};
	march_t m = {0};
	ray_marching(&rc, s.camera, d, 0.0, p->viewDistanceMax, 1, &m);
	dir3[0] = d.x;
	dir3[1] = d.y;
	dir3[2] = d.z;
	*depth = m.depth;
	*thr = m.distThresh;
	*found = m.found;
	*iters = m.iters;
	point3[0] = m.point.x;
	point3[1] = m.point.y;
	point3[2] = m.point.z;
	return MBX_OK;
}

/* Normal at a point (CalculateNormals) with the given distance threshold. */
int mbxo_normal(const mbx_params *p, const double *pt3, double distThresh, double *n3)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	uint32_t seed = 1;
	ray_ctx rc = {&s, &seed, 0, 0};
	shin_t in;
	memset(&in, 0, sizeof(in));
	in.point = V3(pt3[0], pt3[1], pt3[2]);
	in.distThresh = distThresh;
	in.delta = calc_delta(&s, in.point);
	v3 n = calculate_normals(&rc, &in);
	n3[0] = n.x;
	n3[1] = n.y;
	n3[2] = n.z;
	return MBX_OK;
}

/* Secondary-ray shaders at one surface point, for tests that pin them: `which` 0 = AuxShadow of
 * light #1 over its distance (delta / resolution for the penetrating light, view_distance_max
 * otherwise; out3[0]), 1 = AmbientOcclusion with the scene's AO vectors, 2 = FastAmbientOcclusion,
 * 3 = GlobalIlumination with the surface colour col3 (RGB).  The shader input is the point, the
 * normal, the distance threshold and delta = CalcDelta at the point; the random stream starts at
 * `seed`. */
int mbxo_shader_probe(const mbx_params *p, int which, const double *pt3, const double *n3, double distThresh,
	uint32_t seed, const float *col3, float *out3)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	uint32_t sd = seed;
	ray_ctx rc = {&s, &sd, 0, 0};
	shin_t in;
	memset(&in, 0, sizeof(in));
	in.point = V3(pt3[0], pt3[1], pt3[2]);
	in.normal = V3(n3[0], n3[1], n3[2]);
	in.distThresh = distThresh;
	in.delta = calc_delta(&s, in.point);
	out3[0] = out3[1] = out3[2] = 0.0f;
	switch (which)
	{
		case 0:
		{
			const double distance = p->light.penetrating ? in.delta / s.resolution : p->viewDistanceMax;
			out3[0] = aux_shadow(&rc, &in, distance, s.lightDirection);
			return MBX_OK;
		}
		case 1:
		case 2:
		{
			const rgba a = which == 1 ? ambient_occlusion(&rc, &in) : fast_ambient_occlusion(&rc, &in);
			out3[0] = a.R;
			out3[1] = a.G;
			out3[2] = a.B;
			return MBX_OK;
		}
		case 3:
		{
			const rgba c = RGBA(col3[0], col3[1], col3[2], 1.0f);
			global_illumination(&rc, &in, c, out3);
			return MBX_OK;
		}
		default: return MBX_ERR_INVALID;
	}
}

/* Colour index (Compute<calcModeColouring> + CalculateColorIndex) of fractal-local points. */
int mbxo_colour_index(const mbx_params *p, const double *pts, int n, double *out)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	for (int k = 0; k < n; k++) out[k] = compute_colour_index(&s, V3(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]));
	return MBX_OK;
}

/* Surface gradient colour at palette positions (GetColorFloat), 3 floats per position. */
int mbxo_gradient(const mbx_params *p, const float *positions, int n, float *rgb)
{
	if (p->surfaceGradientCount < 2 || p->surfaceGradientCount > MBX_MAX_GRADIENT_STOPS) return MBX_ERR_INVALID;
	for (int k = 0; k < n; k++)
	{
		rgba c = gradient_colour(p, positions[k]);
		rgb[3 * k] = c.R;
		rgb[3 * k + 1] = c.G;
		rgb[3 * k + 2] = c.B;
	}
	return MBX_OK;
}

/* SurfaceColour at world points, 3 floats per point. */
int mbxo_surface_colour(const mbx_params *p, const double *pts, int n, float *rgb)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	if (p->useColorsFromPalette && p->surfaceGradientEnable &&
		(p->surfaceGradientCount < 2 || p->surfaceGradientCount > MBX_MAX_GRADIENT_STOPS))
		return MBX_ERR_INVALID;
	for (int k = 0; k < n; k++)
	{
		rgba c = surface_colour(&s, V3(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]));
		rgb[3 * k] = c.R;
		rgb[3 * k + 1] = c.G;
		rgb[3 * k + 2] = c.B;
	}
	return MBX_OK;
}

/* Struct sizes for binding-layout checks. */
size_t mbxo_sizeof_params(void) { return sizeof(mbx_params); }
size_t mbxo_offsetof_fractal(void) { return offsetof(mbx_params, fractal); }

/* Background colour seen along a view direction (BackgroundShader). */
int mbxo_background(const mbx_params *p, const double *dir3, float *rgb3)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	rgba c = background_shader(&s, V3(dir3[0], dir3[1], dir3[2]));
	rgb3[0] = c.R;
	rgb3[1] = c.G;
	rgb3[2] = c.B;
	return MBX_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Differentiable render, FP64 (the function the adjoint differentiates).  Not a reference     */
/* feature: the march above finds the hit, Newton iterations on DE(camera + t dir) = distThresh */
/* refine it, the normal is the analytic DE gradient (3-tangent dual numbers through the        */
/* Mandelbulb FormulaCode, fractal_mandelbulb.cpp:29-43, and the analytic logarithmic DE,      */
/* compute_fractal.cpp:561-568), and Lambert / background shading run in double.  The tests take */
/* central differences of this render's loss as the reference gradient.                         */
/* ------------------------------------------------------------------------------------------ */
typedef struct { double v, d[3]; } dl3;

static inline dl3 dl_c(double v) { dl3 r = {v, {0.0, 0.0, 0.0}}; return r; }
static inline dl3 dl_add(dl3 a, dl3 b)
{
	dl3 r = {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
	return r;
}
static inline dl3 dl_addc(dl3 a, double c) { a.v += c; return a; }
static inline dl3 dl_scale(dl3 a, double s)
{
	dl3 r = {a.v * s, {a.d[0] * s, a.d[1] * s, a.d[2] * s}};
	return r;
}
static inline dl3 dl_mul(dl3 a, dl3 b)
{
	dl3 r;
	r.v = a.v * b.v;
	for (int k = 0; k < 3; k++) r.d[k] = a.v * b.d[k] + a.d[k] * b.v;
	return r;
}
static inline dl3 dl_div(dl3 a, dl3 b)
{
	dl3 r;
	r.v = a.v / b.v;
	for (int k = 0; k < 3; k++) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
	return r;
}
static inline dl3 dl_fn(dl3 a, double f, double f1)
{
	dl3 r = {f, {f1 * a.d[0], f1 * a.d[1], f1 * a.d[2]}};
	return r;
}
static inline dl3 dl_sqrt(dl3 a) { const double s = sqrt(a.v); return dl_fn(a, s, 0.5 / s); }
static inline dl3 dl_log(dl3 a) { return dl_fn(a, log(a.v), 1.0 / a.v); }
static inline dl3 dl_sin(dl3 a) { return dl_fn(a, sin(a.v), cos(a.v)); }
static inline dl3 dl_cos(dl3 a) { return dl_fn(a, cos(a.v), -sin(a.v)); }
static inline dl3 dl_asin(dl3 a) { return dl_fn(a, asin(a.v), 1.0 / sqrt(1.0 - a.v * a.v)); }
static inline dl3 dl_powc(dl3 a, double p) { return dl_fn(a, pow(a.v, p), p * pow(a.v, p - 1.0)); }
static inline dl3 dl_atan2(dl3 y, dl3 x)
{
	const double q = x.v * x.v + y.v * y.v;
	dl3 r;
	r.v = atan2(y.v, x.v);
	for (int k = 0; k < 3; k++) r.d[k] = (x.v * y.d[k] - y.v * x.d[k]) / q;
	return r;
}

/* analytic log DE of the Mandelbulb node at world point pw (tangents carried in pw) */
static dl3 diff_de_o(const scene_t *s, const dl3 pw[3], int *iters, int *ok)
{
	const double *R = s->nodeRot.m;
	dl3 p[3];
	for (int i = 0; i < 3; i++)
	{
		if (s->nodeRotated)
			p[i] = dl_add(dl_add(dl_scale(pw[0], R[3 * i]), dl_scale(pw[1], R[3 * i + 1])), dl_scale(pw[2], R[3 * i + 2]));
		else
			p[i] = pw[i];
	}
	dl3 z[3] = {p[0], p[1], p[2]};
	dl3 DE = dl_c(1.0);
	dl3 r = dl_sqrt(dl_add(dl_add(dl_mul(z[0], z[0]), dl_mul(z[1], z[1])), dl_mul(z[2], z[2])));
	const double P = s->slot[0].bulbPower;
	int i;
	*ok = 1;
	for (i = 0; i < s->formulaMaxiter; i++)
	{
		if (!(r.v > 0.0)) { *ok = 0; break; }
		const dl3 th0 = dl_addc(dl_asin(dl_div(z[2], r)), s->slot[0].bulbBeta);
		const dl3 ph0 = dl_addc(dl_atan2(z[1], z[0]), s->slot[0].bulbAlpha);
		dl3 rp = dl_powc(r, P - 1.0);
		const dl3 th = dl_scale(th0, P);
		const dl3 ph = dl_scale(ph0, P);
		const dl3 cth = dl_cos(th);
		DE = dl_addc(dl_scale(dl_mul(rp, DE), P), 1.0);
		rp = dl_mul(rp, r);
		z[0] = dl_add(dl_mul(dl_mul(cth, dl_cos(ph)), rp), dl_scale(p[0], s->constantMultiplier.x));
		z[1] = dl_add(dl_mul(dl_mul(cth, dl_sin(ph)), rp), dl_scale(p[1], s->constantMultiplier.y));
		z[2] = dl_add(dl_mul(dl_sin(th), rp), dl_scale(p[2], s->constantMultiplier.z));
		r = dl_sqrt(dl_add(dl_add(dl_mul(z[0], z[0]), dl_mul(z[1], z[1])), dl_mul(z[2], z[2])));
		if (!isfinite(r.v)) { *ok = 0; break; }
		if (r.v > s->slot[0].bailout) break;
	}
	*iters = i + 1;
	if (!(DE.v > 0.0) || !(r.v > 1.0)) { *ok = 0; return dl_c(0.0); }
	return dl_scale(dl_div(dl_mul(r, dl_log(r)), DE), 0.5);
}

static double diff_thresh_o(const scene_t *s, double t)
{
	return s->p->constantDEThreshold ? s->p->DEThresh : t * s->resolution * s->fov / s->p->detailLevel;
}

/* DE value and gradient at a world point */
static double diff_grad_o(const scene_t *s, v3 q, double g[3], int *iters, int *ok)
{
	dl3 pw[3] = {{q.x, {1, 0, 0}}, {q.y, {0, 1, 0}}, {q.z, {0, 0, 1}}};
	const dl3 de = diff_de_o(s, pw, iters, ok);
	for (int k = 0; k < 3; k++) g[k] = de.d[k];
	return de.v;
}

/* branch (optional): which side of each kink of the shading the pixel is on -- bit 0 n.l > 0,
 * bit 1 shade clamped at 500, bit 2 (misses) below the horizon of the background gradient,
 * bits 3-5 channel clamped at 0; a finite difference across a kink is no derivative */
int mbxo_render_diff(const mbx_params *p, int y0, int y1, int threads, double *rgb, double *depth,
	int32_t *iters, uint8_t *mask, double *normal, int32_t *refined, int32_t *branch)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	if (p->fractal.formula != MBX_FORMULA_MANDELBULB || s.DEType != MBX_DE_ANALYTIC ||
		s.DEAnalyticFunction != MBX_ANALYTIC_LOGARITHMIC || !s.slot[0].addC || s.juliaEnabled || s.isHybrid)
		return MBX_ERR_UNSUPPORTED;
	const int W = p->imageWidth, H = p->imageHeight;
	const double dthr = p->constantDEThreshold ? 0.0 : s.resolution * s.fov / p->detailLevel;
	if (y0 < 0) y0 = 0;
	if (y1 > H) y1 = H;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
	for (int ys = y0; ys < y1; ys++)
	{
		for (int xs = 0; xs < W; xs++)
		{
			const size_t i = (size_t)ys * W + xs;
			uint32_t seed = mbxo_pixel_seed((uint32_t)i, p->randomSeed);
			ray_ctx rc = {&s, &seed, 0, 0};
			const double ix = ((1.0 / W) * xs + (-0.5)) * s.aspectRatio;
			const double iy = (-1.0 / H) * ys + 0.5;
			const v3 d = v3_normalize(view_vector(&s, ix, iy));
			march_t m = {0};
			ray_marching(&rc, s.camera, d, 0.0, p->viewDistanceMax, 1, &m);
			double col[3], n[3] = {0, 0, 0}, t = m.depth;
			int it = m.iters, ref = 0, br = 0;
			if (m.found)
			{
				/* Newton on G(t) = DE(camera + t d) - distThresh(t) */
				for (int k = 0; k < 24; k++)
				{
					double g[3];
					int ok, nit;
					const v3 q = v3_add(s.camera, v3_mul(d, t));
					const double de = diff_grad_o(&s, q, g, &nit, &ok);
					if (!ok) break;
					const double thr = diff_thresh_o(&s, t);
					const double Gt = g[0] * d.x + g[1] * d.y + g[2] * d.z - dthr;
					if (!(Gt < 0.0)) break;
					const double dt = -(de - thr) / Gt;
					if (!(fabs(dt) <= 4.0 * thr)) break;
					t += dt;
					it = nit;
					if (fabs(dt) <= 1e-15 * t) { ref = 1; break; }
				}
				if (!ref) t = m.depth;
				double g[3];
				int ok, nit;
				diff_grad_o(&s, v3_add(s.camera, v3_mul(d, t)), g, &nit, &ok);
				const double gl = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
				if (ok && gl > 0.0)
					for (int k = 0; k < 3; k++) n[k] = g[k] / gl;
				else
				{
					n[0] = 1.0;
					ref = 0;
				}
				/* Lambert, as ObjectShader with shadows and AO off (shader_object.cpp:40-229) */
				const double ndotl = n[0] * s.lightDirection.x + n[1] * s.lightDirection.y + n[2] * s.lightDirection.z;
				double shade = 1.0 - p->shading + (ndotl > 0.0 ? ndotl : 0.0) * p->shading;
				shade *= p->light.intensity;
				br |= ndotl > 0.0 ? 1 : 0;
				if (shade > 500.0)
				{
					shade = 500.0;
					br |= 2;
				}
				for (int k = 0; k < 3; k++)
				{
					const double light = p->light.enabled ? p->light.color[k] * shade : 0.0;
					const double v = p->fillLightColor[k] * p->surfaceColor[k] + light * p->surfaceColor[k] +
						p->luminosity * p->luminosityColor[k];
					col[k] = v > 0.0 ? v : 0.0;
					br |= v > 0.0 ? 0 : 8 << k;
				}
			}
			else
			{
				/* BackgroundShader in double: 3-colour gradient and the light's disk */
				const double grad = d.z + 1.0;
				br |= grad < 1.0 ? 4 : 0;
				for (int k = 0; k < 3; k++)
				{
					double c;
					if (p->background3ColorsEnable)
						c = grad < 1.0 ? p->backgroundColor3[k] * (1.0 - grad) + p->backgroundColor2[k] * grad
									   : p->backgroundColor2[k] * (2.0 - grad) + p->backgroundColor1[k] * (grad - 1.0);
					else
						c = p->backgroundColor1[k];
					col[k] = c * p->backgroundBrightness;
				}
				if (p->light.enabled)
				{
					const double dl = d.x * s.lightDirection.x + d.y * s.lightDirection.y + d.z * s.lightDirection.z;
					const double a = (1.0 - dl) * (360.0 / p->light.size);
					const double inten = 1.0 / (pow(a, 6.0 * p->light.contourSharpness) + 1.0) * (p->light.visibility * p->light.intensity);
					for (int k = 0; k < 3; k++) col[k] += p->light.color[k] * inten;
				}
			}
			if (rgb) for (int k = 0; k < 3; k++) rgb[3 * i + k] = col[k];
			if (depth) depth[i] = m.found ? t : 1e20;
			if (iters) iters[i] = it;
			if (mask) mask[i] = (uint8_t)m.found;
			if (normal) for (int k = 0; k < 3; k++) normal[3 * i + k] = n[k];
			if (refined) refined[i] = ref;
			if (branch) branch[i] = br;
		}
	}
	return MBX_OK;
}

/* DE and its analytic gradient (the differentiable render's function) at n world points */
int mbxo_diff_de(const mbx_params *p, const double *pts, int n, double *de, double *grad, int32_t *iters,
	int32_t *ok)
{
	scene_t s;
	int rc0 = scene_init(&s, p);
	if (rc0) return rc0;
	for (int i = 0; i < n; i++)
	{
		int it, okk;
		double g[3];
		const double d = diff_grad_o(&s, V3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), g, &it, &okk);
		if (de) de[i] = d;
		if (grad) for (int k = 0; k < 3; k++) grad[3 * i + k] = g[k];
		if (iters) iters[i] = it;
		if (ok) ok[i] = okk;
	}
	return MBX_OK;
}

/* cImage::CalculateGammaTable + CalculatePixel + ConvertTo8bitChar (cimage.cpp:317-376, 378-397,
 * 460-470) over n pixels of a float RGB image, in float as the reference computes them:
 * brightness, contrast about 0.5, the clamp at 0, tanhf with hdr, saturation about
 * V = sqrtf(R R 0.299 + G G 0.587 + B B 0.114), the clamp to [0, 1], quint16(x * 65535) and the
 * table int(powf(i / 65536, 1 / gamma) * 65535); 8-bit = 16-bit / 256.  out16 / out8 may be NULL. */
int mbxo_compile_image(const mbx_image_adjustments *adj, const float *rgb, int n, uint16_t *out16,
	uint8_t *out8)
{
	if (!adj || !rgb || n < 0 || !(adj->gamma >= 0.0f)) return MBX_ERR_INVALID;
	static int table[65536];
	const float invGamma = 1.0f / adj->gamma;
	for (int i = 0; i < 65536; i++) table[i] = (int)(powf((float)i / 65536.0f, invGamma) * 65535.0f);
	for (int i = 0; i < n; i++)
	{
		float R = rgb[3 * i] * adj->brightness;
		float G = rgb[3 * i + 1] * adj->brightness;
		float B = rgb[3 * i + 2] * adj->brightness;
		R = (R - 0.5f) * adj->contrast + 0.5f;
		G = (G - 0.5f) * adj->contrast + 0.5f;
		B = (B - 0.5f) * adj->contrast + 0.5f;
		if (R < 0.0f) R = 0.0f;
		if (G < 0.0f) G = 0.0f;
		if (B < 0.0f) B = 0.0f;
		if (adj->hdr)
		{
			R = tanhf(R);
			G = tanhf(G);
			B = tanhf(B);
		}
		const float rFactor = 0.299f, gFactor = 0.587f, bFactor = 0.114f;
		const float V = sqrtf(R * R * rFactor + G * G * gFactor + B * B * bFactor);
		R = V + (R - V) * adj->saturation;
		G = V + (G - V) * adj->saturation;
		B = V + (B - V) * adj->saturation;
		R = R < 0.0f ? 0.0f : (R > 1.0f ? 1.0f : R);
		G = G < 0.0f ? 0.0f : (G > 1.0f ? 1.0f : G);
		B = B < 0.0f ? 0.0f : (B > 1.0f ? 1.0f : B);
		const uint16_t r16 = (uint16_t)table[(uint16_t)(R * 65535.0f)];
		const uint16_t g16 = (uint16_t)table[(uint16_t)(G * 65535.0f)];
		const uint16_t b16 = (uint16_t)table[(uint16_t)(B * 65535.0f)];
		if (out16)
		{
			out16[3 * i] = r16;
			out16[3 * i + 1] = g16;
			out16[3 * i + 2] = b16;
		}
		if (out8)
		{
			out8[3 * i] = (uint8_t)(r16 / 256);
			out8[3 * i + 1] = (uint8_t)(g16 / 256);
			out8[3 * i + 2] = (uint8_t)(b16 / 256);
		}
	}
	return MBX_OK;
}

/* The z-buffer image of ImageFileSavePNG::SavePNG (file_image.cpp:628-742) for n depths: the
 * range (smallest depth, largest depth below 1e19; or the constant one), the logarithmic or
 * linear scale in float, int(z1 * 254 or 65534), misses 255 / 65535, optional inversion.
 * `out` holds uint8_t (bits 8) or uint16_t (bits 16) samples. */
int mbxo_compile_zbuffer(const float *depth, int n, int bits, int logarithmic, int invert, int constantRange,
	double minDepth, double maxDepth, void *out)
{
	if (!depth || !out || n < 0 || (bits != 8 && bits != 16)) return MBX_ERR_INVALID;
	float minZ = (float)1.0e50;
	float maxZ = 0.0f;
	if (constantRange)
	{
		minZ = (float)minDepth;
		maxZ = (float)maxDepth;
	}
	else
		for (int i = 0; i < n; i++)
		{
			const float z = depth[i];
			if (z > maxZ && z < 1e19f) maxZ = z;
			if (z < minZ) minZ = z;
		}
	const float kZ = logf(maxZ / minZ);
	for (int i = 0; i < n; i++)
	{
		const float z = depth[i];
		float z1 = logarithmic ? logf(z / minZ) / kZ : (z - minZ) / (maxZ - minZ);
		if (z1 < 0) z1 = 0.0f;
		if (z1 > 1.0f) z1 = 1.0f;
		if (bits == 16)
		{
			int intZ = (int)(z1 * 65534);
			if (z > 1e19f) intZ = 65535;
			if (invert) intZ = 65535 - intZ;
			((uint16_t *)out)[i] = (uint16_t)intZ;
		}
		else
		{
			int intZ = (int)(z1 * 254);
			if (z > 1e19) intZ = 255;
			if (invert) intZ = 255 - intZ;
			((uint8_t *)out)[i] = (uint8_t)intZ;
		}
	}
	return MBX_OK;
}

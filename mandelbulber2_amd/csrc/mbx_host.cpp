// mbx_host.cpp - C ABI of the raymarcher (include/mbx.h): context, scene upload, launches.
//
// The host side of what cRenderJob / cOpenClEngineRenderFractal do before a launch
// (render_job.cpp:399-477, opencl_engine_render_fractal.cpp:1114-1245): turn the user-level
// parameters into per-frame constants in double precision (camera frame, light vector, formula
// constants), then hand them to the device once per frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "../../include/mbx.h"
#include "mbx_frame.h"
#include "mbx_diff.h"
#include "mbx_launch.h"

using mbx::FrameConst;

struct mbx_ctx
{
	int device = 0;
	hipStream_t stream = nullptr;
	FrameConst frame{};
	FrameConst *d_frame = nullptr;
	unsigned long long *d_counters = nullptr; // two banks of kCounterBytes
	int bank = 0;           // counter bank of the next launch
	bool bankClean = false; // that bank is known to be zero
	int statsBank = 0;      // bank holding the last launch's counters
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	bool sceneReady = false;
	bool launched = false;
	double mRot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
	double fov = 0.0;
	double lightDir[3] = {0, 0, 0};
	uint64_t rays = 0;
	std::string err;
	// persistent / batched launches
	int kernelMode = 0; // 0 grid batch kernel, 1 plain kernel per frame, 2 persistent tile queues
	int waves = 0;      // 0 = default
	int useGraph = 0;  // replay a captured HIP graph (MBX_TUNE_GRAPH); direct launches by default
	int deltaTangent = 0;
	int numCUs = 256;
	FrameConst *d_frames = nullptr;
	mbx::FrameJob *d_jobs = nullptr;
	int capacity = 0;
	unsigned int *d_queue = nullptr;
	std::vector<mbx::FrameJob> jobsHost; // what d_jobs holds
	std::vector<FrameConst> framesHost;  // what d_frames holds
	hipGraphExec_t graphExec = nullptr;
	int graphKind = -1, graphFrames = -1, graphWaves = -1, graphTile = -1, graphFull = -1, graphMode = -1;
	unsigned char *graphHits = nullptr;
	unsigned int *graphHitCounts = nullptr;
	unsigned int graphTiles = 0;
	int tileShift = 3; // 8x8-pixel tiles
	int farField = 1;  // far-field march steps without DE evaluation (MBX_TUNE_FAR_FIELD)
	int precision = 0; // Mandelbulb march precision (MBX_TUNE_PRECISION): 0 auto, 1 FP32, 2 FP64
	int timing = 1;     // time renders with events (MBX_TUNE_TIMING)
	bool timedLast = false; // the last launch was bracketed by ev0/ev1
	int batchFull = 1; // kernel variant the staged frames need (mbx::kVariantLite .. kVariantLens)
	// split render (MBX_TUNE_SPLIT): hit records and per-frame hit counts, grown on demand
	int split = 1;
	bool batchSplit = false; // the staged batch takes the split render
	size_t batchPixels = 0;
	unsigned char *d_hits = nullptr;
	size_t hitsBytes = 0;
	unsigned int *d_hitCounts = nullptr;
	int hitCountsCap = 0;
	// differentiable render
	mbx::DiffConst diff{};
	std::string diffWhy = "no scene uploaded"; // empty when the scene is differentiable
	mbx::DiffConst *d_diff = nullptr;
	double *d_grad = nullptr;
	bool diffUploaded = false;
	// image post-processing: the gamma table of the last gamma (cImage::CalculateGammaTable)
	uint16_t *d_gamma = nullptr;
	float gammaOf = 0.0f;
	// sin / cos of the lens angles and the soft-shadow cube roots, computed once on the host
	double *d_mcTables = nullptr;
	bool gammaReady = false;
	std::vector<uint16_t> gammaHost;
	uint32_t *d_zrange = nullptr; // z-buffer image: the frame's depth range
	// allocation-free batches (mbx_render_batch_noalloc): the device's static scratch, host staging
	// sized at creation, and an event marking the end of the last copy out of the staging
	unsigned char *d_scratch = nullptr;
	std::vector<FrameConst> naFrames;
	std::vector<mbx::FrameJob> naJobs;
	hipEvent_t naCopied = nullptr;
	bool naCopyPending = false;
	unsigned long long *statsCounters = nullptr; // counters of the last launch when not a bank
};

namespace
{

constexpr size_t kMcTableDoubles = 2 * (size_t)mbx::kLensAngles + (size_t)mbx::kSoftRadii;

// M_PI_180 of mandelbulber2/src/fractal.h:46, the factor RecalculateFractalParams uses
constexpr double kPi180 = 0.01745329251994329576923690768489;

int fail(mbx_ctx *ctx, int code, const std::string &msg)
{
	if (ctx) ctx->err = msg;
	return code;
}

int hip_fail(mbx_ctx *ctx, hipError_t e, const char *where)
{
	return fail(ctx, MBX_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

// Event around a render launch, when kernel timing is on (each marker costs the stream ~10 us).
hipError_t mark(mbx_ctx *ctx, hipEvent_t ev)
{
	return ctx->timing ? hipEventRecord(ev, ctx->stream) : hipSuccess;
}

unsigned long long *counter_bank(mbx_ctx *ctx, int b)
{
	return ctx->d_counters + (size_t)b * (mbx::kCounterBytes / sizeof(unsigned long long));
}

// Counters of a launch that clears its own bank first (clear kernel): the bank is dirty afterwards.
unsigned long long *counters_cleared_by_launch(mbx_ctx *ctx)
{
	ctx->statsBank = ctx->bank;
	ctx->bankClean = false;
	return counter_bank(ctx, ctx->bank);
}

// ---- double-precision vector helpers for the host-side derivation -------------------------
struct D3
{
	double x, y, z;
};
D3 d3(const double *a) { return D3{a[0], a[1], a[2]}; }
D3 add(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 mul(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double len(D3 a) { return std::sqrt(dot(a, a)); }
D3 cross(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
D3 unit(D3 a)
{
	const double k = 1.0 / len(a);
	return mul(a, k);
}
// Rodrigues rotation of v about a unit axis (CVector3::RotateAroundVectorByAngle, algebra.cpp:71-77)
// (same operation order as the reference: ((v cos) + (axis x v) sin) + (axis (axis.v)) (1 - cos))
D3 rotate_about(D3 v, D3 axis, double angle)
{
	D3 r = mul(v, std::cos(angle));
	r = add(r, mul(cross(axis, v), std::sin(angle)));
	r = add(r, mul(mul(axis, dot(axis, v)), 1 - std::cos(angle)));
	return r;
}

// 3x3 rotation composed by right-multiplication as CRotationMatrix::RotateX/Y/Z do
// (algebra.cpp:198-247)
struct Rot
{
	double m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
	void mulRight(const double b[9])
	{
		double c[9];
		for (int r = 0; r < 3; r++)
			for (int k = 0; k < 3; k++)
				c[3 * r + k] = m[3 * r] * b[k] + m[3 * r + 1] * b[3 + k] + m[3 * r + 2] * b[6 + k];
		std::memcpy(m, c, sizeof(c));
	}
	void x(double a)
	{
		if (a == 0.0) return;
		const double s = std::sin(a), c = std::cos(a);
		const double b[9] = {1, 0, 0, 0, c, -s, 0, s, c};
		mulRight(b);
	}
	void y(double a)
	{
		if (a == 0.0) return;
		const double s = std::sin(a), c = std::cos(a);
		const double b[9] = {c, 0, s, 0, 1, 0, -s, 0, c};
		mulRight(b);
	}
	void z(double a)
	{
		if (a == 0.0) return;
		const double s = std::sin(a), c = std::cos(a);
		const double b[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
		mulRight(b);
	}
	D3 apply(D3 v) const
	{
		return D3{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
			m[6] * v.x + m[7] * v.y + m[8] * v.z};
	}
};

double wrap_angle(double a) { return std::fmod(a + 3.0 * M_PI, 2.0 * M_PI) - M_PI; }

// ---- forward-mode dual numbers for the camera and node-rotation Jacobians -------------------
struct Dual1
{
	double v, d;
};
Dual1 operator+(Dual1 a, Dual1 b) { return Dual1{a.v + b.v, a.d + b.d}; }
Dual1 operator-(Dual1 a, Dual1 b) { return Dual1{a.v - b.v, a.d - b.d}; }
Dual1 operator-(Dual1 a) { return Dual1{-a.v, -a.d}; }
Dual1 operator*(Dual1 a, Dual1 b) { return Dual1{a.v * b.v, a.v * b.d + a.d * b.v}; }
Dual1 operator/(Dual1 a, Dual1 b) { return Dual1{a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
Dual1 dsin(Dual1 a) { return Dual1{std::sin(a.v), std::cos(a.v) * a.d}; }
Dual1 dcos(Dual1 a) { return Dual1{std::cos(a.v), -std::sin(a.v) * a.d}; }
Dual1 dsqrt(Dual1 a)
{
	const double s = std::sqrt(a.v);
	return Dual1{s, 0.5 * a.d / s};
}
Dual1 datan2(Dual1 y, Dual1 x)
{
	return Dual1{std::atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / (x.v * x.v + y.v * y.v)};
}
Dual1 dwrap(Dual1 a) { return Dual1{wrap_angle(a.v), a.d}; }
Dual1 dc(double v) { return Dual1{v, 0.0}; }

struct DD3
{
	Dual1 x, y, z;
};
DD3 dd_sub(DD3 a, DD3 b) { return DD3{a.x - b.x, a.y - b.y, a.z - b.z}; }
Dual1 dd_dot(DD3 a, DD3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DD3 dd_mul(DD3 a, Dual1 s) { return DD3{a.x * s, a.y * s, a.z * s}; }
DD3 dd_add(DD3 a, DD3 b) { return DD3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DD3 dd_unit(DD3 a) { return dd_mul(a, dc(1.0) / dsqrt(dd_dot(a, a))); }
DD3 dd_cross(DD3 a, DD3 b)
{
	return DD3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
DD3 dd_rotate_about(DD3 v, DD3 axis, Dual1 angle)
{
	DD3 r = dd_mul(v, dcos(angle));
	r = dd_add(r, dd_mul(dd_cross(axis, v), dsin(angle)));
	r = dd_add(r, dd_mul(dd_mul(axis, dd_dot(axis, v)), dc(1.0) - dcos(angle)));
	return r;
}
// Rot with dual entries; an axis rotation is skipped only when its angle and tangent are 0
struct DRot
{
	Dual1 m[9] = {dc(1), dc(0), dc(0), dc(0), dc(1), dc(0), dc(0), dc(0), dc(1)};
	void mulRight(const Dual1 b[9])
	{
		Dual1 c[9];
		for (int r = 0; r < 3; r++)
			for (int k = 0; k < 3; k++) c[3 * r + k] = m[3 * r] * b[k] + m[3 * r + 1] * b[3 + k] + m[3 * r + 2] * b[6 + k];
		for (int i = 0; i < 9; i++) m[i] = c[i];
	}
	void axis(int ax, Dual1 a)
	{
		if (a.v == 0.0 && a.d == 0.0) return;
		const Dual1 s = dsin(a), c = dcos(a), o = dc(1.0), z = dc(0.0);
		if (ax == 0)
		{
			const Dual1 b[9] = {o, z, z, z, c, -s, z, s, c};
			mulRight(b);
		}
		else if (ax == 1)
		{
			const Dual1 b[9] = {c, z, s, z, o, z, -s, z, c};
			mulRight(b);
		}
		else
		{
			const Dual1 b[9] = {c, -s, z, s, c, z, z, z, o};
			mulRight(b);
		}
	}
};

// The camera rotation of derive_frame (camera_frame + PrepareMainVectors) in dual numbers: the
// tangents of the result are its derivatives along the tangents of camera and target.
void camera_rotation_dual(const Dual1 cam[3], const Dual1 tgt[3], const double top[3], double sweetH,
	double sweetV, Dual1 out[9])
{
	const DD3 c{cam[0], cam[1], cam[2]}, t{tgt[0], tgt[1], tgt[2]};
	const DD3 f = dd_unit(dd_sub(t, c));
	Dual1 yaw = datan2(f.y, f.x) - dc(0.5 * M_PI);
	Dual1 pitch = datan2(f.z, dsqrt(f.x * f.x + f.y * f.y));
	DD3 tv{dc(top[0]), dc(top[1]), dc(top[2])};
	tv = dd_unit(tv);
	tv = dd_rotate_about(tv, DD3{dc(0), dc(0), dc(1)}, -yaw);
	tv = dd_rotate_about(tv, DD3{dc(1), dc(0), dc(0)}, -pitch);
	Dual1 roll = dc(0.5 * M_PI) - datan2(tv.z, tv.x);
	yaw = dwrap(yaw);
	pitch = dwrap(pitch);
	roll = dwrap(roll);
	DRot m;
	m.axis(2, yaw);
	m.axis(0, pitch);
	m.axis(1, roll);
	m.axis(2, dc(-(sweetH / 180.0 * M_PI)));
	m.axis(0, dc(sweetV / 180.0 * M_PI));
	for (int i = 0; i < 9; i++) out[i] = m.m[i];
}

// Camera frame from (camera, target, top): yaw/pitch/roll as cCameraTarget::SetCameraTargetTop
// (camera_target.cpp:86-126) and the forward/top/right vectors.
struct CameraFrame
{
	D3 forward, top, right;
	double yaw, pitch, roll;
};

CameraFrame camera_frame(D3 camera, D3 target, D3 top)
{
	CameraFrame cf;
	const bool same = camera.x == target.x && camera.y == target.y && camera.z == target.z;
	const D3 f = unit(same ? D3{0.0, 1.0, 0.0} : sub(target, camera));
	double yaw = std::atan2(f.y, f.x) - 0.5 * M_PI;
	double pitch = std::atan2(f.z, std::sqrt(f.x * f.x + f.y * f.y));
	D3 t = unit(top);
	t = rotate_about(t, D3{0, 0, 1}, -yaw);
	t = rotate_about(t, D3{1, 0, 0}, -pitch);
	double roll = 0.5 * M_PI - std::atan2(t.z, t.x);
	yaw = wrap_angle(yaw);
	pitch = wrap_angle(pitch);
	roll = wrap_angle(roll);
	D3 up{0, 0, 1};
	up = rotate_about(up, D3{0, 1, 0}, roll);
	up = rotate_about(up, D3{1, 0, 0}, pitch);
	up = rotate_about(up, D3{0, 0, 1}, yaw);
	cf.forward = f;
	cf.top = up;
	cf.right = cross(f, up);
	cf.yaw = yaw;
	cf.pitch = pitch;
	cf.roll = roll;
	return cf;
}

void set3f(float *dst, D3 v)
{
	dst[0] = (float)v.x;
	dst[1] = (float)v.y;
	dst[2] = (float)v.z;
}
void set3f(float *dst, const double *v)
{
	dst[0] = (float)v[0];
	dst[1] = (float)v[1];
	dst[2] = (float)v[2];
}
void setRot(float *dst, const Rot &r)
{
	for (int i = 0; i < 9; i++) dst[i] = (float)r.m[i];
}

// Far-field radius of the FP32 Mandelbulb march.  Beyond radius r the first formula iteration
// gives |z_1| >= r^P - c r (|T(p)| = r^P for every angle offset, c = max |constantMultiplier|) and
// DE' = P r^(P-1) + 1, so the analytic DE is at least
//   log:    f(r) = 0.5 (r^P - c r) ln(r^P - c r) / (P r^(P-1) + 1)
//   linear: f(r) = (r^P - c r) / (P r^(P-1) + 1)
// once r^P - c r exceeds the bailout (escape at i = 0).  The march's step
// (dist - 0.5 thr) DEFactor (1 - Random(1000)/10000) is then >= 3, i.e. exactly the clamp 3,
// when f(r) >= (3 + 0.5 thrMax) / (0.9 DEFactor); the clamp of dist at 10 must not cap below
// that.  Returns the smallest radius (squared, with 1% margins) from which the bound holds on a
// dense logarithmic grid up to r = 1e4 (farther than any march reaches), or +inf.
float far_field_r2_uncached(double P, bool logDE, double bailout, double c, double DEFactor, double thrMax);

// memoised: batches derive the same bound for frame after frame
float far_field_r2(double P, bool logDE, double bailout, double c, double DEFactor, double thrMax)
{
	thread_local std::map<std::array<double, 6>, float> memo;
	const std::array<double, 6> key{P, logDE ? 1.0 : 0.0, bailout, c, DEFactor, thrMax};
	auto it = memo.find(key);
	if (it != memo.end()) return it->second;
	const float r2 = far_field_r2_uncached(P, logDE, bailout, c, DEFactor, thrMax);
	if (memo.size() > 4096) memo.clear();
	memo.emplace(key, r2);
	return r2;
}

float far_field_r2_uncached(double P, bool logDE, double bailout, double c, double DEFactor, double thrMax)
{
	const float inf = std::numeric_limits<float>::infinity();
	if (!(P >= 2.0) || !(DEFactor > 0.0)) return inf;
	const double need = (3.0 + 0.5 * thrMax) / (0.9 * DEFactor) * 1.01;
	if (need >= 10.0) return inf;
	auto ok = [&](double r) {
		const double r1 = std::pow(r, P) - c * r;
		if (!(r1 > bailout * 1.01) || !(r1 > 1.0)) return false;
		const double de = P * std::pow(r, P - 1.0) + 1.0;
		const double f = logDE ? 0.5 * r1 * std::log(r1) / de : r1 / de;
		return f >= need;
	};
	// the largest grid radius at which the bound fails, scanning down from 1e4
	double rFail = 0.0;
	for (double r = 1e4; r >= 1.0; r /= 1.0005)
		if (!ok(r))
		{
			rFail = r;
			break;
		}
	if (rFail >= 1e3) return inf;
	const double r0 = rFail * 1.01;
	return (float)(r0 * r0);
}

// Mandelbulb DE at one point in FP64 (FormulaCode of fractal_mandelbulb.cpp:29-43 inside the
// Compute loop), used to choose the march precision: the DE at the camera bounds the distance to
// the nearest surface, and with it the smallest distance threshold of the view.
double host_bulb_de(const double p[3], const mbx_fractal &fr, double bailout, int analyticFunction,
	bool addc, bool julia)
{
	const double power = fr.bulb.power;
	const double alpha = fr.bulb.alphaAngleOffset * kPi180, beta = fr.bulb.betaAngleOffset * kPi180;
	double z[3] = {p[0], p[1], p[2]};
	double r = std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
	double DE = 1.0;
	for (int i = 0; i < fr.formulaMaxiter; i++)
	{
		const double th = (std::asin(z[2] / r) + beta) * power;
		const double ph = (std::atan2(z[1], z[0]) + alpha) * power;
		double rp = std::pow(r, power - 1.0);
		DE = rp * DE * power + 1.0;
		rp *= r;
		z[0] = std::cos(th) * std::cos(ph) * rp;
		z[1] = std::cos(th) * std::sin(ph) * rp;
		z[2] = std::sin(th) * rp;
		for (int k = 0; k < 3; k++)
			if (addc) z[k] += julia ? fr.juliaC[k] * fr.constantFactor[k] : p[k] * fr.constantFactor[k];
		r = std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
		if (!std::isfinite(r) || r > bailout) break;
	}
	if (!std::isfinite(r) || !(DE > 0.0)) return 0.0;
	if (analyticFunction == MBX_ANALYTIC_LOGARITHMIC) return r > 1.0 ? 0.5 * r * std::log(r) / DE : 0.0;
	return r / DE;
}

// Kaleidoscopic IFS DE at one point in FP64 (FormulaCode of fractal_kaleidoscopic_ifs.cpp:30-74
// inside the Compute loop, IFS finish (r - 2) / DE, compute_fractal.cpp:606-610), for the same
// precision choice as host_bulb_de.
double host_ifs_de(const double p[3], const mbx_fractal &fr, double bailout, bool addc, bool julia)
{
	const mbx_fractal_ifs &ifs = fr.IFS;
	Rot rots[MBX_IFS_VECTOR_COUNT], mainRot;
	D3 dirs[MBX_IFS_VECTOR_COUNT];
	for (int i = 0; i < MBX_IFS_VECTOR_COUNT; i++)
	{
		rots[i].z(ifs.rotations[i][0] * M_PI / 180.0);
		rots[i].y(ifs.rotations[i][1] * M_PI / 180.0);
		rots[i].x(ifs.rotations[i][2] * M_PI / 180.0);
		const D3 d = d3(ifs.direction[i]);
		const double l = len(d);
		dirs[i] = l > 0.0 ? mul(d, 1.0 / l) : d;
	}
	mainRot.z(ifs.rotation[0] * M_PI / 180.0);
	mainRot.y(ifs.rotation[1] * M_PI / 180.0);
	mainRot.x(ifs.rotation[2] * M_PI / 180.0);
	const D3 off = d3(ifs.offset);
	D3 z = d3(p);
	double DE = 1.0, r = len(z);
	const double sm1 = ifs.scale - 1.0;
	for (int it = 0; it < fr.formulaMaxiter; it++)
	{
		if (ifs.absX) z.x = std::fabs(z.x);
		if (ifs.absY) z.y = std::fabs(z.y);
		if (ifs.absZ) z.z = std::fabs(z.z);
		for (int i = 0; i < MBX_IFS_VECTOR_COUNT; i++)
		{
			if (!ifs.enabled[i]) continue;
			z = rots[i].apply(z);
			const double length = dot(z, dirs[i]);
			if (length < ifs.distance[i]) z = sub(z, mul(dirs[i], 2.0 * (length - ifs.distance[i]) * ifs.intensity[i]));
		}
		if (ifs.rotationEnabled) z = add(mainRot.apply(sub(z, off)), off);
		if (ifs.edgeEnabled)
		{
			if (ifs.edge[0] > 0.0) z.x = ifs.edge[0] - std::fabs(ifs.edge[0] - z.x);
			if (ifs.edge[1] > 0.0) z.y = ifs.edge[1] - std::fabs(ifs.edge[1] - z.y);
			if (ifs.edge[2] > 0.0) z.z = ifs.edge[2] - std::fabs(ifs.edge[2] - z.z);
		}
		z = mul(z, ifs.scale);
		if (ifs.mengerSpongeMode)
		{
			z.x -= off.x * sm1;
			z.y -= off.y * sm1;
			if (z.z > 0.5 * off.z * sm1) z.z -= off.z * sm1;
		}
		else
			z = sub(z, mul(off, sm1));
		DE *= std::fabs(ifs.scale);
		if (addc)
		{
			const double *c = julia ? fr.juliaC : p;
			z = D3{z.x + c[0] * fr.constantFactor[0], z.y + c[1] * fr.constantFactor[1], z.z + c[2] * fr.constantFactor[2]};
		}
		r = len(z);
		if (!std::isfinite(r) || r > bailout) break;
	}
	if (!std::isfinite(r) || !(DE > 0.0)) return 0.0;
	return (r - 2.0) / DE;
}

// Whether FP32 rounding of a coordinate near the nearest hit (|camera| + DE at the camera, the DE a
// bound on the distance to the nearest surface) exceeds 1/256 of the view's smallest distance
// threshold: the FP32 kinds then lose hits and iteration counts, and the view takes an FP64 kernel.
bool view_needs_fp64(const mbx_params *p, double de0, double threshK)
{
	de0 = std::fabs(de0);
	const double thrMin = p->constantDEThreshold ? p->DEThresh : de0 * threshK;
	const double mag = std::max(std::fabs(p->camera[0]), std::max(std::fabs(p->camera[1]), std::fabs(p->camera[2]))) + de0;
	return mag * 0x1p-24 > thrMin / 256.0;
}

// cAbstractFractal metadata of the formulas of the general path, as the constructors of
// formula/definition/fractal_<name>.cpp set it (abstract_fractal.h:85-93): DE type, DE
// function, c addition, default bailout, analytic finish, colour index function, names
struct FormulaMeta
{
	int id, deType, deFunction, cpixel;
	double defaultBailout;
	int analytic, coloring;
	const char *internalName, *nameInComboBox;
};
constexpr FormulaMeta kFormulaTable[] = {
	{MBX_FORMULA_MANDELBULB, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LOGARITHMIC, MBX_CPIXEL_ENABLED_BY_DEFAULT, 10.0,
		MBX_ANALYTIC_LOGARITHMIC, MBX_COLORING_DEFAULT, "mandelbulb", "Mandelbulb"},
	{MBX_FORMULA_MANDELBULB_POWER2, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LOGARITHMIC, MBX_CPIXEL_ENABLED_BY_DEFAULT,
		10.0, MBX_ANALYTIC_LOGARITHMIC, MBX_COLORING_DEFAULT, "mandelbulb_power2", "Mandelbulb - Power 2"},
	{MBX_FORMULA_HYPERCOMPLEX, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LOGARITHMIC, MBX_CPIXEL_ENABLED_BY_DEFAULT, 10.0,
		MBX_ANALYTIC_LOGARITHMIC, MBX_COLORING_DEFAULT, "hypercomplex", "Hypercomplex"},
	{MBX_FORMULA_QUATERNION, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LOGARITHMIC, MBX_CPIXEL_ENABLED_BY_DEFAULT, 10.0,
		MBX_ANALYTIC_LOGARITHMIC, MBX_COLORING_DEFAULT, "quaternion", "Quaternion"},
	{MBX_FORMULA_MENGER_SPONGE, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LINEAR, MBX_CPIXEL_DISABLED_BY_DEFAULT, 10.0,
		MBX_ANALYTIC_IFS, MBX_COLORING_IFS, "menger_sponge", "Menger sponge"},
	{MBX_FORMULA_MANDELBOX, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LINEAR, MBX_CPIXEL_ENABLED_BY_DEFAULT, 100.0,
		MBX_ANALYTIC_LINEAR, MBX_COLORING_ABOX, "mandelbox", "Mandelbox"},
	{MBX_FORMULA_KALEIDOSCOPIC_IFS, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LINEAR, MBX_CPIXEL_DISABLED_BY_DEFAULT, 100.0,
		MBX_ANALYTIC_IFS, MBX_COLORING_IFS, "kaleidoscopic_ifs", "Kaleidoscopic IFS"},
	{MBX_FORMULA_XENODREAMBUIE, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LOGARITHMIC, MBX_CPIXEL_ENABLED_BY_DEFAULT, 10.0,
		MBX_ANALYTIC_LOGARITHMIC, MBX_COLORING_DEFAULT, "xenodreambuie", "Xenodreambuie"},
	{MBX_FORMULA_BOX_FOLD_BULB_POW2, MBX_DE_DELTA, MBX_DE_FUNCTION_LINEAR, MBX_CPIXEL_ENABLED_BY_DEFAULT, 10000.0,
		MBX_ANALYTIC_LINEAR, MBX_COLORING_DEFAULT, "box_fold_bulb_pow2", "Box Fold Bulb Pow 2"},
	{MBX_FORMULA_AEXION, MBX_DE_ANALYTIC, MBX_DE_FUNCTION_LOGARITHMIC, MBX_CPIXEL_ALREADY_HAS, 10000.0,
		MBX_ANALYTIC_LOGARITHMIC, MBX_COLORING_DEFAULT, "aexion", "Aexion"},
};
constexpr int kFormulaCount = (int)(sizeof(kFormulaTable) / sizeof(kFormulaTable[0]));

const FormulaMeta *formula_meta(int id)
{
	for (const FormulaMeta &m : kFormulaTable)
		if (m.id == id) return &m;
	return nullptr;
}

// SetRotation2 of angles in degrees (algebra.cpp:297-303: Z(z), Y(y), X(x)), as
// RecalculateFractalParams sets mandelbox.mainRot (fractal.cpp:663); returns false for the
// identity (all angles zero), which RotateVector skips
bool rot2_deg(const double a[3], double out[9])
{
	Rot r;
	r.z(a[2] * kPi180);
	r.y(a[1] * kPi180);
	r.x(a[0] * kPi180);
	std::memcpy(out, r.m, sizeof(r.m));
	return !(a[0] == 0.0 && a[1] == 0.0 && a[2] == 0.0);
}

// SetRotation3 of angles in degrees (algebra.cpp:305-311); identity when all are zero
void rot3_deg(const double a[3], double out[9], int32_t &identity)
{
	Rot r;
	r.z(a[0] * kPi180);
	r.y(a[1] * kPi180);
	r.x(a[2] * kPi180);
	std::memcpy(out, r.m, sizeof(r.m));
	identity = (a[0] == 0.0 && a[1] == 0.0 && a[2] == 0.0) ? 1 : 0;
}

// One container of the general path: sFractal after RecalculateFractalParams (fractal.cpp:41-120,
// 653-678)
void gen_slot(const mbx_fractal &f, mbx::FractalParams &s)
{
	std::memset(&s, 0, sizeof(s));
	s.formula = f.formula;
	s.weight = f.formulaWeight;
	s.checkBailout = f.checkForBailout ? 1 : 0;
	s.power = f.bulb.power;
	s.alpha = f.bulb.alphaAngleOffset * kPi180;
	s.beta = f.bulb.betaAngleOffset * kPi180;
	const mbx_fractal_mandelbox &mb = f.mandelbox;
	s.mbScale = mb.scale;
	s.mbAbsScale = std::fabs(mb.scale);
	s.mbFoldingLimit = mb.foldingLimit;
	s.mbFoldingValue = mb.foldingValue;
	s.mbFR2 = mb.foldingSphericalFixed * mb.foldingSphericalFixed;
	s.mbMR2 = mb.foldingSphericalMin * mb.foldingSphericalMin;
	s.mbFactor1 = s.mbFR2 / s.mbMR2;
	for (int i = 0; i < 3; i++)
	{
		s.mbOffset[i] = mb.offset[i];
		s.mbColor[i] = mb.colorFactor[i];
		s.ifsOffset[i] = f.IFS.offset[i];
		s.ifsEdge[i] = f.IFS.edge[i];
	}
	s.mbSp1 = mb.colorFactorSp1;
	s.mbSp2 = mb.colorFactorSp2;
	s.mbMainRotOn = (rot2_deg(mb.rotationMain, s.mbMainRot) && mb.mainRotationEnabled) ? 1 : 0;
	// rotated folds: rot[fold][axis] = SetRotation2, rotinv = its transpose (fractal.cpp:665-672)
	s.mbRotOn = mb.rotationsEnabled ? 1 : 0;
	for (int fold = 0; fold < MBX_MANDELBOX_FOLDS; fold++)
		for (int axis = 0; axis < 3; axis++)
		{
			double *m = s.mbRot[fold][axis];
			s.mbRotId[fold][axis] = rot2_deg(mb.rotation[fold][axis], m) ? 0 : 1;
			double *t = s.mbRotInv[fold][axis];
			for (int r = 0; r < 3; r++)
				for (int c = 0; c < 3; c++) t[3 * r + c] = m[3 * c + r];
		}
	const mbx_fractal_ifs &ifs = f.IFS;
	s.ifsScale = ifs.scale;
	rot3_deg(ifs.rotation, s.ifsMainRot, s.ifsMainRotId);
	for (int i = 0; i < MBX_IFS_VECTOR_COUNT; i++)
	{
		rot3_deg(ifs.rotations[i], s.ifsRot[i], s.ifsRotId[i]);
		// CVector4::Normalize with w = 0 (fractal.cpp:653-660)
		const double d[4] = {ifs.direction[i][0], ifs.direction[i][1], ifs.direction[i][2], 0.0};
		const double norm = 1.0 / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + d[3] * d[3]);
		for (int k = 0; k < 4; k++) s.ifsDirection[i][k] = d[k] * norm;
		s.ifsDistance[i] = ifs.distance[i];
		s.ifsIntensity[i] = ifs.intensity[i];
		s.ifsEnabled[i] = ifs.enabled[i] ? 1 : 0;
	}
	s.ifsAbs[0] = ifs.absX;
	s.ifsAbs[1] = ifs.absY;
	s.ifsAbs[2] = ifs.absZ;
	s.ifsMenger = ifs.mengerSpongeMode;
	s.ifsRotationEnabled = ifs.rotationEnabled;
	s.ifsEdgeEnabled = ifs.edgeEnabled;
	const mbx_fractal_transform &t = f.transform;
	s.scale3 = t.scale3;
	s.cadd = t.aexionCadd;
	s.fipFoldFactor = t.foldingIntPowFoldFactor;
	s.fipZFactor = t.foldingIntPowZFactor;
	s.adeScale1 = t.analyticDEScale1;
	s.adeOffset1 = t.analyticDEOffset1;
	s.adeOffset2 = t.analyticDEOffset2;
	s.adeEnabled = t.analyticDEEnabled ? 1 : 0;
	s.adeEnabledFalse = t.analyticDEEnabledFalse ? 1 : 0;
}

// The sequence of the general path: cHybridFractalSequences::CollectSequenceData and
// CreateSequence (hybrid_fractal_sequences.cpp:181-287, 290-503) for one container (hybridCount 0)
// or the hybrid of p->fractal and p->hybridFractals[0 .. hybridCount - 2].
int derive_generic(mbx_ctx *ctx, const mbx_params *p, FrameConst &fc)
{
	const mbx_fractal &node = p->fractal;
	const bool hybrid = p->hybridCount >= 1;
	const int n = hybrid ? p->hybridCount : 1;
	if (n > MBX_MAX_FRACTALS) return fail(ctx, MBX_ERR_INVALID, "hybridCount exceeds MBX_MAX_FRACTALS");
	const mbx_fractal *c[MBX_MAX_FRACTALS];
	c[0] = &p->fractal;
	for (int k = 1; k < n; k++) c[k] = &p->hybridFractals[k - 1];
	fc.genSlots = n;
	fc.genHybrid = hybrid ? 1 : 0;
	fc.genInitialW = node.initialWAxis;
	fc.genLinearDEOffset = p->linearDEOffset;
	double maxBailout = 0.0;
	bool any = false;
	for (int k = 0; k < n; k++)
	{
		mbx::FractalParams &s = fc.gen[k];
		gen_slot(*c[k], s);
		if (c[k]->formula == MBX_FORMULA_NONE) continue;
		const FormulaMeta *m = formula_meta(c[k]->formula);
		if (!m) return fail(ctx, MBX_ERR_UNSUPPORTED, "formula not supported");
		if (c[k]->formulaIterations < 1) return fail(ctx, MBX_ERR_INVALID, "formula_iterations < 1");
		any = true;
		if (!hybrid) s.checkBailout = 1;
		bool addc = false;
		if (m->cpixel != 2)
		{
			addc = !c[k]->dontAddCConstant;
			if (m->cpixel == 1) addc = !addc;
		}
		s.addC = addc ? 1 : 0;
		if (node.useDefaultBailout)
		{
			if (hybrid)
				maxBailout = std::max(maxBailout, m->defaultBailout);
			else
				s.bailout = m->defaultBailout;
		}
		else
			s.bailout = node.bailout;
	}
	if (!any) return fail(ctx, MBX_ERR_INVALID, "every formula of the sequence is none");
	if (hybrid && node.useDefaultBailout)
		for (int k = 0; k < n; k++) fc.gen[k].bailout = maxBailout;

	// the formula of each iteration: formula_iterations of each container in turn, repeating from
	// repeat_from, skipping none formulas and containers outside their iteration range
	const int repeatFromIndex = std::max(0, std::min(p->repeatFrom - 1, n - 1));
	int fractalNo = 0, counter = 0, last = 0;
	bool rapidEnd = false;
	fc.genSeqLength = mbx::kGenSeqLength;
	for (int i = 0; i < mbx::kGenSeqLength; i++)
	{
		counter++;
		int search = 0;
		while ((c[fractalNo]->formula == MBX_FORMULA_NONE || i < c[fractalNo]->formulaStartIteration ||
				   i > c[fractalNo]->formulaStopIteration) && search < n)
		{
			fractalNo++;
			if (fractalNo >= n) fractalNo = repeatFromIndex;
			search++;
		}
		if (search >= n)
		{
			rapidEnd = true;
			break;
		}
		fc.genSeq[i] = (uint8_t)fractalNo;
		last = i;
		if (counter >= c[fractalNo]->formulaIterations)
		{
			counter = 0;
			fractalNo++;
			if (fractalNo >= n) fractalNo = repeatFromIndex;
		}
	}
	if (rapidEnd) fc.genSeqLength = last + 1;

	// DE type and function (hybrid_fractal_sequences.cpp:379-503)
	bool forceDelta = node.deltaDEMethod == 1;
	const bool forceAnalytic = node.deltaDEMethod == 2;
	int deType = MBX_DE_ANALYTIC, deFunction = 0, analytic = MBX_ANALYTIC_LINEAR, coloring = 0;
	if (hybrid)
	{
		if (node.deltaDEFunction == 0)
		{
			// the DE function used for the most iterations; linear first on a tie
			int count[7] = {0, 0, 0, 0, 0, 0, 0};
			for (int k = 0; k < n; k++)
			{
				if (c[k]->formula == MBX_FORMULA_NONE) continue;
				const FormulaMeta *m = formula_meta(c[k]->formula);
				count[m->deFunction] += c[k]->formulaIterations;
				if (!forceDelta && !forceAnalytic && m->deType == MBX_DE_DELTA)
				{
					deType = MBX_DE_DELTA;
					forceDelta = true;
				}
			}
			int maxCount = -1;
			for (int i = 1; i <= 6; i++)
				if (count[i] > maxCount)
				{
					maxCount = count[i];
					deFunction = i;
				}
		}
		else
		{
			deFunction = node.deltaDEFunction;
			for (int k = 0; k < n; k++)
			{
				const FormulaMeta *m = formula_meta(c[k]->formula);
				if (m && m->deType == MBX_DE_DELTA)
				{
					deType = MBX_DE_DELTA;
					break;
				}
			}
		}
		if (forceDelta) deType = MBX_DE_DELTA;
		if (forceAnalytic) deType = MBX_DE_ANALYTIC;
	}
	else
	{
		const FormulaMeta *m = formula_meta(c[0]->formula);
		deType = m->deType;
		deFunction = m->deFunction;
		analytic = m->analytic;
		coloring = m->coloring;
		if (forceDelta) deType = MBX_DE_DELTA;
		if (forceAnalytic) deType = MBX_DE_ANALYTIC;
		if (node.deltaDEFunction != 0)
		{
			deFunction = node.deltaDEFunction;
			analytic = deFunction == MBX_DE_FUNCTION_LOGARITHMIC ? MBX_ANALYTIC_LOGARITHMIC : MBX_ANALYTIC_LINEAR;
		}
	}
	if (deFunction != MBX_DE_FUNCTION_LINEAR && deFunction != MBX_DE_FUNCTION_LOGARITHMIC)
		return fail(ctx, MBX_ERR_UNSUPPORTED, "DE function not supported (pseudo-Kleinian, custom, max-axis)");
	fc.genDEType = deType;
	fc.genDEFunction = deFunction;
	fc.genAnalytic = analytic;
	fc.genColoring = coloring;
	return MBX_OK;
}

// Derive the per-frame constants; returns MBX_OK or an error with ctx->err set.  forceGeneric
// routes the frame to the general formula kernel whatever its formula (batches that mix kinds).
int derive_frame(mbx_ctx *ctx, const mbx_params *p, FrameConst &fc, bool forceGeneric)
{
	std::memset(&fc, 0, sizeof(fc));
	if (p->imageWidth <= 0 || p->imageHeight <= 0 || p->imageWidth > 65536 || p->imageHeight > 65536)
		return fail(ctx, MBX_ERR_INVALID, "image size out of range");
	if (p->perspectiveType < 0 || p->perspectiveType > 3)
		return fail(ctx, MBX_ERR_INVALID, "perspectiveType must be 0 .. 3");
	// MonteCarloDOF offsets the lens in another plane for the other projections (render_worker.cpp:1845-1866)
	if (p->perspectiveType != 0 && p->DOFMonteCarlo && p->DOFEnabled)
		return fail(ctx, MBX_ERR_UNSUPPORTED, "Monte Carlo depth of field with a fish-eye or equirectangular projection");
	if (p->antialiasingEnabled && (p->antialiasingSize < 1 || p->antialiasingSize > 10))
		return fail(ctx, MBX_ERR_INVALID, "antialiasingSize must be 1 .. 10");
	// the projections and antialiasing are built into the full kernel variant, not the lens
	// variant that carries the Monte Carlo lens, chromatic aberration and fog
	if ((p->perspectiveType != 0 || p->antialiasingEnabled) &&
		((p->DOFMonteCarlo && (p->DOFEnabled || p->DOFMonteCarloChromaticAberration)) || p->basicFogEnabled ||
			p->volFogEnabled))
		return fail(ctx, MBX_ERR_UNSUPPORTED,
			"a fish-eye or equirectangular projection or antialiasing with the Monte Carlo lens, chromatic "
			"aberration or fog");
	if (p->maxRaymarchingSteps < 1) return fail(ctx, MBX_ERR_INVALID, "max_raymarching_steps < 1");
	if (!(p->detailLevel > 0.0)) return fail(ctx, MBX_ERR_INVALID, "detail_level must be > 0");
	const mbx_fractal &fr = p->fractal;
	if (fr.formulaMaxiter < 1) return fail(ctx, MBX_ERR_INVALID, "formula maxiter < 1");
	if (p->hybridCount < 0 || p->hybridCount > MBX_MAX_FRACTALS)
		return fail(ctx, MBX_ERR_INVALID, "hybridCount out of range");
	// the specialised kinds cover one Mandelbulb, Mandelbox or IFS with w = 0; everything else
	// takes the general FP64 path
	// the general kernel also takes the Mandelbox with rotated folds
	// and interior / iteration-threshold modes
	bool generic = forceGeneric || p->hybridCount >= 1 || fr.initialWAxis != 0.0 || p->boxFolding ||
		p->sphericalFolding || p->interiorMode || p->iterThreshMode ||
		(fr.formula == MBX_FORMULA_MANDELBOX && fr.mandelbox.rotationsEnabled) ||
		(fr.formula != MBX_FORMULA_MANDELBULB && fr.formula != MBX_FORMULA_MANDELBOX &&
			fr.formula != MBX_FORMULA_KALEIDOSCOPIC_IFS);

	// camera: PrepareMainVectors (render_worker.cpp:539-564)
	const CameraFrame cf = camera_frame(d3(p->camera), d3(p->target), d3(p->topVector));
	Rot m;
	m.z(cf.yaw);
	m.x(cf.pitch);
	m.y(cf.roll);
	m.z(-(p->sweetSpotHAngle / 180.0 * M_PI));
	m.x(p->sweetSpotVAngle / 180.0 * M_PI);
	std::memcpy(ctx->mRot, m.m, sizeof(m.m));
	// CalcFOV (projection_3d.cpp:140-159); the equirectangular image spans 2 x 1 of the image plane
	const double fov = p->perspectiveType == 0 ? 2.0 * std::tan(p->fov / 360.0 * M_PI) : p->fov / 180.0 * M_PI;
	ctx->fov = fov;
	const double resolution = 1.0 / p->imageHeight;
	const double aspect = p->perspectiveType == 2 ? 2.0 : (double)p->imageWidth / p->imageHeight;
	// the equirectangular projection halves the threshold and the delta (render_worker.cpp:669, 688):
	// a factor of a power of two, exact wherever it is folded in
	const double tscale = p->perspectiveType == 2 ? 0.5 : 1.0;
	fc.perspective = p->perspectiveType;
	fc.threshScaleD = tscale;
	fc.fishCutD = M_PI * 0.5 / fov;
	fc.aaSize = p->antialiasingEnabled ? p->antialiasingSize : 0;

	fc.width = p->imageWidth;
	fc.height = p->imageHeight;
	fc.invWidth = (float)(1.0 / p->imageWidth);
	fc.invHeight = (float)(1.0 / p->imageHeight);
	fc.aspectRatio = (float)aspect;
	fc.fov = (float)fov;
	setRot(fc.mRot, m);
	set3f(fc.camera, p->camera);
	for (int i = 0; i < 3; i++) fc.cameraD[i] = p->camera[i];
	for (int i = 0; i < 9; i++) fc.mRotD[i] = m.m[i];
	fc.fovD = fov;
	fc.aspectRatioD = aspect;
	fc.invWidthD = 1.0 / p->imageWidth;
	fc.invHeightD = 1.0 / p->imageHeight;
	fc.threshKD = resolution * fov / p->detailLevel * tscale;
	fc.resolutionD = resolution;
	fc.detailLevelD = p->detailLevel;
	fc.DEFactorD = p->DEFactor;
	fc.deltaKD = resolution * fov * tscale;
	fc.threshK = (float)fc.threshKD;
	fc.deltaK = (float)fc.deltaKD;
	// iteration-threshold mode takes precedence over the constant threshold (render_worker.cpp:656-665)
	fc.constantDEThreshold = (p->constantDEThreshold && !p->iterThreshMode) ? 1 : 0;
	fc.interiorMode = p->interiorMode ? 1 : 0;
	fc.iterThreshMode = p->iterThreshMode ? 1 : 0;
	fc.DEThresh = (float)(p->DEThresh * tscale);
	fc.DEThreshD = p->DEThresh * tscale;
	fc.DEFactor = (float)p->DEFactor;
	fc.searchLimit = (float)(1.0 - 0.001 * p->detailLevel);
	fc.viewDistanceMax = (float)p->viewDistanceMax;
	fc.smoothness = (float)p->smoothness;
	fc.smoothnessD = p->smoothness;
	fc.maxRaymarchingSteps = p->maxRaymarchingSteps;
	fc.minN = p->minN;
	fc.advancedQuality = p->advancedQuality ? 1 : 0;
	fc.absMaxStepD = p->absMaxMarchingStep;
	fc.absMinStepD = p->absMinMarchingStep;
	fc.relMaxStepD = p->relMaxMarchingStep;
	fc.relMinStepD = p->relMinMarchingStep;
	fc.detailSizeMaxD = p->detailSizeMax;
	fc.detailSizeMinD = p->detailSizeMin;
	fc.absMaxStep = (float)p->absMaxMarchingStep;
	fc.absMinStep = (float)p->absMinMarchingStep;
	fc.relMaxStep = (float)p->relMaxMarchingStep;
	fc.relMinStep = (float)p->relMinMarchingStep;
	fc.detailSizeMax = (float)p->detailSizeMax;
	fc.detailSizeMin = (float)p->detailSizeMin;
	fc.deltaRel = p->advancedQuality ? p->deltaDERelativeDelta : 1e-5;

	// fractal sequence data (hybrid_fractal_sequences.cpp:181-287, 471-503)
	int deType = 0, deFunction = 0, analyticFunction = 0;
	double defaultBailout = 0.0;
	bool cpixelDisabled = false;
	switch (fr.formula)
	{
		case MBX_FORMULA_MANDELBULB:
			deFunction = MBX_DE_FUNCTION_LOGARITHMIC;
			analyticFunction = MBX_ANALYTIC_LOGARITHMIC;
			defaultBailout = 10.0;
			break;
		case MBX_FORMULA_MANDELBOX:
			deFunction = MBX_DE_FUNCTION_LINEAR;
			analyticFunction = MBX_ANALYTIC_LINEAR;
			defaultBailout = 100.0;
			break;
		case MBX_FORMULA_KALEIDOSCOPIC_IFS:
			deFunction = MBX_DE_FUNCTION_LINEAR;
			analyticFunction = MBX_ANALYTIC_IFS;
			defaultBailout = 100.0;
			cpixelDisabled = true;
			break;
		default:
			if (!generic) return fail(ctx, MBX_ERR_UNSUPPORTED, "formula not supported");
			break;
	}
	if (fr.deltaDEMethod == 1) deType = MBX_DE_DELTA;
	if (fr.deltaDEMethod == 2) deType = MBX_DE_ANALYTIC;
	if (fr.deltaDEFunction != 0)
	{
		deFunction = fr.deltaDEFunction;
		if (deFunction == MBX_DE_FUNCTION_LOGARITHMIC)
			analyticFunction = MBX_ANALYTIC_LOGARITHMIC;
		else if (deFunction == MBX_DE_FUNCTION_LINEAR)
			analyticFunction = MBX_ANALYTIC_LINEAR;
		else
			return fail(ctx, MBX_ERR_UNSUPPORTED, "delta_DE_function not supported");
	}
	bool addc = !fr.dontAddCConstant;
	if (cpixelDisabled) addc = !addc;
	fc.addC = addc;
	fc.julia = fr.juliaMode;
	for (int i = 0; i < 3; i++)
	{
		fc.juliaC[i] = (float)(fr.juliaC[i] * fr.constantFactor[i]);
		fc.constMul[i] = (float)fr.constantFactor[i];
	}
	fc.initialW = 0.0f;
	fc.bailout = (float)(fr.useDefaultBailout ? defaultBailout : fr.bailout);
	fc.bailoutD = fr.useDefaultBailout ? defaultBailout : fr.bailout;
	fc.formulaMaxiter = fr.formulaMaxiter;
	fc.deFunction = deFunction;
	fc.analyticFunction = analyticFunction;
	fc.power = (float)fr.bulb.power;
	fc.alpha = (float)(fr.bulb.alphaAngleOffset * M_PI / 180.0);
	fc.beta = (float)(fr.bulb.betaAngleOffset * M_PI / 180.0);
	// RecalculateFractalParams (fractal.cpp:677-678)
	fc.powerD = fr.bulb.power;
	fc.alphaD = fr.bulb.alphaAngleOffset * kPi180;
	fc.betaD = fr.bulb.betaAngleOffset * kPi180;

	const mbx_fractal_mandelbox &mb = fr.mandelbox;
	fc.mbScale = (float)mb.scale;
	fc.mbAbsScale = (float)std::fabs(mb.scale);
	fc.mbFoldingLimit = (float)mb.foldingLimit;
	fc.mbFoldingValue = (float)mb.foldingValue;
	const double fR2 = mb.foldingSphericalFixed * mb.foldingSphericalFixed;
	const double mR2 = mb.foldingSphericalMin * mb.foldingSphericalMin;
	fc.mbFR2 = (float)fR2;
	fc.mbMR2 = (float)mR2;
	fc.mbFactor1 = (float)(fR2 / mR2);
	set3f(fc.mbOffset, mb.offset);
	fc.deltaTangent = ctx->deltaTangent;
	fc.mbScaleD = mb.scale;
	fc.mbMainRotOn = (rot2_deg(mb.rotationMain, fc.mbMainRotD) && mb.mainRotationEnabled) ? 1 : 0;
	fc.mbAbsScaleD = std::fabs(mb.scale);
	fc.mbFoldingLimitD = mb.foldingLimit;
	fc.mbFoldingValueD = mb.foldingValue;
	fc.mbFR2D = fR2;
	fc.mbMR2D = mR2;
	fc.mbFactor1D = fR2 / mR2;
	for (int i = 0; i < 3; i++)
	{
		fc.mbOffsetD[i] = mb.offset[i];
		fc.juliaCD[i] = fr.juliaC[i] * fr.constantFactor[i];
		fc.constMulD[i] = fr.constantFactor[i];
	}

	const mbx_fractal_ifs &ifs = fr.IFS;
	fc.ifsScale = (float)ifs.scale;
	fc.ifsAbsScale = (float)std::fabs(ifs.scale);
	set3f(fc.ifsOffset, ifs.offset);
	set3f(fc.ifsEdge, ifs.edge);
	fc.ifsAbs[0] = ifs.absX;
	fc.ifsAbs[1] = ifs.absY;
	fc.ifsAbs[2] = ifs.absZ;
	fc.ifsMenger = ifs.mengerSpongeMode;
	fc.ifsRotationEnabled = ifs.rotationEnabled;
	fc.ifsEdgeEnabled = ifs.edgeEnabled;
	{
		// SetRotation3: Z(x), Y(y), X(z) (algebra.cpp:305-311)
		Rot r;
		r.z(ifs.rotation[0] * M_PI / 180.0);
		r.y(ifs.rotation[1] * M_PI / 180.0);
		r.x(ifs.rotation[2] * M_PI / 180.0);
		setRot(fc.ifsMainRot, r);
	}
	int enabledCount = 0;
	for (int i = 0; i < MBX_IFS_VECTOR_COUNT; i++)
	{
		Rot r;
		r.z(ifs.rotations[i][0] * M_PI / 180.0);
		r.y(ifs.rotations[i][1] * M_PI / 180.0);
		r.x(ifs.rotations[i][2] * M_PI / 180.0);
		setRot(fc.ifsRot[i], r);
		const D3 d = d3(ifs.direction[i]);
		const double l = len(d);
		set3f(fc.ifsDirection[i], l > 0.0 ? mul(d, 1.0 / l) : d);
		fc.ifsDistance[i] = (float)ifs.distance[i];
		fc.ifsIntensity[i] = (float)ifs.intensity[i];
		fc.ifsEnabled[i] = ifs.enabled[i];
		enabledCount += ifs.enabled[i] ? 1 : 0;
	}
	fc.ifsEnabledCount = enabledCount;

	// the IFS kind marches in FP32; a close-up that FP32 cannot resolve takes the general FP64
	// kernel (MBX_TUNE_PRECISION forces either)
	if (!generic && fr.formula == MBX_FORMULA_KALEIDOSCOPIC_IFS && deType == MBX_DE_ANALYTIC && ctx->precision != 1)
	{
		bool fp64 = ctx->precision == 2;
		if (!fp64) fp64 = view_needs_fp64(p, host_ifs_de(p->camera, fr, fc.bailoutD, fc.addC, fc.julia), fc.threshKD);
		generic = fp64;
	}

	// kernel specialisation
	int kind = -1;
	if (generic)
	{
		const int rc = derive_generic(ctx, p, fc);
		if (rc != MBX_OK) return rc;
		kind = mbx::kGeneric;
	}
	else if (fr.formula == MBX_FORMULA_MANDELBULB && deType == MBX_DE_ANALYTIC)
	{
		const bool pow8 = fr.bulb.power == 8.0 && fr.bulb.alphaAngleOffset == 0.0 &&
			fr.bulb.betaAngleOffset == 0.0 && analyticFunction == MBX_ANALYTIC_LOGARITHMIC &&
			fc.addC && !fc.julia; // the power-8 kernel adds c = p * constantMultiplier
		kind = pow8 ? mbx::kBulb8 : mbx::kBulbTrig;
		// FP64 march when FP32 rounding of a coordinate near the nearest hit (|camera| + DE at the
		// camera) exceeds 1/256 of the smallest distance threshold of the view
		bool fp64 = ctx->precision == 2;
		if (ctx->precision == 0)
			fp64 = view_needs_fp64(p, host_bulb_de(p->camera, fr, fc.bailoutD, analyticFunction, fc.addC, fc.julia),
				fc.threshKD);
		if (fp64) kind = mbx::kBulbD;
		if (analyticFunction != MBX_ANALYTIC_LOGARITHMIC && analyticFunction != MBX_ANALYTIC_LINEAR)
			kind = -1;
	}
	else if (fr.formula == MBX_FORMULA_MANDELBOX)
	{
		if (mb.rotationsEnabled)
			return fail(ctx, MBX_ERR_UNSUPPORTED, "Mandelbox rotated folds are not supported on the device");
		kind = deType == MBX_DE_DELTA ? mbx::kBoxDelta : mbx::kBoxAnalytic;
	}
	else if (fr.formula == MBX_FORMULA_KALEIDOSCOPIC_IFS && deType == MBX_DE_ANALYTIC)
		kind = mbx::kIfsAnalytic;
	if (kind < 0) return fail(ctx, MBX_ERR_UNSUPPORTED, "formula / DE mode combination not supported");
	fc.deKind = kind;
	fc.farR2 = std::numeric_limits<float>::infinity();
	// (not with the Monte Carlo lens: its rays start off the camera, outside the bound's premise)
	if (ctx->farField && (kind == mbx::kBulb8 || kind == mbx::kBulbTrig) && !fc.julia && fc.addC && !p->advancedQuality &&
		!(p->DOFMonteCarlo && p->DOFEnabled))
	{
		const double c = std::max(std::fabs(fr.constantFactor[0]),
			std::max(std::fabs(fr.constantFactor[1]), std::fabs(fr.constantFactor[2])));
		const double thrMax = p->constantDEThreshold ? p->DEThresh
													 : (p->viewDistanceMax + 3.0) * fc.threshKD * 1.01;
		fc.farR2 = far_field_r2(fr.bulb.power, analyticFunction == MBX_ANALYTIC_LOGARITHMIC, fc.bailoutD, c,
			p->DEFactor, thrMax);
	}

	// light #1 (light.cpp:80-195)
	const mbx_light &L = p->light;
	D3 rotation{L.rotation[0] / 180.8 * M_PI, L.rotation[1] / 180.8 * M_PI, L.rotation[2] / 180.8 * M_PI};
	rotation.y = -rotation.y; // directional light (light.cpp:92-95)
	D3 ld;
	if (L.relativePosition)
	{
		ld = mul(cf.forward, -1.0);
		ld = rotate_about(ld, cf.forward, rotation.z);
		ld = rotate_about(ld, cf.right, rotation.y);
		ld = rotate_about(ld, cf.top, rotation.x);
	}
	else
	{
		Rot lr; // SetRotation(CVector3): Z(x), X(y), Y(z)
		lr.z(rotation.x);
		lr.x(rotation.y);
		lr.y(rotation.z);
		ld = lr.apply(D3{0.0, -1.0, 0.0});
	}
	ctx->lightDir[0] = ld.x;
	ctx->lightDir[1] = ld.y;
	ctx->lightDir[2] = ld.z;
	fc.lightEnabled = L.enabled;
	fc.castShadows = L.castShadows;
	fc.penetrating = L.penetrating;
	set3f(fc.lightDir, ld);
	fc.lightDirD[0] = ld.x;
	fc.lightDirD[1] = ld.y;
	fc.lightDirD[2] = ld.z;
	fc.lightIntensity = (float)L.intensity;
	fc.lightVisibility = (float)L.visibility;
	fc.lightSize = (float)L.size;
	fc.lightContourSharpness = (float)L.contourSharpness;
	const double softRange = std::tan(L.softShadowCone / 180.0 * M_PI);
	fc.softRange = (float)softRange;
	// no soft shadows in iteration-threshold or interior mode (shader_aux_shadow.cpp:73-76)
	// and none with Monte Carlo soft shadows, which jitter the light instead (:76-87)
	fc.softShadows = softRange > 0.0 && !p->iterThreshMode && !p->interiorMode &&
		!(p->DOFMonteCarlo && p->monteCarloSoftShadows);
	fc.softRangeD = softRange;
	set3f(fc.lightColor, L.color);
	fc.viewDistanceMaxLight = (float)p->viewDistanceMax;

	// material, fill light
	set3f(fc.surfaceColor, p->surfaceColor);
	fc.shading = (float)p->shading;
	for (int i = 0; i < 3; i++)
	{
		fc.luminosity[i] = (float)(p->luminosity * p->luminosityColor[i]);
		fc.fillLight[i] = (float)p->fillLightColor[i];
		// ObjectShader's float products (shader_object.cpp:171-201)
		fc.lumEmissive[i] = (float)p->luminosity * (float)p->luminosityColor[i] * (float)p->luminosityEmissive;
	}

	// palette colouring: the material's colouring speed and offset are floats (material.h:88-89)
	fc.palette = p->useColorsFromPalette ? 1 : 0;
	fc.paletteGradient = p->surfaceGradientEnable ? 1 : 0;
	fc.colorAlgorithm = p->coloringAlgorithm;
	fc.colorPreV215 = p->coloringPreV215 ? 1 : 0;
	fc.color4D = p->coloring4D ? 1 : 0;
	fc.coloringSpeedD = (double)(float)p->coloringSpeed;
	fc.paletteOffsetD = (double)(float)p->coloringPaletteOffset;
	fc.colorSphereRadiusD = p->coloringSphereRadius;
	fc.colorSphereRadius = (float)p->coloringSphereRadius;
	for (int i = 0; i < 3; i++)
	{
		fc.colorLineDirD[i] = p->coloringLineDirection[i];
		fc.colorLineDir[i] = (float)p->coloringLineDirection[i];
		fc.mbColorD[i] = mb.colorFactor[i];
	}
	fc.mbSp1D = mb.colorFactorSp1;
	fc.mbSp2D = mb.colorFactorSp2;
	fc.mbColorRD = mb.colorFactorR;
	if (fc.palette && fc.paletteGradient)
	{
		if (p->coloringAlgorithm < 0 || p->coloringAlgorithm > 4)
			return fail(ctx, MBX_ERR_INVALID, "coloringAlgorithm must be 0 .. 4");
		const int n = p->surfaceGradientCount;
		if (n < 2 || n > MBX_MAX_GRADIENT_STOPS)
			return fail(ctx, MBX_ERR_INVALID, "surfaceGradientCount must be 2 .. MBX_MAX_GRADIENT_STOPS");
		for (int i = 1; i < n; i++)
			if (!((float)p->surfaceGradientPosition[i] >= (float)p->surfaceGradientPosition[i - 1]))
				return fail(ctx, MBX_ERR_INVALID, "surfaceGradientPosition must be sorted (non-decreasing)");
	}
	fc.gradientCount = p->surfaceGradientCount;
	for (int i = 0; i < MBX_MAX_GRADIENT_STOPS; i++)
	{
		const bool used = i < p->surfaceGradientCount;
		fc.gradientPos[i] = used ? (float)p->surfaceGradientPosition[i] : 0.0f;
		for (int c = 0; c < 3; c++) fc.gradientColor[i][c] = used ? (float)p->surfaceGradientColor[i][c] : 0.0f;
	}

	// specular highlights: widths, intensities and roughness are the material's floats
	fc.specPlastic = p->specularPlasticEnable ? 1 : 0;
	fc.specMetal = p->metallic ? 1 : 0;
	if ((fc.specPlastic && !(p->specularWidth > 0.0)) || (fc.specMetal && !(p->specularMetallicWidth > 0.0)))
		return fail(ctx, MBX_ERR_INVALID, "specular widths must be > 0");
	fc.specPlasticExp = 30.0f / (float)p->specularWidth;
	fc.specMetalExp = 30.0f / (float)p->specularMetallicWidth;
	fc.specPlasticK = (float)p->specular;
	fc.specMetalK = (float)p->specularMetallic;
	fc.specRoughness = (float)p->specularMetallicRoughness;
	set3f(fc.specColor, p->specularColor);

	// glow and fog: VolumetricShader over the primary march's steps (shader_volumetric.cpp:70-700)
	fc.glow = p->glowEnabled ? 1 : 0;
	fc.basicFog = p->basicFogEnabled ? 1 : 0;
	fc.volFog = p->volFogEnabled ? 1 : 0;
	fc.fog = (fc.basicFog || fc.volFog) ? 1 : 0;
	if (fc.glow || fc.fog)
	{
		// every march step is at least 0.2 thresholds * DEFactor (0.45 outside interior mode), or
		// relMinMarchingStep thresholds with advanced quality: with these bounds no step falls
		// below the 1e-10 thresholds at which the reference merges steps (shader_volumetric.cpp:178)
		if (!(p->DEFactor >= 1e-9) || (p->advancedQuality && !(p->relMinMarchingStep >= 1e-9)))
			return fail(ctx, MBX_ERR_UNSUPPORTED,
				"glow and fog need DEFactor >= 1e-9 and, with advanced quality, relMinMarchingStep >= 1e-9 "
				"(the step-merging case is not covered)");
		if (p->volumetricLightDEFactor != 1.0)
			return fail(ctx, MBX_ERR_UNSUPPORTED,
				"a volumetric light DE factor other than 1 (the volumetric shader's own march) is not rendered");
	}
	if (p->iterFogEnabled || p->cloudsEnabled)
		return fail(ctx, MBX_ERR_UNSUPPORTED, "iteration fog and clouds are not rendered");
	if ((fc.basicFog && p->basicFogCastShadows) || (fc.volFog && p->distanceFogShadows))
		return fail(ctx, MBX_ERR_UNSUPPORTED, "shadows in fog are not rendered");
	fc.glowIntensity = (float)p->glowIntensity;
	set3f(fc.glowColor1, p->glowColor1);
	set3f(fc.glowColor2, p->glowColor2);
	{
		const float go = (float)p->glowIntensity / 512.0f * (float)p->DEFactor;
		fc.glowStepOpacity = go > 1.0f ? 1.0f : go;
	}
	fc.basicFogVisibilityD = p->basicFogVisibility;
	set3f(fc.basicFogColor, p->basicFogColor);
	fc.volFogDensity = (float)p->volFogDensity;
	fc.volFogFactorD = p->volFogDistanceFactor;
	fc.volFogFromSurfaceD = p->volFogDistanceFromSurface;
	fc.volFogColour1DistD = p->volFogColour1Distance;
	fc.volFogColour2DistD = p->volFogColour2Distance;
	for (int k = 0; k < 3; k++) set3f(fc.volFogColour[k], p->volFogColour[k]);

	// ambient occlusion; in screen-space mode the object shader takes no AO term
	// (shader_object.cpp:108-121) and mbx_render_ssao adds it to the finished image
	fc.aoEnabled = p->ambientOcclusionEnabled && p->ambientOcclusionMode != MBX_AO_SCREEN_SPACE;
	fc.aoMode = p->ambientOcclusionMode;
	fc.aoQuality = p->ambientOcclusionQuality;
	fc.aoCount = p->aoVectorsCount;
	fc.aoIntensity = (float)p->ambientOcclusion;
	for (int i = 0; i < 3; i++) fc.aoColor[i] = (float)(p->ambientOcclusion * p->ambientOcclusionColor[i]);
	fc.aoFastTune = (float)p->ambientOcclusionFastTune;
	if (p->ambientOcclusionEnabled && p->ambientOcclusionMode == MBX_AO_MULTIPLE_RAYS &&
		(p->aoVectorsCount < 1 || p->aoVectorsCount > MBX_MAX_AO_VECTORS))
		return fail(ctx, MBX_ERR_INVALID, "aoVectorsCount out of range");
	if (p->ambientOcclusionEnabled && p->ambientOcclusionMode != MBX_AO_MULTIPLE_RAYS &&
		p->ambientOcclusionMode != MBX_AO_FAST && p->ambientOcclusionMode != MBX_AO_SCREEN_SPACE)
		return fail(ctx, MBX_ERR_INVALID, "ambientOcclusionMode must be 0, 1 or 2");
	for (int i = 0; i < MBX_MAX_AO_VECTORS; i++)
	{
		set3f(fc.aoVectors[i], p->aoVectors[i]);
		for (int k = 0; k < 3; k++) fc.aoVectorsD[i][k] = p->aoVectors[i][k];
		set3f(fc.aoColors[i], p->aoColors[i]);
	}

	// background
	fc.background3Colors = p->background3ColorsEnable;
	set3f(fc.bg1, p->backgroundColor1);
	set3f(fc.bg2, p->backgroundColor2);
	set3f(fc.bg3, p->backgroundColor3);
	fc.bgBrightness = (float)p->backgroundBrightness;
	fc.bgInvGamma = (float)(1.0 / p->backgroundGamma);

	// Monte Carlo
	fc.monteCarlo = p->DOFMonteCarlo;
	fc.samples = p->DOFMonteCarlo ? p->DOFSamples : 1;
	fc.DOFMinSamples = p->DOFMinSamples;
	fc.dofEnabled = (p->DOFMonteCarlo && p->DOFEnabled) ? 1 : 0;
	fc.chromaticAberration = (p->DOFMonteCarlo && p->DOFMonteCarloChromaticAberration) ? 1 : 0;
	fc.mcSoftShadows = (p->DOFMonteCarlo && p->monteCarloSoftShadows) ? 1 : 0;
	fc.dofFocusD = p->DOFFocus;
	fc.dofRadiusD = p->DOFRadius;
	fc.caCameraDispersion = (float)p->DOFMonteCarloCACameraDispersion;
	fc.lensSinCos = ctx->d_mcTables;
	fc.softCbrt = ctx->d_mcTables ? ctx->d_mcTables + 2 * mbx::kLensAngles : nullptr;
	if (fc.dofEnabled && !(p->DOFFocus > 0.0)) return fail(ctx, MBX_ERR_INVALID, "DOF_focus must be > 0");
	fc.DOFMaxNoise01 = p->DOFMaxNoise * 0.01;
	if (fc.samples < 1) return fail(ctx, MBX_ERR_INVALID, "DOF_samples < 1");
	fc.globalIllumination = p->DOFMonteCarloGlobalIllumination;
	fc.reflectionsMax = p->reflectionsMax;
	fc.giRadianceLimit = (float)p->monteCarloGIRadianceLimit;
	if (p->rngMode != MBX_RNG_PER_PIXEL)
		return fail(ctx, MBX_ERR_UNSUPPORTED, "the device uses per-pixel random streams only");
	fc.randomSeed = p->randomSeed;
	// global foldings (fractparams.cpp:346-351; the general kernel applies them) and slow shading
	fc.boxFold = p->boxFolding ? 1 : 0;
	fc.sphFold = p->sphericalFolding ? 1 : 0;
	fc.boxFoldLimitD = p->boxFoldingLimit;
	fc.boxFoldValueD = p->boxFoldingValue;
	fc.sphFR2D = p->sphericalFoldingOuter * p->sphericalFoldingOuter;
	fc.sphMR2D = p->sphericalFoldingInner * p->sphericalFoldingInner;
	fc.sphFactor1D = fc.sphFR2D / fc.sphMR2D;
	fc.slowShading = p->slowShading ? 1 : 0;
	fc.slowShadingN = p->N;

	// fractal node rotation (objects_tree.cpp:292-382): local-to-world is R^T with
	// R = SetRotation2(rotation * pi/180), so world-to-local (calculate_distance.cpp:296) is R
	{
		const double *fr3 = p->formulaRotation;
		fc.nodeRotated = (fr3[0] != 0.0 || fr3[1] != 0.0 || fr3[2] != 0.0) ? 1 : 0;
		Rot nr;
		nr.z(fr3[2] * (M_PI / 180.0));
		nr.y(fr3[1] * (M_PI / 180.0));
		nr.x(fr3[0] * (M_PI / 180.0));
		setRot(fc.nodeRot, nr);
		for (int k = 0; k < 9; k++) fc.nodeRotD[k] = nr.m[k];
	}
	return MBX_OK;
}

int default_waves(const mbx_ctx *ctx) { return ctx->numCUs * 16; }

int ensure_capacity(mbx_ctx *ctx, int n)
{
	if (n <= ctx->capacity) return MBX_OK;
	hipError_t e;
	if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "sync");
	if (ctx->d_frames) (void)hipFree(ctx->d_frames);
	if (ctx->d_jobs) (void)hipFree(ctx->d_jobs);
	ctx->d_frames = nullptr;
	ctx->d_jobs = nullptr;
	ctx->capacity = 0;
	ctx->framesHost.clear();
	ctx->jobsHost.clear();
	if ((e = hipMalloc((void **)&ctx->d_frames, (size_t)n * sizeof(FrameConst))) != hipSuccess ||
		(e = hipMalloc((void **)&ctx->d_jobs, (size_t)n * sizeof(mbx::FrameJob))) != hipSuccess)
		return hip_fail(ctx, e, "allocate batch buffers");
	ctx->capacity = n;
	ctx->graphFrames = -1; // buffers moved: the captured graph is stale
	return MBX_OK;
}

mbx::FrameJob make_job(const FrameConst &fc, const mbx_outputs &o, int tileBase, int pixBase, int tileShift)
{
	mbx::FrameJob j{};
	const int tw = 1 << tileShift, th = 64 >> tileShift;
	j.tileEnd = tileBase + ((fc.width + tw - 1) / tw) * ((fc.height + th - 1) / th);
	j.pixBase = pixBase;
	j.rgb = o.rgb;
	j.depth = o.depth;
	j.iters = o.iters;
	j.mask = o.mask;
	j.normal = o.normal;
	j.steps = o.steps;
	j.colour = o.colour;
	j.opacity = o.opacity;
	j.world = o.world;
	j.specular = o.specular;
	j.shadow = o.shadow;
	j.gi = o.gi;
	return j;
}

bool same_jobs(const std::vector<mbx::FrameJob> &a, const std::vector<mbx::FrameJob> &b)
{
	return a.size() == b.size() &&
		(a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(mbx::FrameJob)) == 0);
}

bool same_frames(const std::vector<FrameConst> &a, const std::vector<FrameConst> &b)
{
	return a.size() == b.size() &&
		(a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(FrameConst)) == 0);
}

// The kernel variant a frame needs: the Monte Carlo / GI / shadow / AO / palette / specular / glow
// code (full), and the Monte Carlo lens or chromatic aberration besides (lens).
int frame_variant(const FrameConst &f)
{
	if (f.dofEnabled || f.chromaticAberration || f.fog) return mbx::kVariantLens;
	return (f.monteCarlo || f.aoEnabled || (f.lightEnabled && (f.castShadows || f.specPlastic || f.specMetal)) ||
			   (f.palette && f.paletteGradient) || f.glow || f.slowShading || f.perspective || f.aaSize)
		? mbx::kVariantFull
		: mbx::kVariantLite;
}

// Upload frames and jobs when they differ from what the device holds.
int stage_batch(mbx_ctx *ctx, const std::vector<FrameConst> &frames, const std::vector<mbx::FrameJob> &jobs)
{
	int rc = ensure_capacity(ctx, (int)frames.size());
	if (rc != MBX_OK) return rc;
	const bool fSame = same_frames(frames, ctx->framesHost);
	const bool jSame = same_jobs(jobs, ctx->jobsHost);
	if (fSame && jSame) return MBX_OK;
	hipError_t e;
	// an in-flight launch may still read the old batch
	if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "sync");
	if (!fSame &&
		(e = hipMemcpy(ctx->d_frames, frames.data(), frames.size() * sizeof(FrameConst), hipMemcpyHostToDevice)) !=
			hipSuccess)
		return hip_fail(ctx, e, "upload frames");
	if (!jSame &&
		(e = hipMemcpy(ctx->d_jobs, jobs.data(), jobs.size() * sizeof(mbx::FrameJob), hipMemcpyHostToDevice)) !=
			hipSuccess)
		return hip_fail(ctx, e, "upload jobs");
	ctx->framesHost = frames;
	ctx->jobsHost = jobs;
	return MBX_OK;
}

// Enqueue the batch launch (queue clear + persistent kernel), from a captured graph when enabled.
int enqueue_batch(mbx_ctx *ctx, int kind, int nFrames, unsigned int tiles)
{
	const int waves = ctx->waves > 0 ? ctx->waves : default_waves(ctx);
	const int mode = ctx->kernelMode == 2 ? mbx::kBatchPersistent : mbx::kBatchGrid;
	mbx::BatchArgs a{};
	a.frames = ctx->d_frames;
	a.jobs = ctx->d_jobs;
	a.numFrames = nFrames;
	a.queue = ctx->d_queue;
	a.tileShift = ctx->tileShift;
	if (mode == mbx::kBatchGrid && !ctx->useGraph)
	{
		// direct grid launch: accumulate into this bank (cleared by the previous grid launch, else
		// by the clear kernel) while workgroup 0 zeroes the other bank for the next launch
		a.counters = counter_bank(ctx, ctx->bank);
		a.clearFirst = ctx->bankClean ? 0 : 1;
		a.clearNext = counter_bank(ctx, ctx->bank ^ 1);
		ctx->statsBank = ctx->bank;
		ctx->bank ^= 1;
		ctx->bankClean = true;
	}
	else
	{
		// graphs bake the counter pointer in: bank 0, cleared by the clear kernel inside the graph
		if (ctx->useGraph) ctx->bank = 0;
		a.counters = counters_cleared_by_launch(ctx);
		a.clearFirst = 1;
		a.clearNext = nullptr;
	}
	a.full = ctx->batchFull;
	a.mode = mode;
	a.totalTiles = tiles;
	if (mode == mbx::kBatchGrid && ctx->batchSplit)
	{
		// the split render's buffers, grown to the batch
		const size_t need = ctx->batchPixels * mbx::kMaxHitRecordBytes;
		hipError_t e0;
		if (need > ctx->hitsBytes)
		{
			if (ctx->d_hits) (void)hipFree(ctx->d_hits);
			ctx->d_hits = nullptr;
			ctx->hitsBytes = 0;
			if ((e0 = hipMalloc((void **)&ctx->d_hits, need)) != hipSuccess) return hip_fail(ctx, e0, "allocate hit records");
			ctx->hitsBytes = need;
		}
		if (nFrames > ctx->hitCountsCap)
		{
			if (ctx->d_hitCounts) (void)hipFree(ctx->d_hitCounts);
			ctx->d_hitCounts = nullptr;
			ctx->hitCountsCap = 0;
			if ((e0 = hipMalloc((void **)&ctx->d_hitCounts, (size_t)nFrames * sizeof(unsigned int))) != hipSuccess)
				return hip_fail(ctx, e0, "allocate hit counts");
			ctx->hitCountsCap = nFrames;
		}
		a.hits = ctx->d_hits;
		a.hitCounts = ctx->d_hitCounts;
	}
	hipError_t e;
	if (!ctx->useGraph)
	{
		if ((e = mark(ctx, ctx->ev0)) != hipSuccess) return hip_fail(ctx, e, "event");
		if ((e = mbx::launch_batch(kind, a, waves, ctx->stream)) != hipSuccess)
			return hip_fail(ctx, e, "batch launch");
		return MBX_OK;
	}
	if (!ctx->graphExec || ctx->graphKind != kind || ctx->graphFrames != nFrames || ctx->graphWaves != waves ||
		ctx->graphTile != ctx->tileShift || ctx->graphFull != ctx->batchFull || ctx->graphMode != mode ||
		ctx->graphHits != a.hits || ctx->graphHitCounts != a.hitCounts ||
		ctx->graphTiles != tiles)
	{
		if (ctx->graphExec)
		{
			(void)hipGraphExecDestroy(ctx->graphExec);
			ctx->graphExec = nullptr;
		}
		hipStream_t cs = nullptr;
		if ((e = hipStreamCreate(&cs)) != hipSuccess) return hip_fail(ctx, e, "capture stream");
		hipGraph_t g = nullptr;
		int rc = MBX_OK;
		if ((e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal)) != hipSuccess)
			rc = hip_fail(ctx, e, "begin capture");
		if (rc == MBX_OK)
		{
			hipError_t e1 = mbx::launch_batch(kind, a, waves, cs);
			e = hipStreamEndCapture(cs, &g);
			if (e1 != hipSuccess) rc = hip_fail(ctx, e1, "capture batch");
			else if (e != hipSuccess) rc = hip_fail(ctx, e, "end capture");
		}
		if (rc == MBX_OK && (e = hipGraphInstantiate(&ctx->graphExec, g, nullptr, nullptr, 0)) != hipSuccess)
			rc = hip_fail(ctx, e, "instantiate graph");
		if (g) (void)hipGraphDestroy(g);
		(void)hipStreamDestroy(cs);
		if (rc != MBX_OK)
		{
			ctx->graphExec = nullptr;
			return rc;
		}
		ctx->graphKind = kind;
		ctx->graphFrames = nFrames;
		ctx->graphWaves = waves;
		ctx->graphTile = ctx->tileShift;
		ctx->graphFull = ctx->batchFull;
		ctx->graphHits = a.hits;
		ctx->graphHitCounts = a.hitCounts;
		ctx->graphMode = mode;
		ctx->graphTiles = tiles;
	}
	// timing starts after any (re)capture, which is host work
	if ((e = mark(ctx, ctx->ev0)) != hipSuccess) return hip_fail(ctx, e, "event");
	if ((e = hipGraphLaunch(ctx->graphExec, ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "graph launch");
	return MBX_OK;
}

int render_frames(mbx_ctx *ctx, const std::vector<FrameConst> &frames, const mbx_outputs *outs)
{
	hipError_t e = hipSetDevice(ctx->device);
	if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
	int kind = frames[0].deKind;
	uint64_t rays = 0;
	std::vector<mbx::FrameJob> jobs;
	int64_t tiles = 0;
	size_t pixels = 0;
	for (size_t i = 0; i < frames.size(); i++)
	{
		jobs.push_back(make_job(frames[i], outs[i], (int)tiles, (int)pixels, ctx->tileShift));
		tiles = jobs.back().tileEnd;
		pixels += (size_t)frames[i].width * frames[i].height;
		const uint64_t aaRays = frames[i].aaSize ? (uint64_t)frames[i].aaSize * frames[i].aaSize : 1;
		rays += (uint64_t)frames[i].width * frames[i].height * (uint64_t)frames[i].samples * aaRays;
		if (tiles > (int64_t)1 << 30) return fail(ctx, MBX_ERR_INVALID, "batch too large");
	}
	int rc = stage_batch(ctx, frames, jobs);
	if (rc != MBX_OK) return rc;
	ctx->batchFull = 0;
	bool monteCarlo = false;
	for (const auto &f : frames)
	{
		ctx->batchFull = std::max(ctx->batchFull, frame_variant(f));
		// several samples per pixel, or rays that are not the three-point camera's
		monteCarlo = monteCarlo || f.monteCarlo || f.aaSize || f.perspective;
	}
	// the split render takes the full variant without Monte Carlo sampling (one sample per pixel);
	// by default only the IFS kind (measured 1.3-1.4x on the Menger scene with the default
	// material; the Mandelbox kinds within 3%, the Mandelbulb and general kinds 5-18% slower)
	const bool splitKind = ctx->split == 2 || (ctx->split == 1 && kind == mbx::kIfsAnalytic);
	ctx->batchSplit = splitKind && ctx->batchFull == mbx::kVariantFull && !monteCarlo;
	ctx->batchPixels = pixels;
	if (ctx->kernelMode == 1)
	{
		// one plain launch per frame (the reference engine's job-per-frame shape)
		unsigned long long *counters = counters_cleared_by_launch(ctx);
		if ((e = mbx::launch_clear(ctx->d_queue, counters, ctx->stream)) != hipSuccess)
			return hip_fail(ctx, e, "clear counters");
		if ((e = mark(ctx, ctx->ev0)) != hipSuccess) return hip_fail(ctx, e, "event");
		for (size_t i = 0; i < frames.size(); i++)
		{
			const mbx::Outputs o{outs[i].rgb, outs[i].depth, outs[i].iters, outs[i].mask, outs[i].normal,
				outs[i].steps, outs[i].colour, outs[i].opacity, outs[i].world, outs[i].specular, outs[i].shadow,
				outs[i].gi};
			e = mbx::launch_render(kind, ctx->d_frames + i, frames[i].width, frames[i].height, o, counters,
				ctx->tileShift, frame_variant(frames[i]), ctx->stream);
			if (e != hipSuccess) return hip_fail(ctx, e, "render launch");
		}
	}
	else
	{
		rc = enqueue_batch(ctx, kind, (int)frames.size(), (unsigned int)tiles);
		if (rc != MBX_OK) return rc;
	}
	if ((e = mark(ctx, ctx->ev1)) != hipSuccess) return hip_fail(ctx, e, "event");
	ctx->launched = true;
	ctx->timedLast = ctx->timing != 0;
	ctx->rays = rays;
	ctx->statsCounters = nullptr;
	return MBX_OK;
}

// Derive the frames of a batch with one kernel specialisation per launch: Mandelbulb frames of
// mixed powers use the general-power (trigonometric) iteration, and all of them the FP64 one if
// any frame needs it; any other mix of kinds (a Mandelbox animation that turns its rotated folds
// on, a hybrid among single formulas) renders every frame on the general formula kernel.
int derive_batch(mbx_ctx *ctx, const mbx_params *params, FrameConst *frames, int n)
{
	for (int i = 0; i < n; i++)
	{
		int rc = derive_frame(ctx, &params[i], frames[i], false);
		if (rc != MBX_OK) return rc;
	}
	int kind = frames[0].deKind;
	bool generic = false;
	auto bulb = [](int k) { return k == mbx::kBulb8 || k == mbx::kBulbTrig || k == mbx::kBulbD; };
	for (int i = 1; i < n; i++)
	{
		const int k = frames[i].deKind;
		if (k == kind) continue;
		if (!bulb(k) || !bulb(kind))
		{
			generic = true;
			break;
		}
		kind = (k == mbx::kBulbD || kind == mbx::kBulbD) ? mbx::kBulbD : mbx::kBulbTrig;
	}
	if (generic)
	{
		for (int i = 0; i < n; i++)
		{
			int rc = derive_frame(ctx, &params[i], frames[i], true);
			if (rc != MBX_OK) return rc;
		}
		return MBX_OK;
	}
	for (int i = 0; i < n; i++) frames[i].deKind = kind;
	return MBX_OK;
}

// Constants of the differentiable render, or the reason the scene is not differentiable.
std::string derive_diff(const mbx_params *p, const FrameConst &fc, const double lightDir[3], mbx::DiffConst &dc_)
{
	const mbx_fractal &fr = p->fractal;
	if (fr.formula == MBX_FORMULA_MANDELBULB && fc.deKind == mbx::kBulbD)
		return "the differentiable render marches in FP32: this view needs the FP64 Mandelbulb (MBX_TUNE_PRECISION)";
	if (fr.formula != MBX_FORMULA_MANDELBULB || (fc.deKind != mbx::kBulb8 && fc.deKind != mbx::kBulbTrig) ||
		fc.analyticFunction != MBX_ANALYTIC_LOGARITHMIC)
		return "the differentiable render covers the Mandelbulb with the analytic logarithmic DE";
	if (!fc.addC || fc.julia) return "the differentiable render needs c = p (no Julia mode)";
	if (p->light.castShadows || p->ambientOcclusionEnabled || p->DOFMonteCarlo || p->antialiasingEnabled)
		return "the differentiable render has no shadows, AO, Monte Carlo sampling or antialiasing";
	if (p->perspectiveType != 0) return "the differentiable render has the three-point perspective only";
	if ((p->useColorsFromPalette && p->surfaceGradientEnable) || p->glowEnabled ||
		(p->light.enabled && (p->specularPlasticEnable || p->metallic)))
		return "the differentiable render has no palette colouring, specular highlights or glow";
	if (p->basicFogEnabled || p->volFogEnabled) return "the differentiable render has no fog";
	if (p->backgroundGamma != 1.0) return "the differentiable render needs background gamma 1";
	if (p->light.enabled && p->light.relativePosition)
		return "the differentiable render needs the light at an absolute angle (relativePosition 0)";
	mbx::DiffConst &c = dc_;
	std::memset(&c, 0, sizeof(c));
	c.power = fr.bulb.power;
	c.alpha = fr.bulb.alphaAngleOffset * M_PI / 180.0;
	c.beta = fr.bulb.betaAngleOffset * M_PI / 180.0;
	c.bailout = fc.bailoutD;
	for (int i = 0; i < 3; i++) c.cm[i] = fr.constantFactor[i];
	c.maxN = fc.formulaMaxiter;
	c.constThr = fc.constantDEThreshold;
	c.threshK = fc.threshKD;
	c.DEThresh = fc.DEThreshD;
	// node rotation and its derivative per degree
	for (int k = 0; k < 3; k++)
	{
		Dual1 a[3];
		for (int j = 0; j < 3; j++) a[j] = Dual1{p->formulaRotation[j] * (M_PI / 180.0), j == k ? M_PI / 180.0 : 0.0};
		DRot r;
		r.axis(2, a[2]);
		r.axis(1, a[1]);
		r.axis(0, a[0]);
		for (int i = 0; i < 9; i++) c.dR[k][i] = r.m[i].d;
	}
	for (int i = 0; i < 9; i++) c.R[i] = fc.nodeRotD[i];
	// camera rotation and its derivatives per camera / target component
	for (int k = 0; k < 6; k++)
	{
		Dual1 cam[3], tgt[3], m[9];
		for (int j = 0; j < 3; j++)
		{
			cam[j] = Dual1{p->camera[j], k == j ? 1.0 : 0.0};
			tgt[j] = Dual1{p->target[j], k == 3 + j ? 1.0 : 0.0};
		}
		camera_rotation_dual(cam, tgt, p->topVector, p->sweetSpotHAngle, p->sweetSpotVAngle, m);
		for (int i = 0; i < 9; i++) c.dM[k][i] = m[i].d;
	}
	for (int i = 0; i < 9; i++) c.M[i] = fc.mRotD[i];
	for (int i = 0; i < 3; i++) c.cam[i] = p->camera[i];
	c.fov = fc.fovD;
	c.aspect = fc.aspectRatioD;
	c.invW = fc.invWidthD;
	c.invH = fc.invHeightD;
	c.width = fc.width;
	c.height = fc.height;
	// shading and background
	c.lightEnabled = p->light.enabled;
	c.bg3 = p->background3ColorsEnable;
	for (int i = 0; i < 3; i++)
	{
		c.light[i] = lightDir[i];
		c.lightColor[i] = p->light.color[i];
		c.colour[i] = p->surfaceColor[i];
		c.fill[i] = p->fillLightColor[i];
		c.lum[i] = p->luminosity * p->luminosityColor[i];
		c.bg1[i] = p->backgroundColor1[i];
		c.bg2[i] = p->backgroundColor2[i];
		c.bgc3[i] = p->backgroundColor3[i];
	}
	c.intensity = p->light.intensity;
	c.shading = p->shading;
	c.bgBrightness = p->backgroundBrightness;
	c.lightSize = p->light.size;
	c.sunPow = 6.0 * p->light.contourSharpness;
	c.visibility = p->light.visibility;
	return std::string();
}

} // namespace

extern "C" {

int mbx_version(void) { return MBX_ABI_VERSION; }

// why the last mbx_ctx_create on this thread failed (mbx_last_error with a NULL context)
static thread_local std::string g_createError = "null context";

static int create_fail(mbx_ctx *ctx, int rc, const char *what, hipError_t e)
{
	g_createError = std::string("mbx_ctx_create: ") + what + ": " + hipGetErrorString(e);
	if (ctx) mbx_ctx_destroy(ctx);
	return rc;
}

int mbx_ctx_create(int device, void *stream, mbx_ctx **out)
{
	if (!out) return MBX_ERR_INVALID;
	*out = nullptr;
	mbx_ctx *ctx = new (std::nothrow) mbx_ctx();
	if (!ctx) return MBX_ERR_NOMEM;
	ctx->device = device;
	hipError_t e = hipSetDevice(device);
	if (e != hipSuccess)
	{
		delete ctx;
		return create_fail(nullptr, MBX_ERR_HIP, "hipSetDevice", e);
	}
	ctx->stream = (hipStream_t)stream;
	if ((e = hipMalloc((void **)&ctx->d_frame, sizeof(FrameConst))) != hipSuccess ||
		(e = hipMalloc((void **)&ctx->d_counters, 2 * mbx::kCounterBytes)) != hipSuccess ||
		(e = hipMalloc((void **)&ctx->d_queue, mbx::kTileQueueBytes)) != hipSuccess ||
		(e = hipMalloc((void **)&ctx->d_mcTables, kMcTableDoubles * sizeof(double))) != hipSuccess)
		return create_fail(ctx, MBX_ERR_HIP, "hipMalloc", e);
	{
		// the host's sincos and pow: a few thousand cycles once per context instead of the FP64
		// argument reductions' registers in every Monte Carlo kernel
		std::vector<double> t(kMcTableDoubles);
		for (int k = 0; k < mbx::kLensAngles; k++) ::sincos((double)k, &t[2 * k], &t[2 * k + 1]);
		for (int k = 0; k < mbx::kSoftRadii; k++)
			t[2 * mbx::kLensAngles + k] = std::pow((double)k / 10000.0, 1.0 / 3.0);
		if ((e = hipMemcpy(ctx->d_mcTables, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice)) !=
			hipSuccess)
			return create_fail(ctx, MBX_ERR_HIP, "hipMemcpy(tables)", e);
	}
	if ((e = hipEventCreate(&ctx->ev0)) != hipSuccess || (e = hipEventCreate(&ctx->ev1)) != hipSuccess ||
		(e = hipEventCreateWithFlags(&ctx->naCopied, hipEventDisableTiming)) != hipSuccess)
		return create_fail(ctx, MBX_ERR_HIP, "hipEventCreate", e);
	// the allocation-free batch path: this device's static scratch (its address is taken once per
	// device, with the device current) and host staging for its largest batch
	void *scratchAddr = nullptr;
	if ((e = mbx::noalloc_scratch_address(device, &scratchAddr)) != hipSuccess)
		return create_fail(ctx, MBX_ERR_HIP, "hipGetSymbolAddress(scratch)", e);
	ctx->d_scratch = (unsigned char *)scratchAddr;
	ctx->naFrames.resize(mbx::kNoAllocMaxFrames);
	ctx->naJobs.resize(mbx::kNoAllocMaxFrames);
	int cus = 0;
	if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
		ctx->numCUs = cus;
	*out = ctx;
	return MBX_OK;
}

void mbx_ctx_destroy(mbx_ctx *ctx)
{
	if (!ctx) return;
	// teardown is best effort: a failure here has no caller left to report to
	(void)hipSetDevice(ctx->device);
	if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
	else (void)hipDeviceSynchronize();
	if (ctx->d_frame) (void)hipFree(ctx->d_frame);
	if (ctx->d_counters) (void)hipFree(ctx->d_counters);
	if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
	if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
	if (ctx->naCopied) (void)hipEventDestroy(ctx->naCopied);
	if (ctx->graphExec) (void)hipGraphExecDestroy(ctx->graphExec);
	if (ctx->d_frames) (void)hipFree(ctx->d_frames);
	if (ctx->d_jobs) (void)hipFree(ctx->d_jobs);
	if (ctx->d_queue) (void)hipFree(ctx->d_queue);
	if (ctx->d_hits) (void)hipFree(ctx->d_hits);
	if (ctx->d_hitCounts) (void)hipFree(ctx->d_hitCounts);
	if (ctx->d_diff) (void)hipFree(ctx->d_diff);
	if (ctx->d_grad) (void)hipFree(ctx->d_grad);
	if (ctx->d_gamma) (void)hipFree(ctx->d_gamma);
	if (ctx->d_mcTables) (void)hipFree(ctx->d_mcTables);
	if (ctx->d_zrange) (void)hipFree(ctx->d_zrange);
	delete ctx;
}

const char *mbx_last_error(const mbx_ctx *ctx)
{
	if (!ctx) return g_createError.c_str();
	return ctx->err.c_str();
}

int mbx_scene_upload(mbx_ctx *ctx, const mbx_params *params)
{
	if (!ctx || !params) return fail(ctx, MBX_ERR_INVALID, "null argument");
	FrameConst fc;
	int rc = derive_frame(ctx, params, fc, false);
	if (rc != MBX_OK) return rc;
	hipError_t e = hipSetDevice(ctx->device);
	if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
	// the previous launch may still read the old constants
	if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "sync");
	if ((e = hipMemcpy(ctx->d_frame, &fc, sizeof(fc), hipMemcpyHostToDevice)) != hipSuccess)
		return hip_fail(ctx, e, "upload frame constants");
	ctx->frame = fc;
	ctx->sceneReady = true;
	ctx->diffWhy = derive_diff(params, fc, ctx->lightDir, ctx->diff);
	ctx->diffUploaded = false;
	ctx->err.clear();
	return MBX_OK;
}

}

namespace
{
// Validate the scene for the differentiable render and put its constants on the device.
int ensure_diff(mbx_ctx *ctx)
{
	if (!ctx->sceneReady) return fail(ctx, MBX_ERR_INVALID, "no scene uploaded");
	if (!ctx->diffWhy.empty()) return fail(ctx, MBX_ERR_UNSUPPORTED, ctx->diffWhy);
	hipError_t e = hipSetDevice(ctx->device);
	if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
	if (!ctx->d_diff && (e = hipMalloc((void **)&ctx->d_diff, sizeof(mbx::DiffConst))) != hipSuccess)
		return hip_fail(ctx, e, "allocate differentiable-render constants");
	if (!ctx->d_grad && (e = hipMalloc((void **)&ctx->d_grad, mbx::kDiffParams * sizeof(double))) != hipSuccess)
		return hip_fail(ctx, e, "allocate gradient");
	if (!ctx->diffUploaded)
	{
		if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "sync");
		if ((e = hipMemcpy(ctx->d_diff, &ctx->diff, sizeof(mbx::DiffConst), hipMemcpyHostToDevice)) != hipSuccess)
			return hip_fail(ctx, e, "upload differentiable-render constants");
		ctx->diffUploaded = true;
	}
	return MBX_OK;
}
} // namespace

extern "C" {

int mbx_render_diff(mbx_ctx *ctx, const mbx_outputs *out, mbx_ray_checkpoint *checkpoints)
{
	if (!ctx || !out || !checkpoints) return fail(ctx, MBX_ERR_INVALID, "null argument");
	int rc = ensure_diff(ctx);
	if (rc != MBX_OK) return rc;
	hipError_t e;
	if ((e = hipEventRecord(ctx->ev0, ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "event");
	if ((e = mbx::launch_clear(ctx->d_queue, counters_cleared_by_launch(ctx), ctx->stream)) != hipSuccess)
		return hip_fail(ctx, e, "clear");
	e = mbx::launch_render_diff(ctx->frame.deKind, ctx->d_frame, ctx->d_diff, ctx->frame.width,
		ctx->frame.height, out->rgb, out->depth, out->iters, out->mask, out->normal,
		reinterpret_cast<mbx::RayCheckpoint *>(checkpoints), ctx->stream);
	if (e != hipSuccess) return hip_fail(ctx, e, "differentiable render launch");
	if ((e = hipEventRecord(ctx->ev1, ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "event");
	ctx->launched = true;
	ctx->timedLast = true;
	ctx->rays = (uint64_t)ctx->frame.width * ctx->frame.height;
	return MBX_OK;
}

int mbx_adjoint(mbx_ctx *ctx, const mbx_ray_checkpoint *checkpoints, const float *dL_drgb, double *grad,
	double *pixel_grad)
{
	if (!ctx || !checkpoints || !dL_drgb || !grad) return fail(ctx, MBX_ERR_INVALID, "null argument");
	int rc = ensure_diff(ctx);
	if (rc != MBX_OK) return rc;
	hipError_t e;
	if ((e = hipEventRecord(ctx->ev0, ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "event");
	e = mbx::launch_adjoint(ctx->d_frame, ctx->d_diff, ctx->frame.width, ctx->frame.height,
		reinterpret_cast<const mbx::RayCheckpoint *>(checkpoints), dL_drgb, ctx->d_grad, pixel_grad, ctx->stream);
	if (e != hipSuccess) return hip_fail(ctx, e, "adjoint launch");
	if ((e = hipEventRecord(ctx->ev1, ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "event");
	if ((e = hipMemcpyAsync(grad, ctx->d_grad, mbx::kDiffParams * sizeof(double), hipMemcpyDeviceToHost,
			 ctx->stream)) != hipSuccess)
		return hip_fail(ctx, e, "read gradient");
	if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "adjoint");
	return MBX_OK;
}

int mbx_scene_info(const mbx_ctx *ctx, double *mrot9, double *fov, double *light_dir3, int *de_kind)
{
	if (!ctx || !ctx->sceneReady) return MBX_ERR_INVALID;
	if (mrot9) std::memcpy(mrot9, ctx->mRot, 9 * sizeof(double));
	if (fov) *fov = ctx->fov;
	if (light_dir3) std::memcpy(light_dir3, ctx->lightDir, 3 * sizeof(double));
	if (de_kind) *de_kind = ctx->frame.deKind;
	return MBX_OK;
}

int mbx_render(mbx_ctx *ctx, const mbx_outputs *out)
{
	if (!ctx || !out) return fail(ctx, MBX_ERR_INVALID, "null argument");
	if (!ctx->sceneReady) return fail(ctx, MBX_ERR_INVALID, "no scene uploaded");
	std::vector<FrameConst> frames(1, ctx->frame);
	return render_frames(ctx, frames, out);
}

int mbx_render_batch(mbx_ctx *ctx, const mbx_params *params, const mbx_outputs *outs, int n)
{
	if (!ctx || !params || !outs || n < 1) return fail(ctx, MBX_ERR_INVALID, "bad argument");
	std::vector<FrameConst> frames((size_t)n);
	int rc = derive_batch(ctx, params, frames.data(), n);
	if (rc != MBX_OK) return rc;
	return render_frames(ctx, frames, outs);
}

int mbx_render_batch_noalloc(mbx_ctx *ctx, const mbx_params *params, const mbx_outputs *outs, int n)
{
	if (!ctx || !params || !outs) return fail(ctx, MBX_ERR_INVALID, "null argument");
	if (n < 1 || n > mbx::kNoAllocMaxFrames)
		return fail(ctx, MBX_ERR_INVALID,
			"mbx_render_batch_noalloc: " + std::to_string(n) + " frames; the static scratch holds 1 to " +
				std::to_string(mbx::kNoAllocMaxFrames));
	if (!ctx->d_scratch) return fail(ctx, MBX_ERR_HIP, "no static scratch on this device");
	hipError_t e = hipSetDevice(ctx->device);
	if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
	// the host staging is overwritten below: the previous call's copy out of it must have ended
	if (ctx->naCopyPending && (e = hipEventSynchronize(ctx->naCopied)) != hipSuccess)
		return hip_fail(ctx, e, "staging copy");
	ctx->naCopyPending = false;
	FrameConst *frames = ctx->naFrames.data();
	int rc = derive_batch(ctx, params, frames, n);
	if (rc != MBX_OK) return rc;
	const int kind = frames[0].deKind;
	uint64_t rays = 0;
	int64_t tiles = 0;
	size_t pixels = 0;
	int full = 0;
	for (int i = 0; i < n; i++)
	{
		ctx->naJobs[i] = make_job(frames[i], outs[i], (int)tiles, (int)pixels, ctx->tileShift);
		tiles = ctx->naJobs[i].tileEnd;
		pixels += (size_t)frames[i].width * frames[i].height;
		const uint64_t aaRays = frames[i].aaSize ? (uint64_t)frames[i].aaSize * frames[i].aaSize : 1;
		rays += (uint64_t)frames[i].width * frames[i].height * (uint64_t)frames[i].samples * aaRays;
		full = std::max(full, frame_variant(frames[i]));
		if (tiles > (int64_t)1 << 30) return fail(ctx, MBX_ERR_INVALID, "batch too large");
	}
	unsigned char *base = ctx->d_scratch;
	auto *dCounters = (unsigned long long *)(base + mbx::kNoAllocCountersOff);
	auto *dQueue = (unsigned int *)(base + mbx::kNoAllocQueueOff);
	auto *dJobs = (mbx::FrameJob *)(base + mbx::kNoAllocJobsOff);
	auto *dFrames = (FrameConst *)(base + mbx::kNoAllocFramesOff);
	// copies on the context stream, ordered after any launch still reading the scratch
	if ((e = hipMemcpyAsync(dJobs, ctx->naJobs.data(), (size_t)n * sizeof(mbx::FrameJob), hipMemcpyHostToDevice,
			 ctx->stream)) != hipSuccess ||
		(e = hipMemcpyAsync(dFrames, frames, (size_t)n * sizeof(FrameConst), hipMemcpyHostToDevice, ctx->stream)) !=
			hipSuccess)
		return hip_fail(ctx, e, "stage frames");
	if ((e = hipEventRecord(ctx->naCopied, ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "event");
	ctx->naCopyPending = true;
	if ((e = mark(ctx, ctx->ev0)) != hipSuccess) return hip_fail(ctx, e, "event");
	if (ctx->kernelMode == 1)
	{
		if ((e = mbx::launch_clear(dQueue, dCounters, ctx->stream)) != hipSuccess)
			return hip_fail(ctx, e, "clear counters");
		for (int i = 0; i < n; i++)
		{
			const mbx::Outputs o{outs[i].rgb, outs[i].depth, outs[i].iters, outs[i].mask, outs[i].normal,
				outs[i].steps, outs[i].colour, outs[i].opacity, outs[i].world, outs[i].specular, outs[i].shadow,
				outs[i].gi};
			e = mbx::launch_render(kind, dFrames + i, frames[i].width, frames[i].height, o, dCounters,
				ctx->tileShift, frame_variant(frames[i]), ctx->stream);
			if (e != hipSuccess) return hip_fail(ctx, e, "render launch");
		}
	}
	else
	{
		mbx::BatchArgs a{};
		a.frames = dFrames;
		a.jobs = dJobs;
		a.numFrames = n;
		a.queue = dQueue;
		a.tileShift = ctx->tileShift;
		a.counters = dCounters;
		a.clearFirst = 1; // a clear kernel, not a memset: the call may sit in a captured stream
		a.clearNext = nullptr;
		a.full = full;
		a.mode = ctx->kernelMode == 2 ? mbx::kBatchPersistent : mbx::kBatchGrid;
		a.totalTiles = (unsigned int)tiles;
		const int waves = ctx->waves > 0 ? ctx->waves : default_waves(ctx);
		if ((e = mbx::launch_batch(kind, a, waves, ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "batch launch");
	}
	if ((e = mark(ctx, ctx->ev1)) != hipSuccess) return hip_fail(ctx, e, "event");
	ctx->launched = true;
	ctx->timedLast = ctx->timing != 0;
	ctx->rays = rays;
	ctx->statsCounters = dCounters;
	return MBX_OK;
}

int mbx_tune(mbx_ctx *ctx, int key, int value)
{
	if (!ctx) return MBX_ERR_INVALID;
	switch (key)
	{
		case MBX_TUNE_KERNEL:
			if (value < 0 || value > 2) return fail(ctx, MBX_ERR_INVALID, "kernel mode must be 0, 1 or 2");
			ctx->kernelMode = value;
			return MBX_OK;
		case MBX_TUNE_WAVES:
			if (value < 0 || value > (1 << 22)) return fail(ctx, MBX_ERR_INVALID, "wave count out of range");
			ctx->waves = value;
			return MBX_OK;
		case MBX_TUNE_GRAPH:
			ctx->useGraph = value ? 1 : 0;
			return MBX_OK;
		case MBX_TUNE_DELTA_DE:
			ctx->deltaTangent = value ? 1 : 0;
			return MBX_OK;
		case MBX_TUNE_FAR_FIELD:
			ctx->farField = value ? 1 : 0;
			return MBX_OK;
		case MBX_TUNE_PRECISION:
			if (value < 0 || value > 2) return fail(ctx, MBX_ERR_INVALID, "precision must be 0, 1 or 2");
			ctx->precision = value;
			return MBX_OK;
		case MBX_TUNE_TIMING:
			ctx->timing = value ? 1 : 0;
			return MBX_OK;
		case MBX_TUNE_SPLIT:
			if (value < 0 || value > 2) return fail(ctx, MBX_ERR_INVALID, "split must be 0, 1 or 2");
			ctx->split = value;
			return MBX_OK;
		case MBX_TUNE_TILE_WIDTH:
		{
			int shift = -1;
			for (int k = 2; k <= 5; k++)
				if (value == (1 << k)) shift = k;
			if (shift < 0) return fail(ctx, MBX_ERR_INVALID, "tile width must be 4, 8, 16 or 32");
			ctx->tileShift = shift;
			return MBX_OK;
		}
		default: return fail(ctx, MBX_ERR_INVALID, "unknown tuning key");
	}
}

int mbx_get_tune(mbx_ctx *ctx, int key, int *value)
{
	if (!ctx || !value) return fail(ctx, MBX_ERR_INVALID, "null argument");
	switch (key)
	{
		case MBX_TUNE_KERNEL: *value = ctx->kernelMode; return MBX_OK;
		case MBX_TUNE_WAVES: *value = ctx->waves > 0 ? ctx->waves : default_waves(ctx); return MBX_OK;
		case MBX_TUNE_GRAPH: *value = ctx->useGraph; return MBX_OK;
		case MBX_TUNE_DELTA_DE: *value = ctx->deltaTangent; return MBX_OK;
		case MBX_TUNE_TILE_WIDTH: *value = 1 << ctx->tileShift; return MBX_OK;
		case MBX_TUNE_FAR_FIELD: *value = ctx->farField; return MBX_OK;
		case MBX_TUNE_PRECISION: *value = ctx->precision; return MBX_OK;
		case MBX_TUNE_TIMING: *value = ctx->timing; return MBX_OK;
		case MBX_TUNE_SPLIT: *value = ctx->split; return MBX_OK;
		default: return fail(ctx, MBX_ERR_INVALID, "unknown tuning key");
	}
}

int mbx_autotune(mbx_ctx *ctx, const mbx_outputs *out, int reps, int *best_waves, double *best_ms)
{
	if (!ctx || !out || reps < 1) return fail(ctx, MBX_ERR_INVALID, "bad argument");
	if (!ctx->sceneReady) return fail(ctx, MBX_ERR_INVALID, "no scene uploaded");
	// the tuner times its renders with the context's events whatever MBX_TUNE_TIMING says
	struct TimingOn
	{
		mbx_ctx *c;
		int saved;
		~TimingOn() { c->timing = saved; }
	} timingOn{ctx, ctx->timing};
	ctx->timing = 1;
	const int savedMode = ctx->kernelMode;
	// the wave count matters only to the persistent kernel
	const std::vector<int> perCU = ctx->kernelMode == 2 ? std::vector<int>{4, 8, 12, 16, 20, 24, 32}
														: std::vector<int>{0};
	const int shifts[] = {3, 4, 2, 5}; // tile widths 8, 16, 4, 32 (64 pixels each)
	double bestT = 1e30;
	int bestW = ctx->waves, bestShift = ctx->tileShift;
	for (int sh : shifts)
	for (int k : perCU)
	{
		ctx->waves = ctx->numCUs * k;
		ctx->tileShift = sh;
		double total = 0.0;
		for (int r = 0; r < reps + 1; r++)
		{
			int rc = mbx_render(ctx, out);
			if (rc != MBX_OK)
			{
				ctx->kernelMode = savedMode;
				return rc;
			}
			mbx_stats st;
			rc = mbx_get_stats(ctx, &st);
			if (rc != MBX_OK)
			{
				ctx->kernelMode = savedMode;
				return rc;
			}
			if (r > 0) total += st.kernel_ms; // first run of a setting: graph capture, warm-up
		}
		const double t = total / reps;
		if (t < bestT)
		{
			bestT = t;
			bestW = ctx->waves;
			bestShift = sh;
		}
	}
	ctx->waves = bestW;
	ctx->tileShift = bestShift;
	ctx->kernelMode = savedMode;
	if (best_waves) *best_waves = bestW > 0 ? bestW : default_waves(ctx);
	if (best_ms) *best_ms = bestT;
	return MBX_OK;
}

int mbx_get_stats(mbx_ctx *ctx, mbx_stats *stats)
{
	if (!ctx || !stats) return fail(ctx, MBX_ERR_INVALID, "null argument");
	if (!ctx->launched) return fail(ctx, MBX_ERR_INVALID, "nothing rendered yet");
	hipError_t e;
	float ms = 0.0f;
	if (ctx->timedLast)
	{
		if ((e = hipEventSynchronize(ctx->ev1)) != hipSuccess) return hip_fail(ctx, e, "render");
		if ((e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1)) != hipSuccess) return hip_fail(ctx, e, "timing");
	}
	else if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
		return hip_fail(ctx, e, "render");
	std::vector<unsigned long long> slots((size_t)mbx::kCounterSlots * mbx::kCounterStride);
	const unsigned long long *src = ctx->statsCounters ? ctx->statsCounters : counter_bank(ctx, ctx->statsBank);
	if ((e = hipMemcpy(slots.data(), src, mbx::kCounterBytes, hipMemcpyDeviceToHost)) != hipSuccess)
		return hip_fail(ctx, e, "read counters");
	unsigned long long c[mbx::kCounters] = {0, 0, 0};
	for (unsigned int sl = 0; sl < mbx::kCounterSlots; sl++)
		for (unsigned int k = 0; k < mbx::kCounters; k++) c[k] += slots[(size_t)sl * mbx::kCounterStride + k];
	stats->rays = ctx->rays;
	stats->de_evals = c[0];
	stats->iterations = c[1];
	stats->kernel_ms = ms;
	stats->far_steps = c[2];
	return MBX_OK;
}

int mbx_prepare_ao_vectors(mbx_params *p, const float *lightmap, int width, int height)
{
	if (!p || (lightmap && (width <= 0 || height <= 0)) || p->ambientOcclusionQuality < 1) return MBX_ERR_INVALID;
	static const float kWhite[5 * 5 * 3] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
		1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
		1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
	const float *tex = lightmap ? lightmap : kWhite;
	const int W = lightmap ? width : 5, H = lightmap ? height : 5;
	// mRotAmbientOcclusionLightMapRotation: SetRotation of (angle, 0, 0) in radians, a turn about z
	Rot rot;
	rot.z(p->aoLightMapRotation / 180.0 * M_PI);
	const double q = p->ambientOcclusionQuality;
	int count = 0;
	// rows of constant latitude b, each sampled in longitude a with a step widened by 1 / cos(b);
	// the lightmap pixel of (a, b) sets the colour and directions with a black pixel are dropped
	// sin and cos of b, and of a + b, each from one sincos call, as the reference's build folds
	// them (glibc's sincos may differ from sin / cos in the last bit)
	for (double b = -0.49 * M_PI; b < 0.49 * M_PI; b += 1.0 / q)
	{
		double sb, cb;
		::sincos(b, &sb, &cb);
		for (double a = 0.0; a < 2.0 * M_PI; a += ((2.0 / q) / cb))
		{
			double sab, cab;
			::sincos(a + b, &sab, &cab);
			const D3 d = rot.apply(D3{cab * cb, sab * cb, sb});
			const int X = int((a + b) / (2.0 * M_PI) * W + W * 8.5) % W;
			const int Y = int(b / (M_PI)*H + H * 8.5) % H;
			const float *c = tex + 3 * ((size_t)X + (size_t)Y * W);
			if (c[0] > 0.001 || c[1] > 0.001 || c[2] > 0.001)
			{
				if (count >= MBX_MAX_AO_VECTORS) return MBX_ERR_UNSUPPORTED;
				p->aoVectors[count][0] = d.x;
				p->aoVectors[count][1] = d.y;
				p->aoVectors[count][2] = d.z;
				for (int k = 0; k < 3; k++) p->aoColors[count][k] = c[k];
				count++;
			}
		}
	}
	if (count == 0)
	{
		// no lit direction: one null direction with no colour
		for (int k = 0; k < 3; k++) p->aoVectors[0][k] = p->aoColors[0][k] = 0.0;
		count = 1;
	}
	p->aoVectorsCount = count;
	return MBX_OK;
}

int mbx_formula_count(void) { return kFormulaCount; }

int mbx_formula_get(int index, mbx_formula_info *info)
{
	if (!info || index < 0 || index >= kFormulaCount) return MBX_ERR_INVALID;
	const FormulaMeta &m = kFormulaTable[index];
	std::memset(info, 0, sizeof(*info));
	info->internalID = m.id;
	info->DEType = m.deType;
	info->DEFunctionType = m.deFunction;
	info->cpixelAddition = m.cpixel;
	info->defaultBailout = m.defaultBailout;
	info->DEAnalyticFunction = m.analytic;
	info->coloringFunction = m.coloring;
	std::snprintf(info->internalName, sizeof(info->internalName), "%s", m.internalName);
	std::snprintf(info->nameInComboBox, sizeof(info->nameInComboBox), "%s", m.nameInComboBox);
	return MBX_OK;
}

int mbx_formula_iterate(mbx_ctx *ctx, const mbx_fractal *fractal, int n, const double *z_in,
	const mbx_extended_aux *aux_in, double *z_out, mbx_extended_aux *aux_out)
{
	if (!ctx || !fractal || n < 0 || (n > 0 && (!z_in || !aux_in || !z_out || !aux_out)))
		return fail(ctx, MBX_ERR_INVALID, "bad argument");
	if (!formula_meta(fractal->formula)) return fail(ctx, MBX_ERR_UNSUPPORTED, "formula not in the table");
	if (n == 0) return MBX_OK;
	hipError_t e = hipSetDevice(ctx->device);
	if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
	mbx::FractalParams slot;
	gen_slot(*fractal, slot);
	const size_t zBytes = 4 * sizeof(double) * (size_t)n, auxBytes = sizeof(mbx_extended_aux) * (size_t)n;
	const size_t slotBytes = (sizeof(slot) + 255) / 256 * 256, zOff = slotBytes, auxOff = zOff + (zBytes + 255) / 256 * 256;
	unsigned char *d = nullptr;
	int rc = MBX_OK;
	if ((e = hipMalloc((void **)&d, auxOff + auxBytes)) != hipSuccess) return hip_fail(ctx, e, "allocate");
	double *dz = (double *)(d + zOff);
	mbx_extended_aux *daux = (mbx_extended_aux *)(d + auxOff);
	if ((e = hipMemcpyAsync(d, &slot, sizeof(slot), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
		(e = hipMemcpyAsync(dz, z_in, zBytes, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
		(e = hipMemcpyAsync(daux, aux_in, auxBytes, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
		rc = hip_fail(ctx, e, "copy inputs");
	if (rc == MBX_OK && (e = mbx::launch_formula((const mbx::FractalParams *)d, dz, daux, n, ctx->stream)) != hipSuccess)
		rc = hip_fail(ctx, e, "formula launch");
	if (rc == MBX_OK && (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) rc = hip_fail(ctx, e, "formula");
	if (rc == MBX_OK && ((e = hipMemcpy(z_out, dz, zBytes, hipMemcpyDeviceToHost)) != hipSuccess ||
							(e = hipMemcpy(aux_out, daux, auxBytes, hipMemcpyDeviceToHost)) != hipSuccess))
		rc = hip_fail(ctx, e, "copy outputs");
	(void)hipFree(d);
	return rc;
}

int mbx_distance(mbx_ctx *ctx, const double *points, int n, double detail_size, float *dist_out,
	int32_t *iters_out)
{
	if (!ctx || (!points && n > 0) || n < 0) return fail(ctx, MBX_ERR_INVALID, "bad argument");
	if (!ctx->sceneReady) return fail(ctx, MBX_ERR_INVALID, "no scene uploaded");
	if (n == 0) return MBX_OK;
	hipError_t e = hipSetDevice(ctx->device);
	if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
	std::vector<float> pf(3 * (size_t)n);
	for (size_t i = 0; i < pf.size(); i++) pf[i] = (float)points[i];
	float *d_pts = nullptr, *d_dist = nullptr;
	int32_t *d_it = nullptr;
	int rc = MBX_OK;
	if ((e = hipMalloc((void **)&d_pts, pf.size() * sizeof(float))) != hipSuccess ||
		(e = hipMalloc((void **)&d_dist, (size_t)n * sizeof(float))) != hipSuccess ||
		(e = hipMalloc((void **)&d_it, (size_t)n * sizeof(int32_t))) != hipSuccess)
		rc = hip_fail(ctx, e, "allocate");
	if (rc == MBX_OK &&
		(e = hipMemcpyAsync(d_pts, pf.data(), pf.size() * sizeof(float), hipMemcpyHostToDevice,
			 ctx->stream)) != hipSuccess)
		rc = hip_fail(ctx, e, "copy points");
	if (rc == MBX_OK &&
		(e = mbx::launch_distance(ctx->frame.deKind, ctx->d_frame, d_pts, n, (float)detail_size, d_dist,
			 d_it, ctx->stream)) != hipSuccess)
		rc = hip_fail(ctx, e, "distance launch");
	if (rc == MBX_OK && (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) rc = hip_fail(ctx, e, "distance");
	if (rc == MBX_OK && dist_out &&
		(e = hipMemcpy(dist_out, d_dist, (size_t)n * sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess)
		rc = hip_fail(ctx, e, "copy distances");
	if (rc == MBX_OK && iters_out &&
		(e = hipMemcpy(iters_out, d_it, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost)) != hipSuccess)
		rc = hip_fail(ctx, e, "copy iterations");
	if (d_pts) (void)hipFree(d_pts);
	if (d_dist) (void)hipFree(d_dist);
	if (d_it) (void)hipFree(d_it);
	return rc;
}

int mbx_compile_image(mbx_ctx *ctx, const mbx_image_adjustments *adj, const float *rgb, int32_t width,
	int32_t height, uint16_t *rgb16, uint8_t *rgb8)
{
	if (!ctx || !adj || !rgb || width <= 0 || height <= 0 || (!rgb16 && !rgb8))
		return fail(ctx, MBX_ERR_INVALID, "bad argument");
	if ((uint64_t)width * (uint64_t)height > 0xffffffffull) return fail(ctx, MBX_ERR_INVALID, "image too large");
	if (!(adj->gamma >= 0.0f)) return fail(ctx, MBX_ERR_INVALID, "gamma must be >= 0");
	hipError_t e = hipSetDevice(ctx->device);
	if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
	if (!ctx->d_gamma && (e = hipMalloc((void **)&ctx->d_gamma, 65536 * sizeof(uint16_t))) != hipSuccess)
		return hip_fail(ctx, e, "allocate gamma table");
	if (!ctx->gammaReady || ctx->gammaOf != adj->gamma)
	{
		// cImage::CalculateGammaTable (cimage.cpp:363-376): int(powf(i / 65536, 1 / gamma) * 65535)
		// for every 16-bit value, in float as the reference computes it
		if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "sync");
		ctx->gammaHost.resize(65536);
		const float invGamma = 1.0f / adj->gamma;
		for (uint32_t i = 0; i < 65536; i++)
			ctx->gammaHost[i] = (uint16_t)(int)(powf((float)i / 65536.0f, invGamma) * 65535.0f);
		if ((e = hipMemcpyAsync(ctx->d_gamma, ctx->gammaHost.data(), 65536 * sizeof(uint16_t),
				 hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
			return hip_fail(ctx, e, "copy gamma table");
		ctx->gammaOf = adj->gamma;
		ctx->gammaReady = true;
	}
	if ((e = mbx::launch_compile_image(rgb, (uint32_t)((uint64_t)width * (uint64_t)height), ctx->d_gamma,
			 adj->brightness, adj->contrast, adj->saturation, adj->hdr ? 1 : 0, rgb16, rgb8, ctx->stream)) !=
		hipSuccess)
		return hip_fail(ctx, e, "compile image launch");
	return MBX_OK;
}

int mbx_compile_zbuffer(mbx_ctx *ctx, const float *depth, int32_t width, int32_t height, int32_t bits,
	int32_t logarithmic, int32_t invert, int32_t constant_range, double min_depth, double max_depth, void *out)
{
	if (!ctx || !depth || !out || width <= 0 || height <= 0 || (bits != 8 && bits != 16))
		return fail(ctx, MBX_ERR_INVALID, "bad argument");
	if ((uint64_t)width * (uint64_t)height > 0xffffffffull) return fail(ctx, MBX_ERR_INVALID, "image too large");
	hipError_t e = hipSetDevice(ctx->device);
	if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
	if (!ctx->d_zrange && (e = hipMalloc((void **)&ctx->d_zrange, 2 * sizeof(uint32_t))) != hipSuccess)
		return hip_fail(ctx, e, "allocate z range");
	if ((e = mbx::launch_zbuffer(depth, (uint32_t)((uint64_t)width * (uint64_t)height), ctx->d_zrange,
			 constant_range ? 0 : 1, (float)min_depth, (float)max_depth, bits, logarithmic ? 1 : 0, invert ? 1 : 0, out,
			 ctx->stream)) != hipSuccess)
		return hip_fail(ctx, e, "z-buffer launch");
	return MBX_OK;
}

int mbx_render_ssao(mbx_ctx *ctx, const mbx_params *params, const float *depth, const uint8_t *colour,
	const uint16_t *opacity, const float *rgb_in, float *rgb_out)
{
	if (!ctx || !params || !depth || !colour || !rgb_in || !rgb_out) return fail(ctx, MBX_ERR_INVALID, "bad argument");
	const int W = params->imageWidth, H = params->imageHeight;
	if (W <= 0 || H <= 0 || (uint64_t)W * (uint64_t)H > 0xffffffffull)
		return fail(ctx, MBX_ERR_INVALID, "bad image size");
	if (params->ambientOcclusionQuality < 1 || params->ambientOcclusionQuality > 10)
		return fail(ctx, MBX_ERR_INVALID, "ambientOcclusionQuality must be 1 .. 10");
	if (params->perspectiveType < 0 || params->perspectiveType > 3)
		return fail(ctx, MBX_ERR_INVALID, "perspectiveType must be 0 .. 3");
	mbx::SsaoArgs a{};
	a.depth = depth;
	a.colour = colour;
	a.opacity = opacity;
	a.rgbIn = rgb_in;
	a.rgbOut = rgb_out;
	a.width = W;
	a.height = H;
	// render_ssao.cpp:126-136 with no progressive pass: quality = int(q^2 * 1.0), at least 3
	int quality = (int)(params->ambientOcclusionQuality * params->ambientOcclusionQuality * 1.0);
	if (quality < 3) quality = 3;
	a.quality = quality;
	a.perspective = params->perspectiveType;
	a.randomMode = params->SSAORandomMode ? 1 : 0;
	a.seed = params->randomSeed ^ 0x55A0A055u;
	// CalcFOV (projection_3d.cpp:140-159)
	a.fov = params->perspectiveType == 0 ? 2.0 * std::tan(params->fov / 360.0 * M_PI) : params->fov / 180.0 * M_PI;
	a.aspectRatio = params->perspectiveType == 2 ? 2.0 : (double)W / H;
	a.scaleFactor = (double)W / (quality * quality) / 2.0;
	a.intensity = (float)params->ambientOcclusion;
	// ambientOcclusionColor as toRGBFloat gives it (a float per channel)
	for (int k = 0; k < 3; k++) a.aoColor[k] = (float)params->ambientOcclusionColor[k];
	// the worker's angle table (ssao_worker.cpp:75-81), with the host's libm
	for (int i = 0; i < quality; i++)
	{
		a.sinT[i] = std::sin(double(i) / quality * 2.0 * M_PI);
		a.cosT[i] = std::cos(double(i) / quality * 2.0 * M_PI);
	}
	hipError_t e = hipSetDevice(ctx->device);
	if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
	if ((e = mbx::launch_ssao(a, ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "SSAO launch");
	return MBX_OK;
}

size_t mbx_sizeof_params(void) { return sizeof(mbx_params); }

} // extern "C"

// mbx_pixel.h - per-pixel device path shared by the render and differentiable kernels:
// RayMarching with its binary search, normals, shading (light, shadow, AO, background, GI),
// the pixel sample loop and the output writes.
//
// One lane per primary ray.  A 64-lane wave covers a 64-pixel tile so that neighbouring rays,
// which take similar march paths, share a wave.  Mandelbulb and IFS run in FP32, the Mandelbox
// kinds march and iterate in FP64; the per-frame constants are wave-uniform scalar loads.
//
// Path (reference): cRenderWorker::doWork pixel body (render_worker.cpp:155-525) ->
// RayRecursion with reflections off (render_worker.cpp:969-1819) -> RayMarching
// (render_worker.cpp:708-951) -> CalculateDistance -> Compute -> FormulaCode; on a hit
// CalculateNormals (shader_calculate_normals.cpp:40-130), ObjectShader (shader_object.cpp:40-229)
// with LightShading / AuxShadow / AO, and the GI bounce loop (shader_global_illumination.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "mbx_device.h"
#include "mbx_launch.h"

namespace mbx
{


template <int KIND>
using PrecOf = Prec<PosPrecision<KIND>::kFP64>;

template <int KIND>
struct MarchOut
{
	typename PrecOf<KIND>::P point;
	typename PrecOf<KIND>::R depth;
	typename PrecOf<KIND>::R distThresh;
	typename PrecOf<KIND>::R lastDist;
	int found;
	int iters;
	int steps;
	int stepCount; // buffCount: march iterations that ended without a hit (render_worker.cpp:838)
};

// Advanced quality: the threshold clamp of CalcDistThresh (render_worker.cpp:671-675) and the step
// limits of RayMarching (render_worker.cpp:819-834; without advanced quality only the clamp at 3)
__device__ __forceinline__ float clamp_thresh(const FrameConst &f, float t, bool adv)
{
	if (!adv) return t;
	t = t > f.detailSizeMax ? f.detailSizeMax : t;
	return t < f.detailSizeMin ? f.detailSizeMin : t;
}
__device__ __forceinline__ double clamp_thresh(const FrameConst &f, double t, bool adv)
{
	if (!adv) return t;
	t = t > f.detailSizeMaxD ? f.detailSizeMaxD : t;
	return t < f.detailSizeMinD ? f.detailSizeMinD : t;
}
__device__ __forceinline__ float limit_step(const FrameConst &f, float step, float thr, bool adv)
{
#pragma clang fp contract(off)
	if (!adv) return step > 3.0f ? 3.0f : step;
	if (step > f.absMaxStep) step = f.absMaxStep;
	if (step < f.absMinStep) step = f.absMinStep;
	if (thr > f.absMinStep && step > f.relMaxStep * thr) step = f.relMaxStep * thr;
	if (step < f.relMinStep * thr) step = f.relMinStep * thr;
	return step;
}
__device__ __forceinline__ double limit_step(const FrameConst &f, double step, double thr, bool adv)
{
#pragma clang fp contract(off)
	if (!adv) return step > 3.0 ? 3.0 : step;
	if (step > f.absMaxStepD) step = f.absMaxStepD;
	if (step < f.absMinStepD) step = f.absMinStepD;
	if (thr > f.absMinStepD && step > f.relMaxStepD * thr) step = f.relMaxStepD * thr;
	if (step < f.relMinStepD * thr) step = f.relMinStepD * thr;
	return step;
}

// CalcDistThresh (render_worker.cpp:653-679)
template <int KIND>
__device__ __forceinline__ typename PrecOf<KIND>::R dist_thresh(const FrameConst &f, typename PrecOf<KIND>::P p)
{
	using Q = PrecOf<KIND>;
	const bool adv = f.advancedQuality != 0;
	if (f.constantDEThreshold) return clamp_thresh(f, Q::DEThresh(f), adv);
	return clamp_thresh(f, Q::thresh(f, p), adv);
}

// CalcDelta (render_worker.cpp:684-690)
template <int KIND>
__device__ __forceinline__ typename PrecOf<KIND>::R calc_delta(const FrameConst &f, typename PrecOf<KIND>::P p)
{
	using Q = PrecOf<KIND>;
	return Q::len(Q::cam(f) - p) * Q::deltaK(f);
}

// Fog of VolumetricShader (shader_volumetric.cpp:70-700) as the march goes.  The reference stores
// every march step and walks them back from the last to the second, blending per step the glow
// (:196-204), the distance fog (:556-571) and the basic fog (:575-590) into the shaded colour.
// Each step's blend is affine in the colour, so the walk equals the composition of those affine
// maps, which the march accumulates front to back (nearest step first) without a step buffer:
// colour = A * shaded + gc * G + B with A the product of the steps' transmittances, gc the
// weight of the glow colour G (known only once the step count is) and B the fogs' own
// contributions; the fog opacity (opacityOut) is 1 - T with T the product of the fogs'
// transmittances.  The reference's merge of steps below 1e-10 of the threshold never happens:
// the host requires DEFactor >= 1e-9 (each step is at least 0.2 thresholds * DEFactor).
struct FogAcc
{
	float A, T, gc;
	V3 B;
};

__device__ __forceinline__ void fog_init(FogAcc &acc)
{
	acc.A = 1.0f;
	acc.T = 1.0f;
	acc.gc = 0.0f;
	acc.B = v3(0.0f, 0.0f, 0.0f);
}

// One step of the walk: `step` is the march step that reached the point, `distance` the DE there
// (the step buffer's entry, render_worker.cpp:788-838).
__device__ __forceinline__ void fog_step(const FrameConst &f, FogAcc &acc, double step, double distance)
{
#pragma clang fp contract(off)
	const float og = f.glow ? f.glowStepOpacity : 0.0f;
	float od = 0.0f;
	V3 cd = v3(0.0f, 0.0f, 0.0f);
	if (f.volFog && f.volFogDensity > 0.0f)
	{
		// DistanceFogOpacity (shader_distance_fog_opacity.cpp:37-52)
		const double factor = f.volFogFactorD;
		const double shifted = fabs(distance - f.volFogFromSurfaceD) + 0.1 * f.volFogFromSurfaceD;
		const float densityTemp = (float)(step * factor / (shifted * shifted + factor * factor));
		double o = 0.3f * f.volFogDensity * densityTemp / (1.0f + f.volFogDensity * densityTemp) - 0.001;
		o = o < 0.0 ? 0.0 : (o > 1.0 ? 1.0 : o);
		od = (float)o;
		// its colour from the shifted distance (shader_volumetric.cpp:344-355)
		float k = (float)(shifted / f.volFogColour1DistD);
		if (k > 1.0f) k = 1.0f;
		float kn = 1.0f - k;
		const V3 t = v3(f.volFogColour[0][0] * kn + f.volFogColour[1][0] * k,
			f.volFogColour[0][1] * kn + f.volFogColour[1][1] * k, f.volFogColour[0][2] * kn + f.volFogColour[1][2] * k);
		float k2 = (float)(shifted / f.volFogColour2DistD * k);
		if (k2 > 1.0f) k2 = 1.0f;
		kn = 1.0f - k2;
		cd = v3(t.x * kn + f.volFogColour[2][0] * k2, t.y * kn + f.volFogColour[2][1] * k2,
			t.z * kn + f.volFogColour[2][2] * k2);
	}
	float ob = 0.0f;
	if (f.basicFog)
	{
		ob = (float)(step / f.basicFogVisibilityD);
		if (ob > 1.0f) ob = 1.0f;
	}
	// the step's map: x -> (1 - ob) [(1 - od) ((1 - og) x + og G) + od cd] + ob cb
	const float kb = 1.0f - ob, kd = 1.0f - od;
	acc.gc = acc.gc + acc.A * (kb * kd * og);
	const float wd = acc.A * (kb * od), wb = acc.A * ob;
	acc.B = v3(acc.B.x + wd * cd.x + wb * f.basicFogColor[0], acc.B.y + wd * cd.y + wb * f.basicFogColor[1],
		acc.B.z + wd * cd.z + wb * f.basicFogColor[2]);
	acc.A = acc.A * (kb * kd * (1.0f - og));
	acc.T = acc.T * (kb * kd);
}

// RayMarching (render_worker.cpp:708-951), exterior mode, binary search on.  In the FP64
// position precision the arithmetic follows the reference's operation order without FMA
// contraction, so that march decisions at the DE's discontinuities come out the same.
// ADV: 0 = advanced quality off, 1 = on (both compiled into the FP32 kinds, so the default march
// carries no test of the flag), 2 = read from the frame (FP64 kinds: one copy of the march)
// FOG compiles the fog walk in: with fogOn the steps are accumulated into *fog (non-null).
template <int KIND, int ADV, bool FOG = false>
__device__ void ray_marching(const FrameConst &f, typename PrecOf<KIND>::P start, typename PrecOf<KIND>::P dir,
	typename PrecOf<KIND>::R maxScan, uint32_t &seed, MarchOut<KIND> &out, Work &w, bool fromCamera = true,
	FogAcc *fog = nullptr, bool fogOn = false)
{
#pragma clang fp contract(off)
	using Q = PrecOf<KIND>;
	using R = typename Q::R;
	using P = typename Q::P;
	constexpr bool kExact = PosPrecision<KIND>::kFP64;
	P point = Q::from(v3(1e30f, 1e30f, 1e30f));
	bool found = false, dead = false;
	R scan = 0, dist = 0, step = 0, thr = 0;
	int counter = 0, iters = 0, stepCount = 0;
	const R dirLength = Q::len(dir);
	const R invDirLength = Q::rcp(dirLength);
	const R searchLimit = kExact ? (R)(1.0 - 0.001 * f.detailLevelD) : (R)f.searchLimit;
	// FP32: the ray starts at the camera, so |camera - point| = scan |dir|; one multiply, and
	// closer to the FP64 value than the difference of two rounded positions
	const bool constThr = f.constantDEThreshold != 0;
	const R thrScale = kExact ? (R)0 : dirLength * (R)f.threshK;
	const R constThrValue = Q::DEThresh(f);
	const int maxSteps = f.maxRaymarchingSteps;
	const bool adv = ADV == 2 ? f.advancedQuality != 0 : ADV == 1; // advanced-quality threshold and step limits
	// far field (FP32 Mandelbulb kinds): at |p|^2 >= farR2 the host has bounded the DE from the
	// first formula iteration (|z_1| >= |p|^P - |c|, escape at i = 0, iters = 1) above the value
	// that makes the step below clamp to exactly 3, so the step is taken without evaluating it;
	// the jitter's random number is still drawn, so every later value is unchanged
	const float farR2 = (kExact || (FOG && fogOn)) ? __builtin_inff() : f.farR2; // fog needs every DE value
	for (int i = 0; i < maxSteps; i++)
	{
		const P last = point;
		counter++;
		point = Q::along(start, dir, scan);
		// a step below the position resolution leaves the point where it was (dead point).  In
		// FP32 the point is always finite: every step is in [0.45 thr DEFactor, 3].
		bool deadPoint = point.x == last.x && point.y == last.y && point.z == last.z;
		if (kExact) deadPoint = deadPoint || !(isfinite(point.x) && isfinite(point.y) && isfinite(point.z));
		if (deadPoint)
		{
			point = last;
			found = true;
			dead = true;
			break;
		}
		if (!kExact && (float)dot(Q::tof(point), Q::tof(point)) >= farR2)
		{
			(void)random_mod(seed, 1000);
			step = (R)3;
			iters = 1;
			w.far_steps++;
			stepCount = i + 1;
			scan += step * invDirLength;
			if (scan > maxScan) break;
			continue;
		}
		if (kExact || !fromCamera)
			thr = dist_thresh<KIND>(f, point);
		else
			thr = clamp_thresh(f, constThr ? constThrValue : scan * thrScale, adv);
		dist = calculate_distance<KIND>(point, thr, f, iters, w);
		if (dist < thr)
		{
			found = true;
			break;
		}
		// the step buffer's entry i (step that reached the point, DE there); the walk takes 1 ..
		if constexpr (FOG)
			if (fogOn && i > 0) fog_step(f, *fog, (double)step, (double)dist);
		const int rnd = random_mod(seed, 1000);
		if (kExact)
		{
			// interior mode keeps a wider margin to the surface (render_worker.cpp:809-812)
			const R margin = (KIND == kGeneric && f.interiorMode) ? (R)0.8 : (R)0.5;
			step = (dist - margin * thr) * (R)f.DEFactorD * ((R)1 - (R)rnd / (R)10000);
			step = limit_step(f, step, thr, adv);
			scan += step / dirLength;
		}
		else
		{
			step = (dist - (R)0.5 * thr) * (R)f.DEFactor * ((R)1 - (R)rnd * (R)1e-4);
			step = limit_step(f, step, thr, adv);
			scan += step * invDirLength;
		}
		stepCount = i + 1;
		if (scan > maxScan) break;
	}
	point = Q::along(start, dir, scan);
	if (found && !dead)
	{
		step *= (R)0.5;
		for (int i = 0; i < 30; i++)
		{
			counter++;
			if (dist < thr && dist > thr * searchLimit) break;
			if (dist > thr)
				scan += step;
			else if (dist < thr * searchLimit)
				scan -= step;
			point = Q::along(start, dir, scan);
			if (kExact || !fromCamera)
				thr = dist_thresh<KIND>(f, point);
			else
				thr = clamp_thresh(f, constThr ? constThrValue : scan * thrScale, adv);
			dist = calculate_distance<KIND>(point, thr, f, iters, w);
			step *= (R)0.5;
		}
	}
	if constexpr (KIND == kGeneric)
	{
		// iteration-threshold mode backs off by one threshold (render_worker.cpp:913-933)
		if (f.iterThreshMode)
		{
			scan -= thr;
			point = Q::along(start, dir, scan);
			if (found) dist = calculate_distance<KIND>(point, thr, f, iters, w);
		}
	}
	out.point = point;
	out.depth = scan;
	out.distThresh = thr;
	out.lastDist = dist;
	out.found = found;
	out.iters = iters;
	out.steps = counter;
	out.stepCount = stepCount;
}

// CalculateNormals, fast mode: central differences of the DE (shader_calculate_normals.cpp:44-78),
// normalised in the kind's precision
template <int KIND>
__device__ typename PrecOf<KIND>::P calculate_normals(const FrameConst &frame, typename PrecOf<KIND>::P p,
	typename PrecOf<KIND>::R thr, Work &w)
{
#pragma clang fp contract(off)
	const FrameConst &f = uniform_ref(frame);
	using Q = PrecOf<KIND>;
	using R = typename Q::R;
	using P = typename Q::P;
	// delta = distThresh * smoothness, a fifth of it in interior mode (shader_calculate_normals.cpp:46-47)
	const R delta = (KIND == kGeneric && f.interiorMode) ? thr * (R)0.2 * Q::smoothness(f) : thr * Q::smoothness(f);
	int it;
	P e = p;
	e.x = p.x + delta;
	const R sx1 = calculate_distance<KIND>(e, thr, f, it, w, true);
	e.x = p.x - delta;
	const R sx2 = calculate_distance<KIND>(e, thr, f, it, w, true);
	e = p;
	e.y = p.y + delta;
	const R sy1 = calculate_distance<KIND>(e, thr, f, it, w, true);
	e.y = p.y - delta;
	const R sy2 = calculate_distance<KIND>(e, thr, f, it, w, true);
	e = p;
	e.z = p.z + delta;
	const R sz1 = calculate_distance<KIND>(e, thr, f, it, w, true);
	e.z = p.z - delta;
	const R sz2 = calculate_distance<KIND>(e, thr, f, it, w, true);
	const R nx = sx1 - sx2, ny = sy1 - sy2, nz = sz1 - sz2;
	P n;
	if (nx == (R)0 && ny == (R)0 && nz == (R)0)
		n = Q::from(v3(1.0f, 0.0f, 0.0f));
	else
	{
		const R k = (R)1 / Q::len(P{nx, ny, nz});
		n = P{nx * k, ny * k, nz * k};
	}
	if (!finite3(Q::tof(n))) n = Q::from(v3(1.0f, 0.0f, 0.0f));
	return n;
}

// CalculateNormals with slow_shading (shader_calculate_normals.cpp:81-123): the sum of
// q * (1 + N - iterations) over the 11 x 11 x 11 grid q in {-1, -0.8, .., 1}^3 (the loop counters
// accumulate 0.2 in FP64 as the reference's do) of orbits at p + q delta, normalised, with the
// reference's fallback (1, 0, 0).  Out of line: 1331 orbits per hit, only for scenes asking for it.
template <int KIND>
__device__ __attribute__((noinline)) typename PrecOf<KIND>::P slow_normals(const FrameConst &frame,
	typename PrecOf<KIND>::P p, double delta, Work &w)
{
#pragma clang fp contract(off)
	const FrameConst &f = uniform_ref(frame);
	using Q = PrecOf<KIND>;
	using R = typename Q::R;
	using P = typename Q::P;
	const double d = delta * f.smoothnessD;
	double nx = 0.0, ny = 0.0, nz = 0.0;
	for (double qx = -1.0; qx <= 1.0; qx += 0.2)
		for (double qy = -1.0; qy <= 1.0; qy += 0.2)
			for (double qz = -1.0; qz <= 1.0; qz += 0.2)
			{
				P q = p;
				q.x = p.x + (R)(qx * d);
				q.y = p.y + (R)(qy * d);
				q.z = p.z + (R)(qz * d);
				const double pseudo = (double)(1 + f.slowShadingN - orbit_iterations<KIND>(q, f, w));
				nx = nx + qx * pseudo;
				ny = ny + qy * pseudo;
				nz = nz + qz * pseudo;
			}
	if (fabs(nx) < 1e-50 && fabs(ny) < 1e-50 && fabs(nz) < 1e-50) return Q::from(v3(1.0f, 0.0f, 0.0f));
	const double inv = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
	const P n = Q::fromD(nx * inv, ny * inv, nz * inv);
	return finite3(Q::tof(n)) ? n : Q::from(v3(1.0f, 0.0f, 0.0f));
}

// AuxShadow for the directional light (shader_aux_shadow.cpp:39-323): hard, or the soft
// cone variant when soft_shadow_cone > 0
template <int KIND>
__device__ float aux_shadow(const FrameConst &f, typename PrecOf<KIND>::P p, typename PrecOf<KIND>::R thr,
	typename PrecOf<KIND>::R distance, uint32_t &seed, Work &w)
{
#pragma clang fp contract(off)
	using Q = PrecOf<KIND>;
	using R = typename Q::R;
	using P = typename Q::P;
	P lv = Q::lightDir(f);
	if (f.mcSoftShadows)
	{
		// a random point of the ball of radius softRange (uniform in volume through the cube root
		// of the radius) added to the light vector (shader_aux_shadow.cpp:78-87)
		P rv;
		rv.x = (R)((double)random_mod(seed, 10000) / 5000.0 - 1.0);
		rv.y = (R)((double)random_mod(seed, 10000) / 5000.0 - 1.0);
		rv.z = (R)((double)random_mod(seed, 10000) / 5000.0 - 1.0);
		const R radius = (R)f.softCbrt[random_mod(seed, kSoftRadii - 1)];
		const R k = Q::softRange(f) * radius / Q::len(rv);
		lv = Q::normalize_p(P{lv.x + rv.x * k, lv.y + rv.y * k, lv.z + rv.z * k});
	}
	R shadowTemp = 1;
	R maxSoft = 0;
	R step = thr;
	int count = 0;
	int it;
	// interior mode starts one DE-factor threshold out (shader_aux_shadow.cpp:53-54)
	const R start = (KIND == kGeneric && f.interiorMode) ? thr * Q::DEFactor(f) : thr;
	// the loop compares with the threshold as a float (shader_aux_shadow.cpp:98-105); it
	// matters where the DE returns the threshold itself (interior mode)
	const R thrF = (R)(float)thr;
	for (R i = start; i < distance; i += step)
	{
		const P p2 = p + lv * i;
		const R dist = calculate_distance<KIND>(p2, thr, f, it, w);
		if (f.softShadows)
		{
			R angle = (dist - thrF) / i;
			if (angle < (R)0 || dist < thrF) angle = 0;
			R soft = (R)1 - angle / (R)f.softRange;
			if (f.penetrating) soft *= (distance - i) / distance;
			soft = soft < (R)0 ? (R)0 : soft;
			maxSoft = soft > maxSoft ? soft : maxSoft;
		}
		shadowTemp = 1;
		if (dist < thrF)
		{
			if (f.penetrating)
			{
				shadowTemp = shadowTemp - (distance - i) / distance;
				shadowTemp = shadowTemp < (R)0 ? (R)0 : shadowTemp;
			}
			else
				shadowTemp = 0;
			break;
		}
		step = (dist < (R)1e6 ? dist : (R)1e6) * Q::DEFactor(f);
		step = step > (R)1e-15 ? step : (R)1e-15;
		count++;
		if (count > f.maxRaymarchingSteps) break;
	}
	return f.softShadows ? (float)((R)1 - maxSoft) : (float)shadowTemp;
}

// AmbientOcclusion, multiple rays (shader_ambient_occlusion.cpp:41-122)
template <int KIND>
__device__ V3 ambient_occlusion(const FrameConst &f, typename PrecOf<KIND>::P p, typename PrecOf<KIND>::R thr,
	typename PrecOf<KIND>::R delta, uint32_t &seed, Work &w)
{
	using Q = PrecOf<KIND>;
	using R = typename Q::R;
	using P = typename Q::P;
	V3 ao = v3(0.0f, 0.0f, 0.0f);
	const R end_dist = delta * (R)f.height; // delta / resolution
	int start = 0, end = f.aoCount - 1;
	if (f.monteCarlo)
	{
		const int pick = random_mod(seed, f.aoCount - 1);
		start = pick;
		end = pick;
	}
	int it;
	for (int i = start; i <= end; i++)
	{
		const P v = Q::aoVector(f, i);
		R shadowTemp = 1;
		R dist;
		for (R r = thr; r < end_dist; r += dist * (R)2)
		{
			dist = calculate_distance<KIND>(p + v * r, thr, f, it, w);
			if (dist < thr)
			{
				shadowTemp = shadowTemp - (end_dist - r) / end_dist;
				shadowTemp = shadowTemp < (R)0 ? (R)0 : shadowTemp;
				break;
			}
		}
		ao = ao + v3(f.aoColors[i][0], f.aoColors[i][1], f.aoColors[i][2]) * (float)shadowTemp;
	}
	if (!f.monteCarlo) ao = ao * frcp((float)f.aoCount);
	return ao;
}

// FastAmbientOcclusion (shader_fast_ambient_occlusion.cpp:40-66)
template <int KIND>
__device__ float fast_ambient_occlusion(const FrameConst &f, typename PrecOf<KIND>::P p,
	typename PrecOf<KIND>::P n, typename PrecOf<KIND>::R thr, Work &w)
{
	using Q = PrecOf<KIND>;
	using R = typename Q::R;
	float aoTemp = 0.0f;
	R lastDist = (R)1e20;
	const int q2 = f.aoQuality * f.aoQuality;
	int it;
	for (int i = 1; i < q2; i++)
	{
		const R scan = (R)(i * i) * thr;
		R dist = calculate_distance<KIND>(p + n * scan, thr, f, it, w);
		if (dist > lastDist * (R)2) dist = lastDist * (R)2;
		lastDist = dist;
		aoTemp += (float)((R)exp2f(-(float)i) * (scan - (R)f.aoFastTune * dist) / thr);
	}
	return fmaxf(1.0f - 0.2f * aoTemp, 0.0f);
}

// BackgroundShader: 3-colour gradient and the directional light's disk (shader_background.cpp)
__device__ V3 background_shader(const FrameConst &frame, V3 viewVector)
{
	const FrameConst &f = uniform_ref(frame);
	const V3 vn = normalize(viewVector);
	float c[3];
	if (f.background3Colors)
	{
		const float grad = vn.z + 1.0f;
#pragma unroll
		for (int k = 0; k < 3; k++)
		{
			float v;
			if (grad < 1.0f)
				v = f.bg3[k] * (1.0f - grad) + f.bg2[k] * grad;
			else
				v = f.bg2[k] * (2.0f - grad) + f.bg1[k] * (grad - 1.0f);
			c[k] = v * f.bgBrightness;
		}
	}
	else
	{
#pragma unroll
		for (int k = 0; k < 3; k++) c[k] = f.bg1[k] * f.bgBrightness;
	}
	V3 out;
	if (f.bgInvGamma == 1.0f)
		out = v3(c[0], c[1], c[2]);
	else
		out = v3(powf(c[0], f.bgInvGamma), powf(c[1], f.bgInvGamma), powf(c[2], f.bgInvGamma));
	if (f.lightEnabled)
	{
		const V3 ld = v3(f.lightDir[0], f.lightDir[1], f.lightDir[2]);
		const float a = -(dot(vn, ld) - 1.0f) * 360.0f * frcp(f.lightSize);
		const float pw = 6.0f * f.lightContourSharpness;
		const float ap = (pw == 6.0f) ? (a * a) * (a * a) * (a * a) : powf(a, pw);
		const float intensity = frcp(1.0f + ap) * f.lightVisibility * f.lightIntensity;
		out = out + v3(f.lightColor[0], f.lightColor[1], f.lightColor[2]) * intensity;
	}
	return out;
}

// SurfaceColour of the fractal (shader_surface_color.cpp:39-108): the gradient colour at the
// palette position of the fractal-local point's colour index, or the material colour
template <int KIND>
__device__ V3 surface_colour(const FrameConst &frame, typename PrecOf<KIND>::P p, const Work &w)
{
#pragma clang fp contract(off)
	const FrameConst &f = uniform_ref(frame);
	using Q = PrecOf<KIND>;
	if (!f.palette || !f.paletteGradient) return v3(f.surfaceColor[0], f.surfaceColor[1], f.surfaceColor[2]);
	const typename Q::P local = w.nodeRotated ? Q::node_local(f, p) : p;
	const double colorIndex = colour_index<KIND>(local, f, w);
	const double nrCol = fmod(fabs(colorIndex), 248.0 * 256.0);
	const double position = fmod(nrCol / 256.0 / 10.0 * f.coloringSpeedD + f.paletteOffsetD, 1.0);
	return gradient_colour(f, (float)position);
}

// SpecularHighlight (shader_specular_highlight.cpp:38-73) with diffuse 1 and no texture: the
// Blinn term pow(n.h, 30 / width), the metallic roughness factor (one draw of the pixel's
// random stream), clamped at 15
__device__ __forceinline__ float specular_shade(V3 n, V3 view, V3 lv, float expo, float roughness, uint32_t &seed)
{
#pragma clang fp contract(off)
	const V3 half = normalize(lv - view);
	float s = dot(n, half);
	if (s < 0.0f) s = 0.0f;
	s = powf(s, expo);
	if (roughness > 0.0f)
		s = (float)((double)s * (1.0 + (double)((float)random_mod(seed, 1000) / 1000.0f * roughness)));
	return s > 15.0f ? 15.0f : s;
}

// ObjectShader with one directional light, Lambert, optional shadow and AO, and the combined
// plastic + metallic specular highlight (shader_object.cpp:40-229, shader_light_shading.cpp:40-122,
// shader_specular_highlight_combined.cpp:39-68).  colour is the surface colour, view the ray
// direction; the highlight (already shadowed) goes to spec.  FULL = false compiles the shadow,
// AO and specular paths out (frames that use none of them), which lowers the kernel's registers.
// shadowOut (optional) receives light 1's shadow factor as the reference's outShadow holds it:
// the AuxShadow value, 0 when neither the shade nor the highlight passes 0.001, 1 without cast
// shadows, 0 with the light off (shader_light_shading.cpp:94-119, render_worker.cpp:1410-1411).
template <int KIND, bool FULL = true>
__device__ V3 object_shader(const FrameConst &f, typename PrecOf<KIND>::P p, typename PrecOf<KIND>::P nP,
	typename PrecOf<KIND>::R thr, typename PrecOf<KIND>::R delta, V3 view, V3 colour, V3 &spec, uint32_t &seed,
	Work &w, float *shadowOut = nullptr)
{
	using R = typename PrecOf<KIND>::R;
	const V3 n = PrecOf<KIND>::tof(nP);
	spec = v3(0.0f, 0.0f, 0.0f);
	V3 ambient = v3(0.0f, 0.0f, 0.0f);
	if (FULL && f.aoEnabled)
	{
		if (f.aoMode == MBX_AO_FAST)
		{
			const float a = fast_ambient_occlusion<KIND>(f, p, nP, thr, w);
			ambient = v3(a, a, a);
		}
		else
			ambient = ambient_occlusion<KIND>(f, p, thr, delta, seed, w);
	}
	V3 light = v3(0.0f, 0.0f, 0.0f);
	if (f.lightEnabled)
	{
		const V3 lv = v3(f.lightDir[0], f.lightDir[1], f.lightDir[2]);
		float shade = fmaxf(dot(n, lv), 0.0f);
		shade = 1.0f - f.shading + shade * f.shading;
		shade = fminf(shade * f.lightIntensity, 500.0f);
		V3 specular = v3(0.0f, 0.0f, 0.0f);
		if (FULL && (f.specPlastic || f.specMetal))
		{
#pragma clang fp contract(off)
			const V3 sc = v3(f.specColor[0], f.specColor[1], f.specColor[2]);
			if (f.specPlastic)
			{
				const float s = specular_shade(n, view, lv, f.specPlasticExp, 0.0f, seed);
				specular = v3(s * sc.x * f.specPlasticK, s * sc.y * f.specPlasticK, s * sc.z * f.specPlasticK);
			}
			if (f.specMetal)
			{
				const float s = specular_shade(n, view, lv, f.specMetalExp, f.specRoughness, seed);
				specular = specular + v3(s * sc.x * (f.specMetalK * colour.x), s * sc.y * (f.specMetalK * colour.y),
										  s * sc.z * (f.specMetalK * colour.z));
			}
			specular = specular * f.lightIntensity;
		}
		float shadow = 1.0f;
		if (FULL && f.castShadows)
		{
			const float specMax = fmaxf(specular.x, fmaxf(specular.y, specular.z));
			if (shade > 0.001f || specMax > 0.001f)
			{
				const R distance = f.penetrating ? delta * (R)f.height : (R)f.viewDistanceMaxLight;
				shadow = aux_shadow<KIND>(f, p, thr, distance, seed, w);
				specular = specular * shadow;
			}
			else
			{
				shadow = 0.0f;
				specular = v3(0.0f, 0.0f, 0.0f);
			}
		}
		light = v3(f.lightColor[0], f.lightColor[1], f.lightColor[2]) * (shade * shadow);
		if (FULL) spec = v3(specular.x * f.lightColor[0], specular.y * f.lightColor[1], specular.z * f.lightColor[2]);
		if (shadowOut) *shadowOut = shadow;
	}
	else if (shadowOut)
		*shadowOut = 0.0f;
	V3 out;
	out.x = (f.fillLight[0] + ambient.x * f.aoColor[0]) * colour.x + light.x * colour.x + f.luminosity[0];
	out.y = (f.fillLight[1] + ambient.y * f.aoColor[1]) * colour.y + light.y * colour.y + f.luminosity[1];
	out.z = (f.fillLight[2] + ambient.z * f.aoColor[2]) * colour.z + light.z * colour.z + f.luminosity[2];
	return out;
}

// GlobalIlumination (shader_global_illumination.cpp:41-209).  colour0 is the primary hit's
// surface colour; each bounce is tinted by the colour of the surface it left.
template <int KIND>
__device__ V3 global_illumination(const FrameConst &frame, typename PrecOf<KIND>::P p0, typename PrecOf<KIND>::P n0,
	typename PrecOf<KIND>::R thr0, V3 colour0, uint32_t &seed, Work &w)
{
	const FrameConst &f = uniform_ref(frame);
	using Q = PrecOf<KIND>;
	using R = typename Q::R;
	using P = typename Q::P;
	V3 out = v3(0.0f, 0.0f, 0.0f);
	P p = p0;
	P n = n0;
	R thrCur = thr0;
	V3 colourTemp = colour0;
	V3 newColour = colour0;
	for (int depth = 0; depth < f.reflectionsMax; depth++)
	{
		P dir;
		if constexpr (PosPrecision<KIND>::kFP64)
		{
			// the reference's FP64 bounce direction (shader_global_illumination.cpp:63-71)
#pragma clang fp contract(off)
			const double rx = (double)(random_mod(seed, 20000) - 10000) / 10000.0;
			const double ry = (double)(random_mod(seed, 20000) - 10000) / 10000.0;
			const double rz = (double)(random_mod(seed, 20000) - 10000) / 10000.0;
			const P v = P{n.x + rx * 1.2, n.y + ry * 1.2, n.z + rz * 1.2};
			const double norm = 1.0 / Q::len(v);
			dir = P{v.x * norm, v.y * norm, v.z * norm};
		}
		else
		{
			const float rx = (float)(random_mod(seed, 20000) - 10000) * 1e-4f;
			const float ry = (float)(random_mod(seed, 20000) - 10000) * 1e-4f;
			const float rz = (float)(random_mod(seed, 20000) - 10000) * 1e-4f;
			dir = Q::from(normalize(n + v3(rx * 1.2f, ry * 1.2f, rz * 1.2f)));
		}
		const V3 dirf = Q::tof(dir);
		R dist = 0;
		bool found = false;
		int stepCount = 0;
		int it;
		P hit = p;
		for (R scan = thrCur; scan < (R)f.viewDistanceMax; scan += dist * Q::DEFactor(f))
		{
			if (stepCount >= f.maxRaymarchingSteps) break;
			const P q = p + dir * scan;
			const R thr = dist_thresh<KIND>(f, q);
			thrCur = thr;
			dist = calculate_distance<KIND>(q, thr, f, it, w);
			if (dist < thr)
			{
				if (scan < thr * (R)2) return out;
				hit = q;
				found = true;
				break;
			}
			stepCount++;
		}
		V3 result;
		bool finished = false;
		if (found)
		{
			p = hit;
			// the bounce's shader input has delta = distThresh (shader_global_illumination.cpp:86)
			n = f.slowShading ? slow_normals<KIND>(f, p, (KIND == kGeneric && f.interiorMode) ? (double)thrCur * 0.2
																								 : (double)thrCur, w)
							  : calculate_normals<KIND>(f, p, thrCur, w);
			const V3 colour = surface_colour<KIND>(f, p, w);
			V3 spec;
			result = object_shader<KIND>(f, p, n, thrCur, thrCur, dirf, colour, spec, seed, w);
			newColour = colour;
			out = out + v3((result.x + spec.x + f.lumEmissive[0]) * colourTemp.x,
							(result.y + spec.y + f.lumEmissive[1]) * colourTemp.y,
							(result.z + spec.z + f.lumEmissive[2]) * colourTemp.z);
			result = result + spec;
		}
		else
		{
			result = background_shader(f, dirf);
			finished = true;
		}
		const float lim = f.giRadianceLimit;
		result = v3(fminf(fmaxf(result.x, 0.0f), lim), fminf(fmaxf(result.y, 0.0f), lim),
			fminf(fmaxf(result.z, 0.0f), lim));
		out = out + v3(result.x * colourTemp.x, result.y * colourTemp.y, result.z * colourTemp.z);
		colourTemp = newColour;
		if (finished) break;
	}
	return out;
}

// The glow of VolumetricShader (shader_volumetric.cpp:86-92, 196-204) with the other volumetric
// effects off: stepCount - 1 identical blends towards the glow colour.  The reference skips a
// blend only when a march step is below 1e-10 of the distance threshold, which needs a DE factor
// below 1e-9 (the host rejects glow there).
__device__ __forceinline__ V3 glow_colour(const FrameConst &f, int stepCount)
{
#pragma clang fp contract(off)
	const float glow = (float)stepCount * f.glowIntensity / 512.0f * f.DEFactor;
	float glowN = 1.0f - glow;
	if (glowN < 0.0f) glowN = 0.0f;
	return v3(f.glowColor1[0] * glowN + f.glowColor2[0] * glow, f.glowColor1[1] * glowN + f.glowColor2[1] * glow,
		f.glowColor1[2] * glowN + f.glowColor2[2] * glow);
}
__device__ __forceinline__ V3 glow_shader(const FrameConst &f, int stepCount, V3 rgb)
{
#pragma clang fp contract(off)
	if (stepCount <= 0) return rgb;
	const float glow = (float)stepCount * f.glowIntensity / 512.0f * f.DEFactor;
	const V3 g = glow_colour(f, stepCount);
	float opacity = glow / (float)stepCount;
	if (opacity > 1.0f) opacity = 1.0f;
	const float keep = 1.0f - opacity;
	for (int k = 1; k < stepCount; k++)
		rgb = v3(opacity * g.x + keep * rgb.x, opacity * g.y + keep * rgb.y, opacity * g.z + keep * rgb.z);
	return rgb;
}

struct PixelOut
{
	V3 rgb;
	V3 normal;
	float depth;
	int found;
	int iters;
	int steps;
	uint32_t colour[3]; // the colour buffer's bytes (integer mean over the samples)
	uint32_t opacity;   // the opacity buffer's ushort (integer mean over the samples)
};

// uchar(c * 255) of the reference's colour buffer (render_worker.cpp:402-404): a float product
// truncated to an integer whose low byte is kept
__device__ __forceinline__ uint32_t colour_byte(float c)
{
#pragma clang fp contract(off)
	return (uint32_t)(uint8_t)(int)(c * 255.0f);
}

// MonteCarloDOFNoiseEstimation (render_worker.cpp:1871-1901) after sample `rep` with the running
// sum: channels in float as sRGBFloat holds them.  mc_deviation accumulates every sample;
// mc_noise, the estimate in double, is only needed once the pixel may stop.
__device__ __forceinline__ void mc_deviation(V3 pixel, int rep, V3 sum, V3 &stdDevSum)
{
#pragma clang fp contract(off)
	const float k = (float)(rep + 1);
	const float aR = sum.x / k, aG = sum.y / k, aB = sum.z / k;
	stdDevSum.x += (pixel.x - aR) * (pixel.x - aR);
	stdDevSum.y += (pixel.y - aG) * (pixel.y - aG);
	stdDevSum.z += (pixel.z - aB) * (pixel.z - aB);
}
__device__ __attribute__((noinline)) double mc_noise(int rep, V3 sum, V3 stdDevSum)
{
#pragma clang fp contract(off)
	double total = sqrt((double)(stdDevSum.x + stdDevSum.y + stdDevSum.z) / ((double)rep + 1.0));
	total /= sqrt((double)(rep + 1));
	const double brightness = (double)(sum.x + sum.y + sum.z) / ((double)rep + 1.0);
	if (brightness > 1.0) total /= brightness;
	return total;
}

// Hsv2rgb (hsv2rgb.cpp:40-60) in float: hue in degrees, six sectors
__device__ __forceinline__ V3 hsv2rgb(float hue, float sat, float val)
{
#pragma clang fp contract(off)
	const float h = hue / 60.0f;
	const int i = (int)floorf(hue / 60.0f);
	const float fr = h - (float)i;
	const float pp = val * (1.0f - sat);
	const float q = val * (1.0f - sat * fr);
	const float t = val * (1.0f - sat * (1.0f - fr));
	switch (i)
	{
		case 0: return v3(val, t, pp);
		case 1: return v3(q, val, pp);
		case 2: return v3(pp, val, t);
		case 3: return v3(pp, q, val);
		case 4: return v3(t, pp, val);
		case 5: return v3(val, pp, q);
		default: return v3(0.0f, 0.0f, 0.0f);
	}
}

// CalculateViewVector (projection_3d.cpp:41-89) of the fish-eye (1, 3) and equirectangular (2)
// projections in FP64, normalised, then turned by the camera rotation in the kind's precision; the
// march normalises the result again as doWork does
template <int KIND>
__device__ __forceinline__ typename PrecOf<KIND>::P persp_view(const FrameConst &f, double ix, double iy)
{
#pragma clang fp contract(off)
	using Q = PrecOf<KIND>;
	DV3 v;
	if (f.perspective == 2)
	{
		const double vx = ix * 0.5, vy = iy * 0.5;
		const double cy = cos(f.fovD * vy);
		v = dv3(sin(f.fovD * vx) * cy, cos(f.fovD * vx) * cy, sin(f.fovD * vy));
	}
	else
	{
		const double r = sqrt(ix * ix + iy * iy);
		if (r == 0.0)
			v = dv3(0.0, 1.0, 0.0);
		else
		{
			const double s = sin(r * f.fovD);
			v = dv3(ix / r * s, cos(r * f.fovD), iy / r * s);
		}
	}
	v = v * (1.0 / sqrt(ddot(v, v)));
	return Q::cam_rot(f, Q::fromD(v.x, v.y, v.z));
}

// One pixel: the sample loop of doWork and RayRecursion at depth 0.  VAR = 0 is the variant for
// frames without Monte Carlo sampling, GI, shadows and AO; 1 carries them; 2 adds the Monte Carlo
// lens, chromatic aberration and fog.
// The optional channels of out (world position, specular, shadow, GI) are stored per sample as the
// reference overwrites its per-pixel values (render_worker.cpp:355-366), the GI plane summed over
// the samples and divided by their count (:474-476), so they hold no registers across the loop.
template <int KIND, int VAR>
__device__ void render_pixel(const FrameConst &frame, int xs, int ys, PixelOut &po, Work &w, const Outputs &out,
	size_t oi)
{
#pragma clang fp contract(off)
	constexpr bool FULL = VAR >= 1;
	constexpr bool LENS = VAR == 2;
	const FrameConst &f = uniform_ref(frame);
	using Q = PrecOf<KIND>;
	using R = typename Q::R;
	using P = typename Q::P;
	const uint32_t pix = (uint32_t)ys * (uint32_t)f.width + (uint32_t)xs;
	uint32_t seed = pixel_seed(pix, f.randomSeed);
	const R aspect = Q::aspect(f);
	const R ox = (Q::invWidth(f) * (R)xs - (R)0.5) * aspect;
	const R oy = (R)0.5 - Q::invHeight(f) * (R)ys;
	const P cam = Q::cam(f);
	int repeats = (FULL && f.monteCarlo) ? f.samples : 1;
	// antialiasing multiplies the samples by the grid (render_worker.cpp:185-186); the projections
	// and the grid come with the full variant only (the host picks it for such frames and refuses
	// them with the lens variant's effects)
	constexpr bool EXTRA = VAR == 1;
	const bool persp = EXTRA && f.perspective != 0;
	const int aa = EXTRA ? f.aaSize : 0;
	if (aa) repeats *= aa * aa;
	// the fish-eye cut marches nothing outside the hemisphere (render_worker.cpp:165-170): every
	// sample is black at depth 1e20 with fog opacity 1 and no channels, whatever its random draws
	// (the pixel's own stream), so the pixel's outputs are known without its sample loop
	if (EXTRA && f.perspective == 3)
	{
		const double oxd = (f.invWidthD * (double)xs - 0.5) * f.aspectRatioD, oyd = 0.5 - f.invHeightD * (double)ys;
		if (sqrt(oxd * oxd + oyd * oyd) > f.fishCutD)
		{
			po.rgb = v3(0.0f, 0.0f, 0.0f);
			po.depth = 1e20f;
			po.found = 0;
			po.iters = 0;
			po.steps = 0;
			po.normal = v3(0.0f, 0.0f, 0.0f);
			po.colour[0] = po.colour[1] = po.colour[2] = 0;
			po.opacity = 65535u;
			for (int k = 0; k < 3; k++)
			{
				if (out.world) out.world[3 * oi + k] = 0.0f;
				if (out.specular) out.specular[3 * oi + k] = 0.0f;
				if (out.shadow) out.shadow[3 * oi + k] = 0.0f;
				if (out.gi) out.gi[3 * oi + k] = 0.0f;
			}
			return;
		}
	}
	V3 acc = v3(0.0f, 0.0f, 0.0f);
	V3 stdDevSum = v3(0.0f, 0.0f, 0.0f);
	po.steps = 0;
	uint32_t colR = 0, colG = 0, colB = 0, opSum = 0;
	for (int rep = 0; rep < repeats; rep++)
	{
		R ix = ox, iy = oy;
		if (aa)
		{
			// the sample's place on the sub-pixel grid (render_worker.cpp:187-199); with Monte Carlo
			// samples the x step runs on past the pixel as the reference's does
			const int xStep = rep / aa, yStep = rep - xStep * aa;
			ix = (R)((double)ox + (double)xStep / aa / (double)f.width * f.aspectRatioD);
			iy = (R)((double)oy + (double)yStep / aa / (double)f.height);
		}
		else if (FULL && f.monteCarlo)
		{
			ix = ox + ((R)random_mod(seed, 1000) * (R)1e-3 - (R)0.5) * Q::invWidth(f) * aspect;
			iy = oy + ((R)random_mod(seed, 1000) * (R)1e-3 - (R)0.5) * Q::invHeight(f);
		}
		P dir, start = cam;
		float hue = 0.0f; // chromatic aberration: the sample's hue (its tint is applied after shading)
		const bool lens = LENS && (f.dofEnabled || f.chromaticAberration);
		if (lens)
		{
			P v = Q::view_raw(f, ix, iy);
			if (f.dofEnabled)
			{
				// MonteCarloDOF (render_worker.cpp:1832-1844): a random point of the lens, turned by
				// the camera rotation, becomes the ray origin; the view vector loses 1/focus of it so
				// that the ray still crosses the undisturbed one at the focal distance
				const double randR =
					0.0015 * f.dofRadiusD * f.dofFocusD * sqrt((double)random_mod(seed, 65536) / 65536.0);
				// the angle is a whole number of radians: its sine and cosine come from the table
				const int randAngle = random_mod(seed, kLensAngles - 1);
				const double sa = f.lensSinCos[2 * randAngle], ca = f.lensSinCos[2 * randAngle + 1];
				const P rr = Q::cam_rot(f, P{(R)(randR * sa), (R)0, (R)(randR * ca)});
				const R invFocus = (R)(1.0 / f.dofFocusD);
				v = P{v.x - rr.x * invFocus, v.y - rr.y * invFocus, v.z - rr.z * invFocus};
				start = P{cam.x + rr.x, cam.y + rr.y, cam.z + rr.z};
			}
			if (f.chromaticAberration)
			{
				// a hue per sample bends the view vector and tints the sample
				// (render_worker.cpp:255-264, 374-380)
				hue = (float)((double)random_mod(seed, 3600) / 10.0);
				const P cv = Q::cam_rot(f, P{(R)0, (R)(double)(hue / 20000.0f * f.caCameraDispersion), (R)0});
				v = Q::normalize_p(P{v.x - cv.x, v.y - cv.y, v.z - cv.z});
			}
			dir = Q::normalize_p(v);
		}
		else if (persp)
			dir = Q::normalize_p(persp_view<KIND>(f, (double)ix, (double)iy));
		else
			dir = Q::view(f, ix, iy);
		MarchOut<KIND> m;
		FogAcc fogAcc;
		fog_init(fogAcc);
		const bool fogOn = LENS && f.fog;
		if constexpr (PosPrecision<KIND>::kFP64)
			ray_marching<KIND, 2, LENS>(f, start, dir, (R)f.viewDistanceMax, seed, m, w, !lens, &fogAcc, fogOn);
		else if (f.advancedQuality)
			ray_marching<KIND, 1, LENS>(f, start, dir, (R)f.viewDistanceMax, seed, m, w, !lens, &fogAcc, fogOn);
		else
			ray_marching<KIND, 0, LENS>(f, start, dir, (R)f.viewDistanceMax, seed, m, w, !lens, &fogAcc, fogOn);
		V3 rgb;
		V3 n = v3(0.0f, 0.0f, 0.0f);
		if (out.world)
		{
			const V3 pt = Q::tof(m.point);
			out.world[3 * oi + 0] = pt.x;
			out.world[3 * oi + 1] = pt.y;
			out.world[3 * oi + 2] = pt.z;
		}
		// the sample's specular, shadow and GI channels: a miss has none, and a shadow of 1
		// (render_worker.cpp:1606-1612)
		V3 specCh = v3(0.0f, 0.0f, 0.0f), giCh = v3(0.0f, 0.0f, 0.0f);
		float shadowCh = 1.0f;
		if (m.found)
		{
			const R delta = calc_delta<KIND>(f, m.point);
			P nP;
			if (FULL && f.slowShading)
				nP = slow_normals<KIND>(f, m.point,
					(KIND == kGeneric && f.interiorMode) ? (double)m.distThresh * 0.2 : (double)delta, w);
			else
				nP = calculate_normals<KIND>(f, m.point, m.distThresh, w);
			n = Q::tof(nP);
			const V3 colour = FULL ? surface_colour<KIND>(f, m.point, w)
								   : v3(f.surfaceColor[0], f.surfaceColor[1], f.surfaceColor[2]);
			V3 spec;
			rgb = object_shader<KIND, FULL>(f, m.point, nP, m.distThresh, delta, Q::tof(dir), colour, spec, seed, w,
				&shadowCh);
			if (FULL && f.monteCarlo && f.globalIllumination)
			{
				giCh = global_illumination<KIND>(f, m.point, nP, m.distThresh, colour, seed, w);
				rgb = rgb + giCh;
			}
			if (FULL)
			{
				rgb = rgb + spec; // resultShader = objectShader + specular (render_worker.cpp:1458-1462)
				specCh = spec;
			}
			rgb = v3(fmaxf(rgb.x, 0.0f), fmaxf(rgb.y, 0.0f), fmaxf(rgb.z, 0.0f));
			// the colour buffer holds the hit's surface colour; a miss keeps (0, 0, 0)
			colR += colour_byte(colour.x);
			colG += colour_byte(colour.y);
			colB += colour_byte(colour.z);
		}
		else
			rgb = background_shader(f, Q::tof(dir));
		if (LENS && f.fog)
		{
			// the volumetric walk as accumulated by the march; the glow colour needs the step count
			const V3 g = glow_colour(f, m.stepCount);
			rgb = v3(fogAcc.A * rgb.x + fogAcc.gc * g.x + fogAcc.B.x, fogAcc.A * rgb.y + fogAcc.gc * g.y + fogAcc.B.y,
				fogAcc.A * rgb.z + fogAcc.gc * g.z + fogAcc.B.z);
			float op = 1.0f - fogAcc.T;
			op = op > 1.0f ? 1.0f : op;
			opSum += (uint32_t)(uint16_t)(int)((double)op * 65535.0);
		}
		else if (FULL && f.glow)
			rgb = glow_shader(f, m.stepCount, rgb);
		if (LENS && f.chromaticAberration)
		{
			const V3 tint = hsv2rgb(fmodf(360.0f + hue - 60.0f, 360.0f), 1.0f, 2.0f);
			rgb = v3(rgb.x * tint.x, rgb.y * tint.y, rgb.z * tint.z);
		}
		if (out.specular)
		{
			out.specular[3 * oi + 0] = specCh.x;
			out.specular[3 * oi + 1] = specCh.y;
			out.specular[3 * oi + 2] = specCh.z;
		}
		if (out.shadow)
		{
			out.shadow[3 * oi + 0] = shadowCh;
			out.shadow[3 * oi + 1] = shadowCh;
			out.shadow[3 * oi + 2] = shadowCh;
		}
		if (out.gi)
		{
			if (rep > 0) giCh = v3(out.gi[3 * oi + 0] + giCh.x, out.gi[3 * oi + 1] + giCh.y, out.gi[3 * oi + 2] + giCh.z);
			out.gi[3 * oi + 0] = giCh.x;
			out.gi[3 * oi + 1] = giCh.y;
			out.gi[3 * oi + 2] = giCh.z;
		}
		acc = acc + rgb;
		if (rep == 0)
		{
			po.found = m.found;
			po.iters = m.iters;
			po.normal = n;
		}
		po.depth = m.found ? (float)m.depth : 1e20f;
		po.steps += m.steps;
		if (FULL && f.monteCarlo)
		{
			mc_deviation(rgb, rep, acc, stdDevSum);
			if (rep > f.DOFMinSamples && mc_noise(rep, acc, stdDevSum) < f.DOFMaxNoise01)
			{
				repeats = rep + 1; // adaptive stop (render_worker.cpp:436-449)
				break;
			}
		}
	}
	po.rgb = repeats > 1 ? v3(acc.x / (float)repeats, acc.y / (float)repeats, acc.z / (float)repeats) : acc;
	if (out.gi && repeats > 1)
		for (int k = 0; k < 3; k++) out.gi[3 * oi + k] = out.gi[3 * oi + k] / (float)repeats;
	// integer means of the colour and opacity buffers (render_worker.cpp:469-472)
	const uint32_t reps = (uint32_t)repeats;
	po.colour[0] = colR / reps;
	po.colour[1] = colG / reps;
	po.colour[2] = colB / reps;
	po.opacity = opSum / reps;
}

__device__ __forceinline__ void wave_add_u64(unsigned long long *dst, uint32_t v)
{
	unsigned long long s = v;
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
	if ((threadIdx.x & 63) == 0) atomicAdd(dst, s);
}

__device__ __forceinline__ void write_pixel(const Outputs &out, size_t i, const PixelOut &po)
{
	if (out.rgb)
	{
		out.rgb[3 * i + 0] = po.rgb.x;
		out.rgb[3 * i + 1] = po.rgb.y;
		out.rgb[3 * i + 2] = po.rgb.z;
	}
	if (out.depth) out.depth[i] = po.depth;
	if (out.iters) out.iters[i] = po.iters;
	if (out.mask) out.mask[i] = (uint8_t)po.found;
	if (out.normal)
	{
		out.normal[3 * i + 0] = po.normal.x;
		out.normal[3 * i + 1] = po.normal.y;
		out.normal[3 * i + 2] = po.normal.z;
	}
	if (out.steps) out.steps[i] = po.steps;
	if (out.colour)
	{
		out.colour[3 * i + 0] = (uint8_t)po.colour[0];
		out.colour[3 * i + 1] = (uint8_t)po.colour[1];
		out.colour[3 * i + 2] = (uint8_t)po.colour[2];
	}
	if (out.opacity) out.opacity[i] = (uint16_t)po.opacity;
}

} // namespace mbx

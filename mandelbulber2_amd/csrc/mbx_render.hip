// mbx_render.hip - the per-pixel raymarch kernels for gfx950.
//
// One lane per primary ray.  A 64-lane workgroup (one wave) covers an 8x8 pixel tile so that
// neighbouring rays, which take similar march paths, share a wave.  Everything runs in FP32;
// the per-frame constants are wave-uniform scalar loads.
//
// Path (reference): cRenderWorker::doWork pixel body (render_worker.cpp:155-525) ->
// RayRecursion with reflections off (render_worker.cpp:969-1819) -> RayMarching
// (render_worker.cpp:708-951) -> CalculateDistance -> Compute -> FormulaCode; on a hit
// CalculateNormals (shader_calculate_normals.cpp:40-130), ObjectShader (shader_object.cpp:40-229)
// with LightShading / AuxShadow / AO, and the GI bounce loop (shader_global_illumination.cpp).
#include <hip/hip_runtime.h>

#include "mbx_device.h"
#include "mbx_launch.h"

namespace mbx
{

struct MarchOut
{
	V3 point;
	float depth;
	float distThresh;
	float lastDist;
	int found;
	int iters;
	int steps;
};

__device__ __forceinline__ float dist_thresh(const FrameConst &f, V3 p)
{
	// CalcDistThresh (render_worker.cpp:653-679)
	if (f.constantDEThreshold) return f.DEThresh;
	return length(v3(f.camera[0], f.camera[1], f.camera[2]) - p) * f.threshK;
}

__device__ __forceinline__ float calc_delta(const FrameConst &f, V3 p)
{
	// CalcDelta (render_worker.cpp:684-690)
	return length(v3(f.camera[0], f.camera[1], f.camera[2]) - p) * f.deltaK;
}

// RayMarching (render_worker.cpp:708-951), exterior mode, binary search on
template <int KIND>
__device__ void ray_marching(const FrameConst &f, V3 start, V3 dir, float maxScan,
	uint32_t &seed, MarchOut &out, Work &w)
{
	V3 point = v3(1e30f, 1e30f, 1e30f);
	bool found = false, dead = false;
	float scan = 0.0f, dist = 0.0f, step = 0.0f, thr = 0.0f;
	int counter = 0, iters = 0;
	const float dirLength = length(dir);
	for (int i = 0; i < f.maxRaymarchingSteps; i++)
	{
		const V3 last = point;
		counter++;
		point = start + dir * scan;
		if ((point.x == last.x && point.y == last.y && point.z == last.z) || !finite3(point))
		{
			point = last;
			found = true;
			dead = true;
			break;
		}
		thr = dist_thresh(f, point);
		dist = calculate_distance<KIND>(point, thr, f, iters, w);
		if (dist < thr)
		{
			found = true;
			break;
		}
		const float jitter = 1.0f - (float)random_mod(seed, 1000) * 1e-4f;
		step = (dist - 0.5f * thr) * f.DEFactor * jitter;
		step = fminf(step, 3.0f);
		scan += step / dirLength;
		if (scan > maxScan) break;
	}
	point = start + dir * scan;
	if (found && !dead)
	{
		step *= 0.5f;
		for (int i = 0; i < 30; i++)
		{
			counter++;
			if (dist < thr && dist > thr * f.searchLimit) break;
			if (dist > thr)
				scan += step;
			else if (dist < thr * f.searchLimit)
				scan -= step;
			point = start + dir * scan;
			thr = dist_thresh(f, point);
			dist = calculate_distance<KIND>(point, thr, f, iters, w);
			step *= 0.5f;
		}
	}
	out.point = point;
	out.depth = scan;
	out.distThresh = thr;
	out.lastDist = dist;
	out.found = found;
	out.iters = iters;
	out.steps = counter;
}

// CalculateNormals, fast mode: central differences of the DE (shader_calculate_normals.cpp:44-78)
template <int KIND>
__device__ V3 calculate_normals(const FrameConst &f, V3 p, float thr, Work &w)
{
	const float delta = thr * f.smoothness;
	int it;
	const float sx1 = calculate_distance<KIND>(p + v3(delta, 0.0f, 0.0f), thr, f, it, w);
	const float sx2 = calculate_distance<KIND>(p - v3(delta, 0.0f, 0.0f), thr, f, it, w);
	const float sy1 = calculate_distance<KIND>(p + v3(0.0f, delta, 0.0f), thr, f, it, w);
	const float sy2 = calculate_distance<KIND>(p - v3(0.0f, delta, 0.0f), thr, f, it, w);
	const float sz1 = calculate_distance<KIND>(p + v3(0.0f, 0.0f, delta), thr, f, it, w);
	const float sz2 = calculate_distance<KIND>(p - v3(0.0f, 0.0f, delta), thr, f, it, w);
	V3 n = v3(sx1 - sx2, sy1 - sy2, sz1 - sz2);
	if (n.x == 0.0f && n.y == 0.0f && n.z == 0.0f)
		n = v3(1.0f, 0.0f, 0.0f);
	else
		n = normalize(n);
	if (!finite3(n)) n = v3(1.0f, 0.0f, 0.0f);
	return n;
}

// AuxShadow for the directional light (shader_aux_shadow.cpp:39-323): hard, or the soft
// cone variant when soft_shadow_cone > 0
template <int KIND>
__device__ float aux_shadow(const FrameConst &f, V3 p, float thr, float distance, V3 lv, Work &w)
{
	float shadowTemp = 1.0f;
	float maxSoft = 0.0f;
	float step = thr;
	int count = 0;
	int it;
	for (float i = thr; i < distance; i += step)
	{
		const V3 p2 = p + lv * i;
		const float dist = calculate_distance<KIND>(p2, thr, f, it, w);
		if (f.softShadows)
		{
			float angle = (dist - thr) / i;
			if (angle < 0.0f || dist < thr) angle = 0.0f;
			float soft = 1.0f - angle / f.softRange;
			if (f.penetrating) soft *= (distance - i) / distance;
			soft = fmaxf(soft, 0.0f);
			maxSoft = fmaxf(maxSoft, soft);
		}
		shadowTemp = 1.0f;
		if (dist < thr)
		{
			if (f.penetrating)
				shadowTemp = fmaxf(shadowTemp - (distance - i) / distance, 0.0f);
			else
				shadowTemp = 0.0f;
			break;
		}
		step = fmaxf(fminf(dist, 1e6f) * f.DEFactor, 1e-15f);
		count++;
		if (count > f.maxRaymarchingSteps) break;
	}
	return f.softShadows ? 1.0f - maxSoft : shadowTemp;
}

// AmbientOcclusion, multiple rays (shader_ambient_occlusion.cpp:41-122)
template <int KIND>
__device__ V3 ambient_occlusion(const FrameConst &f, V3 p, float thr, float delta, uint32_t &seed,
	Work &w)
{
	V3 ao = v3(0.0f, 0.0f, 0.0f);
	const float end_dist = delta * (float)f.height; // delta / resolution
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
		const V3 v = v3(f.aoVectors[i][0], f.aoVectors[i][1], f.aoVectors[i][2]);
		float shadowTemp = 1.0f;
		float dist;
		for (float r = thr; r < end_dist; r += dist * 2.0f)
		{
			dist = calculate_distance<KIND>(p + v * r, thr, f, it, w);
			if (dist < thr)
			{
				shadowTemp = fmaxf(shadowTemp - (end_dist - r) / end_dist, 0.0f);
				break;
			}
		}
		ao = ao + v3(f.aoColors[i][0], f.aoColors[i][1], f.aoColors[i][2]) * shadowTemp;
	}
	if (!f.monteCarlo) ao = ao * (1.0f / (float)f.aoCount);
	return ao;
}

// FastAmbientOcclusion (shader_fast_ambient_occlusion.cpp:40-66)
template <int KIND>
__device__ float fast_ambient_occlusion(const FrameConst &f, V3 p, V3 n, float thr, Work &w)
{
	float aoTemp = 0.0f;
	float lastDist = 1e20f;
	const int q2 = f.aoQuality * f.aoQuality;
	int it;
	for (int i = 1; i < q2; i++)
	{
		const float scan = (float)(i * i) * thr;
		float dist = calculate_distance<KIND>(p + n * scan, thr, f, it, w);
		if (dist > lastDist * 2.0f) dist = lastDist * 2.0f;
		lastDist = dist;
		aoTemp += exp2f(-(float)i) * (scan - f.aoFastTune * dist) / thr;
	}
	return fmaxf(1.0f - 0.2f * aoTemp, 0.0f);
}

// BackgroundShader: 3-colour gradient and the directional light's disk (shader_background.cpp)
__device__ V3 background_shader(const FrameConst &f, V3 viewVector)
{
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
		const float a = -(dot(vn, ld) - 1.0f) * 360.0f / f.lightSize;
		const float pw = 6.0f * f.lightContourSharpness;
		const float ap = (pw == 6.0f) ? (a * a) * (a * a) * (a * a) : powf(a, pw);
		const float intensity = 1.0f / (1.0f + ap) * f.lightVisibility * f.lightIntensity;
		out = out + v3(f.lightColor[0], f.lightColor[1], f.lightColor[2]) * intensity;
	}
	return out;
}

// ObjectShader with one directional light, Lambert, optional shadow and AO
// (shader_object.cpp:40-229, shader_light_shading.cpp:40-122)
template <int KIND>
__device__ V3 object_shader(const FrameConst &f, V3 p, V3 n, float thr, float delta,
	uint32_t &seed, Work &w)
{
	const V3 colour = v3(f.surfaceColor[0], f.surfaceColor[1], f.surfaceColor[2]);
	V3 ambient = v3(0.0f, 0.0f, 0.0f);
	if (f.aoEnabled)
	{
		if (f.aoMode == MBX_AO_FAST)
		{
			const float a = fast_ambient_occlusion<KIND>(f, p, n, thr, w);
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
		float shadow = 1.0f;
		if (f.castShadows)
		{
			if (shade > 0.001f)
			{
				const float distance = f.penetrating ? delta * (float)f.height : f.viewDistanceMaxLight;
				shadow = aux_shadow<KIND>(f, p, thr, distance, lv, w);
			}
			else
				shadow = 0.0f;
		}
		light = v3(f.lightColor[0], f.lightColor[1], f.lightColor[2]) * (shade * shadow);
	}
	V3 out;
	out.x = (f.fillLight[0] + ambient.x * f.aoColor[0]) * colour.x + light.x * colour.x + f.luminosity[0];
	out.y = (f.fillLight[1] + ambient.y * f.aoColor[1]) * colour.y + light.y * colour.y + f.luminosity[1];
	out.z = (f.fillLight[2] + ambient.z * f.aoColor[2]) * colour.z + light.z * colour.z + f.luminosity[2];
	return out;
}

// GlobalIlumination (shader_global_illumination.cpp:41-209)
template <int KIND>
__device__ V3 global_illumination(const FrameConst &f, V3 p0, V3 n0, float thr0, uint32_t &seed,
	Work &w)
{
	V3 out = v3(0.0f, 0.0f, 0.0f);
	V3 p = p0, n = n0;
	float thrCur = thr0;
	const V3 colour = v3(f.surfaceColor[0], f.surfaceColor[1], f.surfaceColor[2]);
	V3 colourTemp = colour;
	for (int depth = 0; depth < f.reflectionsMax; depth++)
	{
		const float rx = (float)(random_mod(seed, 20000) - 10000) * 1e-4f;
		const float ry = (float)(random_mod(seed, 20000) - 10000) * 1e-4f;
		const float rz = (float)(random_mod(seed, 20000) - 10000) * 1e-4f;
		const V3 dir = normalize(n + v3(rx * 1.2f, ry * 1.2f, rz * 1.2f));
		float dist = 0.0f;
		bool found = false;
		int stepCount = 0;
		int it;
		V3 hit = p;
		for (float scan = thrCur; scan < f.viewDistanceMax; scan += dist * f.DEFactor)
		{
			if (stepCount >= f.maxRaymarchingSteps) break;
			const V3 q = p + dir * scan;
			const float thr = dist_thresh(f, q);
			thrCur = thr;
			dist = calculate_distance<KIND>(q, thr, f, it, w);
			if (dist < thr)
			{
				if (scan < thr * 2.0f) return out;
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
			n = calculate_normals<KIND>(f, p, thrCur, w);
			result = object_shader<KIND>(f, p, n, thrCur, thrCur, seed, w);
			out = out + v3(result.x * colourTemp.x, result.y * colourTemp.y, result.z * colourTemp.z);
		}
		else
		{
			result = background_shader(f, dir);
			finished = true;
		}
		const float lim = f.giRadianceLimit;
		result = v3(fminf(fmaxf(result.x, 0.0f), lim), fminf(fmaxf(result.y, 0.0f), lim),
			fminf(fmaxf(result.z, 0.0f), lim));
		out = out + v3(result.x * colourTemp.x, result.y * colourTemp.y, result.z * colourTemp.z);
		colourTemp = colour;
		if (finished) break;
	}
	return out;
}

struct PixelOut
{
	V3 rgb;
	V3 normal;
	float depth;
	int found;
	int iters;
	int steps;
};

// One pixel: the sample loop of doWork and RayRecursion at depth 0
template <int KIND>
__device__ void render_pixel(const FrameConst &f, int xs, int ys, PixelOut &po, Work &w)
{
	const uint32_t pix = (uint32_t)ys * (uint32_t)f.width + (uint32_t)xs;
	uint32_t seed = pixel_seed(pix, f.randomSeed);
	const float ox = (f.invWidth * (float)xs - 0.5f) * f.aspectRatio;
	const float oy = 0.5f - f.invHeight * (float)ys;
	const V3 cam = v3(f.camera[0], f.camera[1], f.camera[2]);
	const int repeats = f.monteCarlo ? f.samples : 1;
	V3 acc = v3(0.0f, 0.0f, 0.0f);
	po.steps = 0;
	for (int rep = 0; rep < repeats; rep++)
	{
		float ix = ox, iy = oy;
		if (f.monteCarlo)
		{
			ix = ox + ((float)random_mod(seed, 1000) * 1e-3f - 0.5f) * f.invWidth * f.aspectRatio;
			iy = oy + ((float)random_mod(seed, 1000) * 1e-3f - 0.5f) * f.invHeight;
		}
		// CalculateViewVector three-point (projection_3d.cpp:77-89), then normalise again
		V3 vv = normalize(v3(ix * f.fov, 1.0f, iy * f.fov));
		const V3 dir = normalize(rot(f.mRot, vv));
		MarchOut m;
		ray_marching<KIND>(f, cam, dir, f.viewDistanceMax, seed, m, w);
		V3 rgb;
		V3 n = v3(0.0f, 0.0f, 0.0f);
		if (m.found)
		{
			n = calculate_normals<KIND>(f, m.point, m.distThresh, w);
			const float delta = calc_delta(f, m.point);
			rgb = object_shader<KIND>(f, m.point, n, m.distThresh, delta, seed, w);
			if (f.monteCarlo && f.globalIllumination)
				rgb = rgb + global_illumination<KIND>(f, m.point, n, m.distThresh, seed, w);
			rgb = v3(fmaxf(rgb.x, 0.0f), fmaxf(rgb.y, 0.0f), fmaxf(rgb.z, 0.0f));
		}
		else
			rgb = background_shader(f, dir);
		acc = acc + rgb;
		if (rep == 0)
		{
			po.found = m.found;
			po.iters = m.iters;
			po.normal = n;
		}
		po.depth = m.found ? m.depth : 1e20f;
		po.steps += m.steps;
	}
	po.rgb = repeats > 1 ? acc * (1.0f / (float)repeats) : acc;
}

__device__ __forceinline__ void wave_add_u64(unsigned long long *dst, uint32_t v)
{
	unsigned long long s = v;
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
	if ((threadIdx.x & 63) == 0) atomicAdd(dst, s);
}

// Plain per-frame kernel: 64 lanes = one 8x8 tile.
template <int KIND>
__global__ __launch_bounds__(64) void render_kernel(const FrameConst *__restrict__ frame,
	Outputs out, unsigned long long *__restrict__ counters)
{
	const FrameConst &f = *frame;
	const int tilesX = (f.width + 7) >> 3;
	const int tile = blockIdx.x;
	const int tx = tile % tilesX, ty = tile / tilesX;
	const int xs = tx * 8 + (threadIdx.x & 7);
	const int ys = ty * 8 + (threadIdx.x >> 3);
	Work w = {0u, 0u};
	const bool active = xs < f.width && ys < f.height;
	if (active)
	{
		PixelOut po;
		render_pixel<KIND>(f, xs, ys, po, w);
		const size_t i = (size_t)ys * (size_t)f.width + (size_t)xs;
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
	}
	if (counters)
	{
		wave_add_u64(&counters[0], w.de_evals);
		wave_add_u64(&counters[1], w.iterations);
	}
}

// Point DE query (distance engine analogue)
template <int KIND>
__global__ __launch_bounds__(64) void distance_kernel(const FrameConst *__restrict__ frame,
	const float *__restrict__ pts, int n, float detailSize, float *__restrict__ dist,
	int32_t *__restrict__ iters)
{
	const int i = blockIdx.x * 64 + threadIdx.x;
	if (i >= n) return;
	Work w = {0u, 0u};
	int it = 0;
	const V3 p = v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
	dist[i] = calculate_distance<KIND>(p, detailSize, *frame, it, w);
	iters[i] = it;
}

template <int KIND>
static hipError_t launch_render_t(const FrameConst *frame, int width, int height, const Outputs &out,
	unsigned long long *counters, hipStream_t stream)
{
	const int tiles = ((width + 7) / 8) * ((height + 7) / 8);
	hipLaunchKernelGGL(render_kernel<KIND>, dim3(tiles), dim3(64), 0, stream, frame, out, counters);
	return hipGetLastError();
}

template <int KIND>
static hipError_t launch_distance_t(const FrameConst *frame, const float *pts, int n, float detail,
	float *dist, int32_t *iters, hipStream_t stream)
{
	hipLaunchKernelGGL(distance_kernel<KIND>, dim3((n + 63) / 64), dim3(64), 0, stream, frame, pts, n,
		detail, dist, iters);
	return hipGetLastError();
}

hipError_t launch_render(int deKind, const FrameConst *frame, int width, int height,
	const Outputs &out, unsigned long long *counters, hipStream_t stream)
{
	switch (deKind)
	{
		case kBulb8: return launch_render_t<kBulb8>(frame, width, height, out, counters, stream);
		case kBulbTrig: return launch_render_t<kBulbTrig>(frame, width, height, out, counters, stream);
		case kBoxAnalytic:
			return launch_render_t<kBoxAnalytic>(frame, width, height, out, counters, stream);
		case kBoxDelta: return launch_render_t<kBoxDelta>(frame, width, height, out, counters, stream);
		case kIfsAnalytic:
			return launch_render_t<kIfsAnalytic>(frame, width, height, out, counters, stream);
		default: return hipErrorInvalidValue;
	}
}

hipError_t launch_distance(int deKind, const FrameConst *frame, const float *pts, int n,
	float detail, float *dist, int32_t *iters, hipStream_t stream)
{
	switch (deKind)
	{
		case kBulb8: return launch_distance_t<kBulb8>(frame, pts, n, detail, dist, iters, stream);
		case kBulbTrig: return launch_distance_t<kBulbTrig>(frame, pts, n, detail, dist, iters, stream);
		case kBoxAnalytic:
			return launch_distance_t<kBoxAnalytic>(frame, pts, n, detail, dist, iters, stream);
		case kBoxDelta: return launch_distance_t<kBoxDelta>(frame, pts, n, detail, dist, iters, stream);
		case kIfsAnalytic:
			return launch_distance_t<kIfsAnalytic>(frame, pts, n, detail, dist, iters, stream);
		default: return hipErrorInvalidValue;
	}
}

} // namespace mbx

// mbx_render.hip - the raymarch kernels for gfx950 (plain per-frame, persistent batch,
// distance query) and their launchers; the per-pixel path is in mbx_pixel.h.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mbx_pixel.h"

namespace mbx
{

// Occupancy targets (min waves per SIMD in __launch_bounds__).  The general formula kind and the
// two Mandelbox kinds take 3 (168 VGPRs): the general kind's out-of-line DE fits that without
// spilling, and 3 measured 1.32x faster than 2 on the hybrid002 scene although its full variant
// then spills ~200 B per lane (4 waves gained 2.5% more at 50x the spill traffic); the Mandelbox
// delta-DE full variant (shadow + AO scene) went from 256 VGPRs and 256 B of spills at 2 waves to
// 168 VGPRs, 1.22x faster.  The IFS and FP64 Mandelbulb kinds take 2 (256 VGPRs), the FP32
// Mandelbulb kinds 1 (no cap needed).  The distance query shares the general kind's out-of-line
// DE, so it carries the same target for that kind.
template <int KIND>
constexpr int occupancy_target()
{
	return (KIND == kGeneric || KIND == kBoxDelta || KIND == kBoxAnalytic) ? 3 : (KIND >= kBoxAnalytic ? 2 : 1);
}

// Plain per-frame kernel: 64 lanes = one tile of (1 << tileShift) x (64 >> tileShift) pixels
// (8x8 by default), one workgroup per tile, one launch per frame.
template <int KIND, int VAR>
__global__ __launch_bounds__(64, occupancy_target<KIND>()) void render_kernel(const FrameConst *__restrict__ frame,
	Outputs out, unsigned long long *__restrict__ counters, int tileShift)
{
	const FrameConst &f = *frame;
	const int tilesX = (f.width + (1 << tileShift) - 1) >> tileShift;
	const int tile = blockIdx.x;
	const int tx = tile % tilesX, ty = tile / tilesX;
	const int xs = (tx << tileShift) + (threadIdx.x & ((1 << tileShift) - 1));
	const int ys = ty * (64 >> tileShift) + (threadIdx.x >> tileShift);
	Work w = make_work(f);
	if (xs < f.width && ys < f.height)
	{
		PixelOut po;
		const size_t pix = (size_t)ys * (size_t)f.width + (size_t)xs;
		render_pixel<KIND, VAR>(f, xs, ys, po, w, out, pix);
		write_pixel(out, pix, po);
	}
	if (counters)
	{
		unsigned long long *slot = counters + (blockIdx.x % kCounterSlots) * kCounterStride;
		wave_add_u64(&slot[0], w.de_evals);
		wave_add_u64(&slot[1], w.iterations);
		wave_add_u64(&slot[2], w.far_steps);
	}
}

// Batch grid kernel: one 64-lane workgroup per tile over every frame of a batch (frames of
// different resolution, camera and formula parameters), in a single launch.  A workgroup finds
// its frame by a binary search over the frames' tile ranges (wave-uniform scalar loads), so the
// frame constants stay in scalar registers.  One launch (replayed from a captured graph)
// replaces the reference's one-job-per-frame animation loop; the hardware dispatcher balances
// the tiles without any atomic.
static_assert(kCounterSlots == 64, "workgroup 0 of batch_kernel clears one counter slot per lane");

// When clearNext is set, workgroup 0 zeroes those counter slots (another bank than `counters`):
// the next launch accumulates there without a clear kernel in front of it.
template <int KIND, int VAR>
__global__ __launch_bounds__(64, occupancy_target<KIND>()) void batch_kernel(const FrameConst *__restrict__ frames,
	const FrameJob *__restrict__ jobs, int numFrames, unsigned long long *__restrict__ counters, int tileShift,
	unsigned long long *__restrict__ clearNext)
{
	const int lane = threadIdx.x;
	const unsigned int t = blockIdx.x;
	if (clearNext && t == 0)
		for (unsigned int k = 0; k < kCounters; k++) clearNext[lane * kCounterStride + k] = 0ull;
	int lo = 0, hi = numFrames - 1;
	while (lo < hi)
	{
		const int mid = (lo + hi) >> 1;
		if (t < (unsigned int)jobs[mid].tileEnd)
			hi = mid;
		else
			lo = mid + 1;
	}
	const int fr = lo;
	const FrameConst &f = frames[fr];
	const FrameJob &job = jobs[fr];
	Work w = make_work(f);
	const unsigned int tileBegin = fr == 0 ? 0u : (unsigned int)jobs[fr - 1].tileEnd;
	const int lt = (int)(t - tileBegin);
	const int tilesX = (f.width + (1 << tileShift) - 1) >> tileShift;
	const int xs = ((lt % tilesX) << tileShift) + (lane & ((1 << tileShift) - 1));
	const int ys = (lt / tilesX) * (64 >> tileShift) + (lane >> tileShift);
	if (xs < f.width && ys < f.height)
	{
		PixelOut po;
		const Outputs out{job.rgb, job.depth, job.iters, job.mask, job.normal, job.steps, job.colour, job.opacity,
			job.world, job.specular, job.shadow, job.gi};
		const size_t pix = (size_t)ys * (size_t)f.width + (size_t)xs;
		render_pixel<KIND, VAR>(f, xs, ys, po, w, out, pix);
		write_pixel(out, pix, po);
	}
	if (counters)
	{
		unsigned long long *slot = counters + (blockIdx.x % kCounterSlots) * kCounterStride;
		wave_add_u64(&slot[0], w.de_evals);
		wave_add_u64(&slot[1], w.iterations);
		wave_add_u64(&slot[2], w.far_steps);
	}
}

// ---- split render for frames with the full shading and no Monte Carlo sampling -------------
// The full variant's shadow, AO, normal and palette work runs only on hits, but in the one-kernel
// render a wave carries the tile's misses through it as idle lanes, and its registers hold the
// march's state besides the shading's.  The split render marches every pixel in
// split_march_kernel, which shades the misses (background, glow) and appends each hit to its
// frame's list (the point, threshold, step count and the pixel's random stream after the march);
// split_shade_kernel then shades the lists, 64 hits per wave.  Each pixel runs the same operations
// in the same order as in render_pixel, so the planes are bit for bit those of the one-kernel
// render.
template <int KIND>
struct HitRec
{
	uint32_t pix;      // pixel index within the frame
	uint32_t seed;     // the pixel's random stream after the march
	int32_t stepCount; // the march's step buffer count (glow)
	typename PrecOf<KIND>::R thr;
	typename PrecOf<KIND>::P point;
};

// the primary ray of pixel (xs, ys) without Monte Carlo jitter (render_pixel's image point)
template <int KIND>
__device__ __forceinline__ typename PrecOf<KIND>::P primary_dir(const FrameConst &f, int xs, int ys)
{
#pragma clang fp contract(off)
	using Q = PrecOf<KIND>;
	using R = typename Q::R;
	const R aspect = Q::aspect(f);
	const R ox = (Q::invWidth(f) * (R)xs - (R)0.5) * aspect;
	const R oy = (R)0.5 - Q::invHeight(f) * (R)ys;
	return Q::view(f, ox, oy);
}

template <int KIND>
__device__ bool split_march_pixel(const FrameConst &frame, int xs, int ys, Work &w, const Outputs &out, size_t oi,
	HitRec<KIND> &rec)
{
#pragma clang fp contract(off)
	const FrameConst &f = uniform_ref(frame);
	using Q = PrecOf<KIND>;
	using R = typename Q::R;
	using P = typename Q::P;
	const uint32_t pix = (uint32_t)ys * (uint32_t)f.width + (uint32_t)xs;
	uint32_t seed = pixel_seed(pix, f.randomSeed);
	const P dir = primary_dir<KIND>(f, xs, ys);
	const P cam = Q::cam(f);
	MarchOut<KIND> m;
	if constexpr (PosPrecision<KIND>::kFP64)
		ray_marching<KIND, 2>(f, cam, dir, (R)f.viewDistanceMax, seed, m, w);
	else if (f.advancedQuality)
		ray_marching<KIND, 1>(f, cam, dir, (R)f.viewDistanceMax, seed, m, w);
	else
		ray_marching<KIND, 0>(f, cam, dir, (R)f.viewDistanceMax, seed, m, w);
	if (out.world)
	{
		const V3 pt = Q::tof(m.point);
		out.world[3 * oi + 0] = pt.x;
		out.world[3 * oi + 1] = pt.y;
		out.world[3 * oi + 2] = pt.z;
	}
	if (out.gi)
		for (int k = 0; k < 3; k++) out.gi[3 * oi + k] = 0.0f;
	if (out.depth) out.depth[oi] = m.found ? (float)m.depth : 1e20f;
	if (out.iters) out.iters[oi] = m.iters;
	if (out.mask) out.mask[oi] = (uint8_t)m.found;
	if (out.steps) out.steps[oi] = m.steps;
	if (out.opacity) out.opacity[oi] = 0;
	if (m.found)
	{
		rec.pix = pix;
		rec.seed = seed;
		rec.stepCount = m.stepCount;
		rec.thr = m.distThresh;
		rec.point = m.point;
		return true;
	}
	V3 rgb = v3(0.0f, 0.0f, 0.0f) + background_shader(f, Q::tof(dir));
	if (f.glow) rgb = glow_shader(f, m.stepCount, rgb);
	if (out.rgb)
	{
		out.rgb[3 * oi + 0] = rgb.x;
		out.rgb[3 * oi + 1] = rgb.y;
		out.rgb[3 * oi + 2] = rgb.z;
	}
	for (int k = 0; k < 3; k++)
	{
		if (out.normal) out.normal[3 * oi + k] = 0.0f;
		if (out.colour) out.colour[3 * oi + k] = 0;
		if (out.specular) out.specular[3 * oi + k] = 0.0f;
		if (out.shadow) out.shadow[3 * oi + k] = 1.0f;
	}
	return false;
}

template <int KIND>
__device__ void split_shade_hit(const FrameConst &frame, const HitRec<KIND> &rec, Work &w, const Outputs &out)
{
#pragma clang fp contract(off)
	const FrameConst &f = uniform_ref(frame);
	using Q = PrecOf<KIND>;
	using P = typename Q::P;
	using R = typename Q::R;
	const int xs = (int)(rec.pix % (uint32_t)f.width), ys = (int)(rec.pix / (uint32_t)f.width);
	const size_t oi = rec.pix;
	uint32_t seed = rec.seed;
	const P dir = primary_dir<KIND>(f, xs, ys);
	const R delta = calc_delta<KIND>(f, rec.point);
	P nP;
	if (f.slowShading)
		nP = slow_normals<KIND>(f, rec.point, (KIND == kGeneric && f.interiorMode) ? (double)rec.thr * 0.2 : (double)delta, w);
	else
		nP = calculate_normals<KIND>(f, rec.point, rec.thr, w);
	const V3 n = Q::tof(nP);
	const V3 colour = surface_colour<KIND>(f, rec.point, w);
	V3 spec;
	float shadowCh = 1.0f;
	V3 rgb = object_shader<KIND, true>(f, rec.point, nP, rec.thr, delta, Q::tof(dir), colour, spec, seed, w, &shadowCh);
	rgb = rgb + spec;
	rgb = v3(fmaxf(rgb.x, 0.0f), fmaxf(rgb.y, 0.0f), fmaxf(rgb.z, 0.0f));
	if (f.glow) rgb = glow_shader(f, rec.stepCount, rgb);
	rgb = v3(0.0f, 0.0f, 0.0f) + rgb;
	if (out.rgb)
	{
		out.rgb[3 * oi + 0] = rgb.x;
		out.rgb[3 * oi + 1] = rgb.y;
		out.rgb[3 * oi + 2] = rgb.z;
	}
	if (out.normal)
	{
		out.normal[3 * oi + 0] = n.x;
		out.normal[3 * oi + 1] = n.y;
		out.normal[3 * oi + 2] = n.z;
	}
	if (out.colour)
	{
		out.colour[3 * oi + 0] = (uint8_t)colour_byte(colour.x);
		out.colour[3 * oi + 1] = (uint8_t)colour_byte(colour.y);
		out.colour[3 * oi + 2] = (uint8_t)colour_byte(colour.z);
	}
	if (out.specular)
	{
		out.specular[3 * oi + 0] = spec.x;
		out.specular[3 * oi + 1] = spec.y;
		out.specular[3 * oi + 2] = spec.z;
	}
	if (out.shadow)
		for (int k = 0; k < 3; k++) out.shadow[3 * oi + k] = shadowCh;
}

// The frame of global tile t: a binary search over the frames' tile ranges (wave-uniform)
__device__ __forceinline__ int frame_of_tile(const FrameJob *__restrict__ jobs, int numFrames, unsigned int t)
{
	int lo = 0, hi = numFrames - 1;
	while (lo < hi)
	{
		const int mid = (lo + hi) >> 1;
		if (t < (unsigned int)jobs[mid].tileEnd)
			hi = mid;
		else
			lo = mid + 1;
	}
	return lo;
}

template <int KIND>
__global__ __launch_bounds__(64, occupancy_target<KIND>()) void split_march_kernel(
	const FrameConst *__restrict__ frames, const FrameJob *__restrict__ jobs, int numFrames,
	unsigned long long *__restrict__ counters, int tileShift, unsigned long long *__restrict__ clearNext,
	unsigned char *__restrict__ hits, unsigned int *__restrict__ hitCounts)
{
	const int lane = threadIdx.x;
	const unsigned int t = blockIdx.x;
	if (clearNext && t == 0)
		for (unsigned int k = 0; k < kCounters; k++) clearNext[lane * kCounterStride + k] = 0ull;
	const int fr = frame_of_tile(jobs, numFrames, t);
	const FrameConst &f = frames[fr];
	const FrameJob &job = jobs[fr];
	Work w = make_work(f);
	const unsigned int tileBegin = fr == 0 ? 0u : (unsigned int)jobs[fr - 1].tileEnd;
	const int lt = (int)(t - tileBegin);
	const int tilesX = (f.width + (1 << tileShift) - 1) >> tileShift;
	const int xs = ((lt % tilesX) << tileShift) + (lane & ((1 << tileShift) - 1));
	const int ys = (lt / tilesX) * (64 >> tileShift) + (lane >> tileShift);
	static_assert(sizeof(HitRec<KIND>) <= kMaxHitRecordBytes, "hit record larger than its allocation");
	HitRec<KIND> rec;
	bool hit = false;
	if (xs < f.width && ys < f.height)
	{
		const Outputs out{job.rgb, job.depth, job.iters, job.mask, job.normal, job.steps, job.colour, job.opacity,
			job.world, job.specular, job.shadow, job.gi};
		hit = split_march_pixel<KIND>(f, xs, ys, w, out, (size_t)ys * (size_t)f.width + (size_t)xs, rec);
	}
	// one atomic per wave appends the wave's hits to the frame's list, in lane order
	const unsigned long long ballot = __ballot(hit);
	if (ballot)
	{
		unsigned int base = 0;
		if (lane == 0) base = atomicAdd(&hitCounts[fr], (unsigned int)__popcll(ballot));
		base = __shfl(base, 0, 64);
		if (hit)
		{
			const unsigned int slot = base + (unsigned int)__popcll(ballot & ((1ull << lane) - 1ull));
			reinterpret_cast<HitRec<KIND> *>(hits)[(size_t)job.pixBase + slot] = rec;
		}
	}
	if (counters)
	{
		unsigned long long *slot = counters + (blockIdx.x % kCounterSlots) * kCounterStride;
		wave_add_u64(&slot[0], w.de_evals);
		wave_add_u64(&slot[1], w.iterations);
		wave_add_u64(&slot[2], w.far_steps);
	}
}

template <int KIND>
__global__ __launch_bounds__(64, occupancy_target<KIND>()) void split_shade_kernel(
	const FrameConst *__restrict__ frames, const FrameJob *__restrict__ jobs, int numFrames,
	unsigned long long *__restrict__ counters, const unsigned char *__restrict__ hits,
	const unsigned int *__restrict__ hitCounts)
{
	const int lane = threadIdx.x;
	const unsigned int t = blockIdx.x;
	const int fr = frame_of_tile(jobs, numFrames, t);
	const unsigned int tileBegin = fr == 0 ? 0u : (unsigned int)jobs[fr - 1].tileEnd;
	const unsigned int first = (t - tileBegin) * 64u;
	const unsigned int count = hitCounts[fr];
	if (first >= count) return; // a whole wave past the frame's hits
	const FrameConst &f = frames[fr];
	const FrameJob &job = jobs[fr];
	Work w = make_work(f);
	if (first + (unsigned int)lane < count)
	{
		const HitRec<KIND> rec = reinterpret_cast<const HitRec<KIND> *>(hits)[(size_t)job.pixBase + first + lane];
		const Outputs out{job.rgb, job.depth, job.iters, job.mask, job.normal, job.steps, job.colour, job.opacity,
			job.world, job.specular, job.shadow, job.gi};
		split_shade_hit<KIND>(f, rec, w, out);
	}
	if (counters)
	{
		unsigned long long *slot = counters + (blockIdx.x % kCounterSlots) * kCounterStride;
		wave_add_u64(&slot[0], w.de_evals);
		wave_add_u64(&slot[1], w.iterations);
		wave_add_u64(&slot[2], w.far_steps);
	}
}

// Zero the split render's per-frame hit counts
__global__ void hit_clear_kernel(unsigned int *counts, int n)
{
	const int i = blockIdx.x * 64 + threadIdx.x;
	if (i < n) counts[i] = 0u;
}

// Persistent work-stealing kernel: a fixed grid of waves drains the 64-pixel tiles of every
// frame of a batch (frames of different resolution, camera and formula parameters).  The
// wave-uniform tile index keeps each tile's frame constants in scalar registers.  One launch
// replaces the reference's one-job-per-frame animation loop, and the fixed grid avoids
// dispatching one workgroup per tile.
// The tiles are split into kTileQueues contiguous ranges, each with its own counter on its own
// 256-byte line: a device-scope atomic on one address serialises at about 8 ns per operation,
// which 4096 waves asking for a tile every ~25 us would saturate.  A wave starts on queue
// (block mod kTileQueues) and, when that range is exhausted, moves on to the next queue.
// (frames and jobs are separate __restrict__ kernel arguments so that the compiler may read
// them with scalar loads)
template <int KIND, int VAR>
__global__ __launch_bounds__(64, occupancy_target<KIND>()) void persistent_kernel(const FrameConst *__restrict__ frames,
	const FrameJob *__restrict__ jobs, int numFrames, unsigned int *__restrict__ queue,
	unsigned long long *__restrict__ counters, int tileShift)
{
	const int lane = threadIdx.x;
	const unsigned int totalTiles = (unsigned int)jobs[numFrames - 1].tileEnd;
	Work w = make_work(frames[0]);
	int fr = 0;
	const unsigned int home = blockIdx.x % kTileQueues;
	for (unsigned int k = 0; k < kTileQueues; k++)
	{
		const unsigned int q = (home + k) % kTileQueues;
		const unsigned int begin = (unsigned int)((unsigned long long)totalTiles * q / kTileQueues);
		const unsigned int count = (unsigned int)((unsigned long long)totalTiles * (q + 1) / kTileQueues) - begin;
		unsigned int *head = queue + q * kTileQueueStride;
		// a plain load never reads more than the counter holds: skip ranges already exhausted
		if (__builtin_amdgcn_readfirstlane(__atomic_load_n(head, __ATOMIC_RELAXED)) >= count) continue;
		while (true)
		{
			unsigned int c = 0;
			if (lane == 0) c = atomicAdd(head, 1u);
			c = __builtin_amdgcn_readfirstlane(c);
			if (c >= count) break;
			const unsigned int t = begin + c;
			if (fr > 0 && t < (unsigned int)jobs[fr - 1].tileEnd) fr = 0;
			while (t >= (unsigned int)jobs[fr].tileEnd) fr++;
			const FrameConst &f = frames[fr];
			const FrameJob &job = jobs[fr];
			load_orbit_constants(w, f);
			const unsigned int tileBegin = fr == 0 ? 0u : (unsigned int)jobs[fr - 1].tileEnd;
			const int lt = (int)(t - tileBegin);
			const int tilesX = (f.width + (1 << tileShift) - 1) >> tileShift;
			const int xs = ((lt % tilesX) << tileShift) + (lane & ((1 << tileShift) - 1));
			const int ys = (lt / tilesX) * (64 >> tileShift) + (lane >> tileShift);
			if (xs < f.width && ys < f.height)
			{
				PixelOut po;
				const Outputs out{job.rgb, job.depth, job.iters, job.mask, job.normal, job.steps, job.colour,
					job.opacity, job.world, job.specular, job.shadow, job.gi};
				const size_t pix = (size_t)ys * (size_t)f.width + (size_t)xs;
				render_pixel<KIND, VAR>(f, xs, ys, po, w, out, pix);
				write_pixel(out, pix, po);
			}
		}
	}
	if (counters)
	{
		unsigned long long *slot = counters + (blockIdx.x % kCounterSlots) * kCounterStride;
		wave_add_u64(&slot[0], w.de_evals);
		wave_add_u64(&slot[1], w.iterations);
		wave_add_u64(&slot[2], w.far_steps);
	}
}

// Static scratch of mbx_render_batch_noalloc (layout in mbx_launch.h): every call on a GPU
// shares this GPU's copy, which keeps the last call's contents.  Host code never passes the
// array's name to a kernel: the name is a host-side stand-in, and the device address comes from
// hipGetSymbolAddress (noalloc_scratch_address) once per device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch_address(int device, void **ptr)
{
	constexpr int kMaxDevices = 64;
	static std::mutex lock;
	static void *address[kMaxDevices] = {};
	*ptr = nullptr;
	if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
	std::lock_guard<std::mutex> guard(lock);
	if (!address[device])
	{
		void *p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess) return e;
		address[device] = p;
	}
	*ptr = address[device];
	return hipSuccess;
}

// Zero the tile queue heads and the work counters (a kernel, so that it can be captured in a
// graph without a memset node).
__global__ void clear_kernel(unsigned int *queue, unsigned long long *counters)
{
	if (threadIdx.x < kTileQueues) queue[threadIdx.x * kTileQueueStride] = 0u;
	if (threadIdx.x < kCounterSlots)
		for (unsigned int k = 0; k < kCounters; k++) counters[threadIdx.x * kCounterStride + k] = 0ull;
}

// Point DE query (distance engine analogue)
template <int KIND>
__global__ __launch_bounds__(64, KIND == kGeneric ? 3 : 1) void distance_kernel(const FrameConst *__restrict__ frame,
	const float *__restrict__ pts, int n, float detailSize, float *__restrict__ dist,
	int32_t *__restrict__ iters)
{
	const int i = blockIdx.x * 64 + threadIdx.x;
	if (i >= n) return;
	Work w = make_work(*frame);
	int it = 0;
	using Q = PrecOf<KIND>;
	const typename Q::P p = Q::from(v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
	dist[i] = (float)calculate_distance<KIND>(p, (typename Q::R)detailSize, *frame, it, w);
	iters[i] = it;
}

template <int KIND>
static hipError_t launch_render_t(const FrameConst *frame, int width, int height, const Outputs &out,
	unsigned long long *counters, int tileShift, int variant, hipStream_t stream)
{
	const int tw = 1 << tileShift, th = 64 >> tileShift;
	const int tiles = ((width + tw - 1) / tw) * ((height + th - 1) / th);
	if (variant == kVariantLens)
		hipLaunchKernelGGL((render_kernel<KIND, kVariantLens>), dim3(tiles), dim3(64), 0, stream, frame, out,
			counters, tileShift);
	else if (variant == kVariantFull)
		hipLaunchKernelGGL((render_kernel<KIND, kVariantFull>), dim3(tiles), dim3(64), 0, stream, frame, out,
			counters, tileShift);
	else
		hipLaunchKernelGGL((render_kernel<KIND, kVariantLite>), dim3(tiles), dim3(64), 0, stream, frame, out,
			counters, tileShift);
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
	const Outputs &out, unsigned long long *counters, int tileShift, int variant, hipStream_t stream)
{
	switch (deKind)
	{
		case kBulb8: return launch_render_t<kBulb8>(frame, width, height, out, counters, tileShift, variant, stream);
		case kBulbTrig: return launch_render_t<kBulbTrig>(frame, width, height, out, counters, tileShift, variant, stream);
		case kBoxAnalytic:
			return launch_render_t<kBoxAnalytic>(frame, width, height, out, counters, tileShift, variant, stream);
		case kBoxDelta: return launch_render_t<kBoxDelta>(frame, width, height, out, counters, tileShift, variant, stream);
		case kIfsAnalytic:
			return launch_render_t<kIfsAnalytic>(frame, width, height, out, counters, tileShift, variant, stream);
		case kBulbD: return launch_render_t<kBulbD>(frame, width, height, out, counters, tileShift, variant, stream);
		case kGeneric: return launch_render_t<kGeneric>(frame, width, height, out, counters, tileShift, variant, stream);
		default: return hipErrorInvalidValue;
	}
}

hipError_t launch_clear(unsigned int *queue, unsigned long long *counters, hipStream_t stream)
{
	hipLaunchKernelGGL(clear_kernel, dim3(1), dim3(64), 0, stream, queue, counters);
	return hipGetLastError();
}

template <int KIND, int VAR>
static void launch_batch_v(const BatchArgs &a, int numWaves, hipStream_t stream)
{
	if (a.mode == kBatchPersistent)
		hipLaunchKernelGGL((persistent_kernel<KIND, VAR>), dim3(numWaves), dim3(64), 0, stream, a.frames, a.jobs,
			a.numFrames, a.queue, a.counters, a.tileShift);
	else
		hipLaunchKernelGGL((batch_kernel<KIND, VAR>), dim3(a.totalTiles), dim3(64), 0, stream, a.frames, a.jobs,
			a.numFrames, a.counters, a.tileShift, a.clearNext);
}

template <int KIND>
static hipError_t launch_batch_t(const BatchArgs &a, int numWaves, int variant, hipStream_t stream)
{
	if (a.hits && a.mode == kBatchGrid && variant == kVariantFull)
	{
		hipLaunchKernelGGL(hit_clear_kernel, dim3((a.numFrames + 63) / 64), dim3(64), 0, stream, a.hitCounts,
			a.numFrames);
		hipLaunchKernelGGL((split_march_kernel<KIND>), dim3(a.totalTiles), dim3(64), 0, stream, a.frames, a.jobs,
			a.numFrames, a.counters, a.tileShift, a.clearNext, a.hits, a.hitCounts);
		hipLaunchKernelGGL((split_shade_kernel<KIND>), dim3(a.totalTiles), dim3(64), 0, stream, a.frames, a.jobs,
			a.numFrames, a.counters, (const unsigned char *)a.hits, (const unsigned int *)a.hitCounts);
		return hipGetLastError();
	}
	if (variant == kVariantLens)
		launch_batch_v<KIND, kVariantLens>(a, numWaves, stream);
	else if (variant == kVariantFull)
		launch_batch_v<KIND, kVariantFull>(a, numWaves, stream);
	else
		launch_batch_v<KIND, kVariantLite>(a, numWaves, stream);
	return hipGetLastError();
}

hipError_t launch_batch(int deKind, BatchArgs a, int numWaves, hipStream_t stream)
{
	if (a.clearFirst)
	{
		hipError_t e = launch_clear(a.queue, a.counters, stream);
		if (e != hipSuccess) return e;
	}
	switch (deKind)
	{
		case kBulb8: return launch_batch_t<kBulb8>(a, numWaves, a.full, stream);
		case kBulbTrig: return launch_batch_t<kBulbTrig>(a, numWaves, a.full, stream);
		case kBoxAnalytic: return launch_batch_t<kBoxAnalytic>(a, numWaves, a.full, stream);
		case kBoxDelta: return launch_batch_t<kBoxDelta>(a, numWaves, a.full, stream);
		case kIfsAnalytic: return launch_batch_t<kIfsAnalytic>(a, numWaves, a.full, stream);
		case kBulbD: return launch_batch_t<kBulbD>(a, numWaves, a.full, stream);
		case kGeneric: return launch_batch_t<kGeneric>(a, numWaves, a.full, stream);
		default: return hipErrorInvalidValue;
	}
}

// One FormulaCode call per point: the table's <Name>Iteration on z and the full sExtendedAux,
// in the layout of mbx_extended_aux
__global__ __launch_bounds__(64) void formula_kernel(const FractalParams *__restrict__ fractal,
	double *__restrict__ zs, mbx_extended_aux *__restrict__ auxs, int n)
{
	const int k = blockIdx.x * 64 + threadIdx.x;
	if (k >= n) return;
	DV4 z = dv4(zs[4 * k], zs[4 * k + 1], zs[4 * k + 2], zs[4 * k + 3]);
	mbx_extended_aux &e = auxs[k];
	ExtendedAux a;
	a.c = dv4(e.c[0], e.c[1], e.c[2], e.c[3]);
	a.const_c = dv4(e.const_c[0], e.const_c[1], e.const_c[2], e.const_c[3]);
	a.old_z = dv4(e.old_z[0], e.old_z[1], e.old_z[2], e.old_z[3]);
	a.pos_neg = e.pos_neg;
	a.r = e.r;
	a.DE = e.DE;
	a.DE0 = e.DE0;
	a.dist = e.dist;
	a.pseudoKleinianDE = e.pseudoKleinianDE;
	a.actualScale = e.actualScale;
	a.actualScaleA = e.actualScaleA;
	a.color = e.color;
	a.colorHybrid = e.colorHybrid;
	a.temp1000 = e.temp1000;
	a.i = e.i;
	formula_iteration(*fractal, z, a);
	zs[4 * k] = z.x;
	zs[4 * k + 1] = z.y;
	zs[4 * k + 2] = z.z;
	zs[4 * k + 3] = z.w;
	const double v4[3][4] = {{a.c.x, a.c.y, a.c.z, a.c.w}, {a.const_c.x, a.const_c.y, a.const_c.z, a.const_c.w},
		{a.old_z.x, a.old_z.y, a.old_z.z, a.old_z.w}};
	for (int j = 0; j < 4; j++)
	{
		e.c[j] = v4[0][j];
		e.const_c[j] = v4[1][j];
		e.old_z[j] = v4[2][j];
	}
	e.pos_neg = a.pos_neg;
	e.r = a.r;
	e.DE = a.DE;
	e.DE0 = a.DE0;
	e.dist = a.dist;
	e.pseudoKleinianDE = a.pseudoKleinianDE;
	e.actualScale = a.actualScale;
	e.actualScaleA = a.actualScaleA;
	e.color = a.color;
	e.colorHybrid = a.colorHybrid;
	e.temp1000 = a.temp1000;
	e.i = a.i;
}

hipError_t launch_formula(const FractalParams *fractal, double *z, mbx_extended_aux *aux, int n,
	hipStream_t stream)
{
	if (n <= 0) return hipSuccess;
	hipLaunchKernelGGL(formula_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, fractal, z, aux, n);
	return hipGetLastError();
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
		case kBulbD: return launch_distance_t<kBulbD>(frame, pts, n, detail, dist, iters, stream);
		case kGeneric: return launch_distance_t<kGeneric>(frame, pts, n, detail, dist, iters, stream);
		default: return hipErrorInvalidValue;
	}
}

} // namespace mbx

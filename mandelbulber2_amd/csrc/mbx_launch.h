// mbx_launch.h - host-callable launchers of the device kernels (mbx_render.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mbx_frame.h"

namespace mbx
{

struct Outputs
{
	float *rgb;
	float *depth;
	int32_t *iters;
	uint8_t *mask;
	float *normal;
	int32_t *steps;
	uint8_t *colour;
	uint16_t *opacity;
	float *world;    // optional channels: hit point, specular, shadow, GI (3 floats per pixel)
	float *specular;
	float *shadow;
	float *gi;
};

// Kernel variants: without Monte Carlo, GI, shadow and AO code; with it; and with it plus the
// Monte Carlo lens and chromatic aberration, whose rays start off the camera (a march with
// per-point thresholds that the other variants do not carry), and the fog walk of the march.
constexpr int kVariantLite = 0;
constexpr int kVariantFull = 1;
constexpr int kVariantLens = 2;

// Plain per-frame launch: one 64-lane workgroup per tile of (1 << tileShift) x (64 >> tileShift)
// pixels, in the given kernel variant.  The work counters are added into `counters`
// (kCounterBytes, slot layout above).
hipError_t launch_render(int deKind, const FrameConst *frame, int width, int height,
	const Outputs &out, unsigned long long *counters, int tileShift, int variant, hipStream_t stream);

// Point DE query at a fixed detail size.
hipError_t launch_distance(int deKind, const FrameConst *frame, const float *pts, int n,
	float detail, float *dist, int32_t *iters, hipStream_t stream);

// One formula iteration per point (device memory): z (4 doubles) and aux updated in place.
hipError_t launch_formula(const FractalParams *fractal, double *z, mbx_extended_aux *aux, int n,
	hipStream_t stream);

// Image post-processing of n pixels (3 floats each) through a 65536-entry gamma table (device);
// out16 / out8 may be null.
hipError_t launch_compile_image(const float *rgb, uint32_t n, const uint16_t *gammaTable, float brightness,
	float contrast, float saturation, int hdr, uint16_t *out16, uint8_t *out8, hipStream_t stream);

// z-buffer grey image of n depths; `range` (2 uint32, device) receives the frame's depth range
// when useRange is set.
hipError_t launch_zbuffer(const float *depth, uint32_t n, uint32_t *range, int useRange, float minDepth,
	float maxDepth, int bits, int logarithmic, int invert, void *out, hipStream_t stream);

// Screen-space AO over a frame's planes (passed by value: the angle table rides in the kernel
// arguments).  quality = max(3, ambient_occlusion_quality^2) <= kSsaoMaxQuality angles.
constexpr int kSsaoMaxQuality = 100;
struct SsaoArgs
{
	const float *depth;
	const uint8_t *colour;
	const uint16_t *opacity; // may be null (opacity 0)
	const float *rgbIn;
	float *rgbOut;
	int width, height;
	int quality;
	int perspective; // enumPerspectiveType: 0 three-point, 1 fish-eye, 2 equirectangular, 3 fish-eye cut
	int randomMode;
	uint32_t seed;   // frame seed of the per-pixel streams (random mode)
	double fov;      // CalcFOV of the perspective
	double aspectRatio;
	double scaleFactor; // width / quality^2 / 2
	float intensity;
	float aoColor[3];
	double cosT[kSsaoMaxQuality]; // cos and sin of 2 pi i / quality, as the host's libm gives them
	double sinT[kSsaoMaxQuality];
};
hipError_t launch_ssao(const SsaoArgs &a, hipStream_t stream);

} // namespace mbx

namespace mbx
{

// One frame of a batch: its tile range in the global tile numbering and its output planes.
struct FrameJob
{
	int tileEnd; // exclusive end of this frame's tiles (64 pixels each)
	int pixBase; // offset of this frame's first pixel in the batch
	float *rgb;
	float *depth;
	int32_t *iters;
	uint8_t *mask;
	float *normal;
	int32_t *steps;
	uint8_t *colour;
	uint16_t *opacity;
	float *world;
	float *specular;
	float *shadow;
	float *gi;
};

// Work counters: kCounterSlots slots of kCounterStride u64, each slot on its own 256-byte line;
// a workgroup adds into slot (block mod kCounterSlots) so that the atomics of thousands of
// workgroups do not queue on one address.  Counter k of the launch = sum over the slots of
// slot[k]: 0 DE evaluations, 1 formula iterations, 2 far-field steps.
constexpr unsigned int kCounterSlots = 64;
constexpr unsigned int kCounterStride = 32;
constexpr unsigned int kCounters = 3;
constexpr size_t kCounterBytes = (size_t)kCounterSlots * kCounterStride * sizeof(unsigned long long);

// Tile queues of the persistent kernel: kTileQueues counters, kTileQueueStride uints apart.
constexpr unsigned int kTileQueues = 32;
constexpr unsigned int kTileQueueStride = 64;
constexpr size_t kTileQueueBytes = (size_t)kTileQueues * kTileQueueStride * sizeof(unsigned int);

// Arguments of the persistent kernel (passed by value).
struct BatchArgs
{
	const FrameConst *frames;
	const FrameJob *jobs;
	int numFrames;
	int tileShift; // tiles are (1 << tileShift) x (64 >> tileShift) pixels, tileShift in 2..5
	int full;      // kernel variant (kVariantLite, kVariantFull, kVariantLens) that every frame fits
	int mode;      // kBatchGrid or kBatchPersistent
	unsigned int totalTiles;
	unsigned int *queue;           // kTileQueues tile counters (kTileQueueBytes)
	unsigned long long *counters;  // work counter slots (kCounterBytes)
	int clearFirst;                // launch the clear kernel (queue + counters) before the batch
	unsigned long long *clearNext; // grid mode: counter slots workgroup 0 zeroes for the next launch
	unsigned char *hits;           // grid mode, full variant without Monte Carlo: the split render's
	                               // hit records (kMaxHitRecordBytes per pixel of the batch), or null
	unsigned int *hitCounts;       // its per-frame hit counts (numFrames)
};

// Upper bound of the split render's hit record (the FP64 kinds': index, seed, step count,
// threshold and point)
constexpr size_t kMaxHitRecordBytes = 48;

// Batch launch modes: one workgroup per tile of every frame (grid), or a fixed grid of
// `numWaves` waves draining tile queues (persistent).
constexpr int kBatchGrid = 0;
constexpr int kBatchPersistent = 2;

// Enqueue the batch: a clear of the queues and counters, then the batch kernel of a.mode.
hipError_t launch_batch(int deKind, BatchArgs a, int numWaves, hipStream_t stream);

// Zero the tile counters and the work counters.
hipError_t launch_clear(unsigned int *queue, unsigned long long *counters, hipStream_t stream);

} // namespace mbx

namespace mbx
{

// Static scratch of the allocation-free batch entry point (mbx_render_batch_noalloc): one
// file-scope __device__ array per GPU, laid out as the work-counter slots, the tile-queue heads,
// the job table and the frame constants of at most kNoAllocMaxFrames frames.
constexpr int kNoAllocMaxFrames = MBX_NOALLOC_MAX_FRAMES;
constexpr size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }
constexpr size_t kNoAllocCountersOff = 0;
constexpr size_t kNoAllocQueueOff = align256(kNoAllocCountersOff + kCounterBytes);
constexpr size_t kNoAllocJobsOff = align256(kNoAllocQueueOff + kTileQueueBytes);
constexpr size_t kNoAllocFramesOff = align256(kNoAllocJobsOff + (size_t)kNoAllocMaxFrames * sizeof(FrameJob));
constexpr size_t kNoAllocScratchBytes = align256(kNoAllocFramesOff + (size_t)kNoAllocMaxFrames * sizeof(FrameConst));

// Device address of the static scratch on `device`, which the caller has made current: taken
// with hipGetSymbolAddress on the first call for the device and kept per device.  Host code uses
// this address, never the array's name, for kernel arguments and copies.
hipError_t noalloc_scratch_address(int device, void **ptr);

} // namespace mbx

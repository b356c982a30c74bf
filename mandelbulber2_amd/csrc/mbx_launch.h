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
};

// Plain per-frame launch: one 64-lane workgroup per tile of (1 << tileShift) x (64 >> tileShift)
// pixels.  counters[0] receives the
// DE evaluations, counters[1] the formula iterations (atomically added).
hipError_t launch_render(int deKind, const FrameConst *frame, int width, int height,
	const Outputs &out, unsigned long long *counters, int tileShift, hipStream_t stream);

// Point DE query at a fixed detail size.
hipError_t launch_distance(int deKind, const FrameConst *frame, const float *pts, int n,
	float detail, float *dist, int32_t *iters, hipStream_t stream);

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
};

// Arguments of the persistent kernel (passed by value).
struct BatchArgs
{
	const FrameConst *frames;
	const FrameJob *jobs;
	int numFrames;
	int tileShift; // tiles are (1 << tileShift) x (64 >> tileShift) pixels, tileShift in 2..5
	unsigned int *queue;           // [0] tile queue head
	unsigned long long *counters;  // [0] DE evaluations, [1] formula iterations
};

// Enqueue the batch: a clear of the queue head and counters, then the persistent kernel with
// `numWaves` 64-lane waves.
hipError_t launch_batch(int deKind, BatchArgs a, int numWaves, hipStream_t stream);

// Zero queue[0] and the work counters.
hipError_t launch_clear(unsigned int *queue, unsigned long long *counters, hipStream_t stream);

} // namespace mbx

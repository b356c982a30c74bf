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

// Plain per-frame launch: one 64-lane workgroup per 8x8 pixel tile.  counters[0] receives the
// DE evaluations, counters[1] the formula iterations (atomically added).
hipError_t launch_render(int deKind, const FrameConst *frame, int width, int height,
	const Outputs &out, unsigned long long *counters, hipStream_t stream);

// Point DE query at a fixed detail size.
hipError_t launch_distance(int deKind, const FrameConst *frame, const float *pts, int n,
	float detail, float *dist, int32_t *iters, hipStream_t stream);

} // namespace mbx

namespace mbx
{

// Feature set compiled into a persistent-kernel instance.
enum Feature : int
{
	kFeatShadow = 1,
	kFeatAO = 2,
	kFeatMC = 4, // Monte Carlo sub-pixel samples and the GI bounce loop
};

// One frame of a batch: its tile range in the global tile numbering and its output planes.
struct FrameJob
{
	int tileEnd; // exclusive end of this frame's tiles (8x8 pixels each)
	int pad;
	float *rgb;
	float *depth;
	int32_t *iters;
	uint8_t *mask;
	float *normal;
	int32_t *steps;
};

// Persistent work-stealing launch over all tiles of `numFrames` frames.  queue[0] must be 0
// and counters zeroed before the launch (launch_clear).
hipError_t launch_persistent(int deKind, int feat, const FrameConst *frames, const FrameJob *jobs,
	int numFrames, unsigned int *queue, unsigned long long *counters, int numWaves, hipStream_t stream);

hipError_t launch_clear(unsigned int *queue, unsigned long long *counters, hipStream_t stream);

} // namespace mbx

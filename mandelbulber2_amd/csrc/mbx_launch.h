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

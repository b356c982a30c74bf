// mbx_diff.h - constants of the differentiable render and its adjoint, shared by the host
// (mbx_host.cpp derives them in FP64) and the kernels (mbx_diff.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mbx.h"
#include "mbx_frame.h"

namespace mbx
{

constexpr int kDiffParams = MBX_DIFF_NPARAMS;

struct DiffConst
{
	// formula: Mandelbulb (fractal_mandelbulb.cpp:29-43), analytic logarithmic DE, c = p * cm
	double power, alpha, beta, bailout;
	double cm[3];
	int32_t maxN;
	int32_t constThr;   // distThresh = DEThresh, else t |dir| threshK
	double threshK, DEThresh;
	// node rotation (world -> local) and its derivatives per degree of rotation x, y, z
	double R[9];
	double dR[3][9];
	// camera position, rotation and the rotation's derivatives per camera x, y, z, target x, y, z
	double cam[3];
	double M[9];
	double dM[6][9];
	double fov, aspect, invW, invH;
	int32_t width, height;
	// Lambert with one directional light (shader_object.cpp, shader_light_shading.cpp)
	int32_t lightEnabled, bg3;
	double light[3];
	double lightColor[3];
	double intensity, shading;
	double colour[3], fill[3], lum[3];
	// background (shader_background.cpp), gamma 1: 3-colour gradient and the light's disk
	double bg1[3], bg2[3], bgc3[3];
	double bgBrightness, lightSize, sunPow, visibility;
};

struct RayCheckpoint
{
	double t;
	int32_t found, refined, iters, pad;
};
static_assert(sizeof(RayCheckpoint) == sizeof(mbx_ray_checkpoint), "checkpoint layout");

// Differentiable render of one frame: the kind's march, FP64 Newton refinement, analytic normal,
// FP64 Lambert shading; one 64-lane workgroup per 8x8 tile.
hipError_t launch_render_diff(int deKind, const FrameConst *frame, const DiffConst *dc, int width,
	int height, float *rgb, float *depth, int32_t *iters, uint8_t *mask, float *normal,
	RayCheckpoint *ckpt, hipStream_t stream);

// Adjoint: accumulates dL/dtheta into grad (kDiffParams doubles, zeroed by the launcher with a
// kernel) and optionally writes per-pixel contributions.
hipError_t launch_adjoint(const FrameConst *frame, const DiffConst *dc, int width, int height,
	const RayCheckpoint *ckpt, const float *dLdI, double *grad, double *pixelGrad, hipStream_t stream);

} // namespace mbx

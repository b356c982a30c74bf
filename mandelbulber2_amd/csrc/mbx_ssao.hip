// mbx_ssao.hip - screen-space ambient occlusion, the post effect the reference runs over the
// finished z-buffer when ambient_occlusion_mode is screen space (render_image.cpp:498-508,
// cRenderSSAO::RenderSSAO render_ssao.cpp:76-200, per pixel cSSAOWorker::doWork
// ssao_worker.cpp:63-261).
//
// One lane per pixel, 16x16-pixel workgroups: for a given angle and radius the 64 lanes of a
// wave read a 16x4 block of the depth plane displaced by the same offset, so the gathers of a
// wave stay within a few cache lines.  FP64 arithmetic in the reference's operation order with
// contraction off (the sample position decides which depth is read, so it must round as the
// reference's does).  A direction that leaves the image never comes back (the sample position
// moves monotonically outwards with the radius), so the radius loop stops there instead of
// testing the remaining radii.
#include <hip/hip_runtime.h>

#include "mbx_device.h"
#include "mbx_launch.h"

namespace mbx
{

__device__ __forceinline__ void ssao_project(const SsaoArgs &a, double px, double py, double z, double &ox, double &oy)
{
#pragma clang fp contract(off)
	const double W = (double)a.width, H = (double)a.height;
	if (a.perspective == 1 || a.perspective == 3)
	{
		// fish-eye: the sample points carry the factor pi that the pixel's own point does not
		// (ssao_worker.cpp:126-136 against :193-204), and the sample's test is on the radius
		// loop's r, which is never 0; this helper serves the samples
		double xx2 = M_PI * (px / W - 0.5) * a.aspectRatio;
		double yy2 = M_PI * (py / H - 0.5);
		const double r2 = sqrt(xx2 * xx2 + yy2 * yy2);
		xx2 = xx2 / r2 * sin(r2 * a.fov) * z;
		yy2 = yy2 / r2 * sin(r2 * a.fov) * z;
		ox = xx2;
		oy = yy2;
	}
	else if (a.perspective == 2)
	{
		double xx2 = M_PI * (px / W - 0.5) * a.aspectRatio;
		double yy2 = M_PI * (py / H - 0.5);
		ox = sin(a.fov * xx2) * cos(a.fov * yy2) * z;
		oy = sin(a.fov * yy2) * z;
	}
	else
	{
		const double xx2 = (px / W - 0.5) * a.aspectRatio;
		const double yy2 = py / H - 0.5;
		ox = xx2 * (z * a.fov);
		oy = yy2 * (z * a.fov);
	}
}

__global__ __launch_bounds__(256) void ssao_kernel(const SsaoArgs a)
{
#pragma clang fp contract(off)
	const int x = blockIdx.x * 16 + (threadIdx.x & 15);
	const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
	if (x >= a.width || y >= a.height) return;
	const size_t i = (size_t)y * (size_t)a.width + (size_t)x;
	const double z = (double)a.depth[i];
	const float opacity = a.opacity ? (float)a.opacity[i] / 65535.0f : 0.0f;
	float totalAmbient = 0.0f;
	if (z < 1e19)
	{
		const double W = (double)a.width, H = (double)a.height;
		double x2, y2;
		if (a.perspective == 1 || a.perspective == 3)
		{
			x2 = ((double)x / W - 0.5) * a.aspectRatio;
			y2 = ((double)y / H - 0.5);
			const double r = sqrt(x2 * x2 + y2 * y2);
			if (r != 0.0)
			{
				x2 = x2 / r * sin(r * a.fov) * z;
				y2 = y2 / r * sin(r * a.fov) * z;
			}
		}
		else if (a.perspective == 2)
		{
			x2 = M_PI * ((double)x / W - 0.5) * a.aspectRatio;
			y2 = M_PI * ((double)y / H - 0.5);
			x2 = sin(a.fov * x2) * cos(a.fov * y2) * z;
			y2 = sin(a.fov * y2) * z;
		}
		else
		{
			x2 = ((double)x / W - 0.5) * a.aspectRatio;
			y2 = (double)y / H - 0.5;
			x2 = x2 * z * a.fov;
			y2 = y2 * z * a.fov;
		}
		float ambient = 0.0f;
		const int quality = a.quality;
		const double angleStep = M_PI * 2.0 / (double)quality;
		const int maxRandom = 62831 / quality;
		double rRandom = 1.0;
		uint32_t seed = 0;
		if (a.randomMode)
		{
			seed = pixel_seed((uint32_t)i, a.seed);
			rRandom = 0.5 + (double)random_mod(seed, 65536) / 65536.0;
		}
		int rayCount = 0;
		const double xMax = (double)(a.width - 1), yMax = (double)(a.height - 1);
		for (int angleIndex = 0; angleIndex < quality; angleIndex++)
		{
			double ca, sa;
			if (a.randomMode)
			{
				const double angle = angleStep * (double)angleIndex + (double)random_mod(seed, maxRandom) / 10000.0;
				ca = cos(angle);
				sa = sin(angle);
			}
			else
			{
				ca = a.cosT[angleIndex];
				sa = a.sinT[angleIndex];
			}
			double maxDiff = -1e50;
			bool wasRay = false;
			for (double r = 1.0; r < (double)quality; r += rRandom)
			{
				const double rr = r * r * a.scaleFactor;
				const double xx = (double)x + rr * ca;
				const double yy = (double)y + rr * sa;
				if ((int)xx == x && (int)yy == y) continue;
				if (xx < 0.0 || xx > xMax || yy < 0.0 || yy > yMax) break; // outside for every larger r
				const double z2 = (double)a.depth[(size_t)(int)yy * (size_t)a.width + (size_t)(int)xx];
				wasRay = true;
				double xx2, yy2;
				ssao_project(a, xx, yy, z2, xx2, yy2);
				const double dx = xx2 - x2;
				const double dy = yy2 - y2;
				const double dz = z2 - z;
				const double dr = sqrt(dx * dx + dy * dy);
				const double diff = -dz / dr;
				if (diff > maxDiff) maxDiff = diff;
			}
			const double maxAngle = atan(maxDiff);
			if (wasRay)
			{
				ambient = (float)((double)ambient + (-maxAngle / M_PI + 0.5));
				rayCount++;
			}
		}
		totalAmbient = ambient / (float)rayCount;
		if (totalAmbient < 0.0f) totalAmbient = 0.0f;
	}
	float r = a.rgbIn[3 * i + 0], g = a.rgbIn[3 * i + 1], b = a.rgbIn[3 * i + 2];
	// the worker's write loop stops short of the last column (ssao_worker.cpp:241-243)
	if (x < a.width - 1)
	{
		const float shadeFactor = 1.0f / 256.0f * totalAmbient * a.intensity * (1.0f - opacity);
		r = r + (float)a.colour[3 * i + 0] * shadeFactor * a.aoColor[0];
		g = g + (float)a.colour[3 * i + 1] * shadeFactor * a.aoColor[1];
		b = b + (float)a.colour[3 * i + 2] * shadeFactor * a.aoColor[2];
	}
	a.rgbOut[3 * i + 0] = r;
	a.rgbOut[3 * i + 1] = g;
	a.rgbOut[3 * i + 2] = b;
}

hipError_t launch_ssao(const SsaoArgs &a, hipStream_t stream)
{
	const dim3 grid((unsigned)((a.width + 15) / 16), (unsigned)((a.height + 15) / 16));
	hipLaunchKernelGGL(ssao_kernel, grid, dim3(256), 0, stream, a);
	return hipGetLastError();
}

} // namespace mbx

// mbx_image.hip - the image post-processing kernel: cImage::CalculatePixel over every pixel of the
// float image (mandelbulber2/src/cimage.cpp:317-361, 378-397) and the 8-bit conversion of
// cImage::ConvertTo8bitChar (cimage.cpp:460-470), in one pass.  HBM-bound: 12 bytes read and
// 6 + 3 bytes written per pixel; the 128 KiB gamma table stays in L2.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "mbx_launch.h"

namespace mbx
{

__global__ __launch_bounds__(256) void compile_image_kernel(const float *__restrict__ rgb, uint32_t n,
	const uint16_t *__restrict__ gammaTable, float brightness, float contrast, float saturation, int hdr,
	uint16_t *__restrict__ out16, uint8_t *__restrict__ out8)
{
#pragma clang fp contract(off)
	const uint32_t stride = gridDim.x * blockDim.x;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		const size_t k = 3 * (size_t)i;
		float c[3] = {rgb[k], rgb[k + 1], rgb[k + 2]};
		for (int j = 0; j < 3; j++)
		{
			c[j] = c[j] * brightness;
			c[j] = (c[j] - 0.5f) * contrast + 0.5f;
			if (c[j] < 0.0f) c[j] = 0.0f;
			if (hdr) c[j] = tanhf(c[j]);
		}
		// saturation about the luma (rFactor 0.299, gFactor 0.587, bFactor 0.114).  sqrtf must be
		// correctly rounded as the host's is; v_sqrt_f32 is not (1 ulp), so the root is taken in
		// FP64 (correctly rounded) and rounded to float, which is the correctly rounded float root
		// (53 >= 2 * 24 + 2 bits).
		const float s = c[0] * c[0] * 0.299f + c[1] * c[1] * 0.587f + c[2] * c[2] * 0.114f;
		const float V = (float)sqrt((double)s);
		uint16_t o[3];
		for (int j = 0; j < 3; j++)
		{
			float x = V + (c[j] - V) * saturation;
			x = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
			o[j] = gammaTable[(uint16_t)(x * 65535.0f)];
		}
		if (out16)
		{
			out16[k] = o[0];
			out16[k + 1] = o[1];
			out16[k + 2] = o[2];
		}
		if (out8)
		{
			out8[k] = (uint8_t)(o[0] / 256);
			out8[k + 1] = (uint8_t)(o[1] / 256);
			out8[k + 2] = (uint8_t)(o[2] / 256);
		}
	}
}

hipError_t launch_compile_image(const float *rgb, uint32_t n, const uint16_t *gammaTable, float brightness,
	float contrast, float saturation, int hdr, uint16_t *out16, uint8_t *out8, hipStream_t stream)
{
	const uint32_t blocks = (n + 255u) / 256u;
	const uint32_t grid = blocks < 65536u ? blocks : 65536u;
	hipLaunchKernelGGL(compile_image_kernel, dim3(grid), dim3(256), 0, stream, rgb, n, gammaTable, brightness,
		contrast, saturation, hdr, out16, out8);
	return hipGetLastError();
}

} // namespace mbx

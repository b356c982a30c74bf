// mbx_image.hip - the image post-processing kernel: cImage::CalculatePixel over every pixel of the
// float image (mandelbulber2/src/cimage.cpp:317-361, 378-397) and the 8-bit conversion of
// cImage::ConvertTo8bitChar (cimage.cpp:460-470), in one pass.  HBM-bound: 12 bytes read and
// 6 + 3 bytes written per pixel; the 128 KiB gamma table stays in L2.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "mbx_launch.h"

namespace mbx
{

// cImage::CalculatePixel of one pixel: 16-bit R, G, B
__device__ __forceinline__ void calculate_pixel(float r, float g, float b, const uint16_t *__restrict__ gammaTable,
	float brightness, float contrast, float saturation, int hdr, uint16_t o[3])
{
#pragma clang fp contract(off)
	float c[3] = {r, g, b};
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
	for (int j = 0; j < 3; j++)
	{
		float x = V + (c[j] - V) * saturation;
		x = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
		o[j] = gammaTable[(uint16_t)(x * 65535.0f)];
	}
}

// Four pixels per thread: three 16-byte loads (12 floats), three 8-byte stores of 16-bit samples
// and three 4-byte stores of 8-bit samples; rgb 16-byte, out16 8-byte and out8 4-byte aligned (the
// host checks).  Pixels past the last whole group of four go through the one-pixel path.
__global__ __launch_bounds__(256) void compile_image_kernel(const float *__restrict__ rgb, uint32_t n,
	const uint16_t *__restrict__ gammaTable, float brightness, float contrast, float saturation, int hdr,
	uint16_t *__restrict__ out16, uint8_t *__restrict__ out8, int vec4)
{
	const uint32_t stride = gridDim.x * blockDim.x;
	const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
	const uint32_t quads = vec4 ? n / 4 : 0;
	for (uint32_t q = tid; q < quads; q += stride)
	{
		const float4 *in = (const float4 *)rgb + 3 * (size_t)q;
		const float4 a = in[0], b = in[1], c = in[2];
		const float f[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
		uint16_t o[12];
		for (int k = 0; k < 4; k++)
			calculate_pixel(f[3 * k], f[3 * k + 1], f[3 * k + 2], gammaTable, brightness, contrast, saturation, hdr,
				o + 3 * k);
		if (out16)
		{
			uint2 *d = (uint2 *)out16 + 3 * (size_t)q;
			for (int w = 0; w < 3; w++)
				d[w] = make_uint2((uint32_t)o[4 * w] | ((uint32_t)o[4 * w + 1] << 16),
					(uint32_t)o[4 * w + 2] | ((uint32_t)o[4 * w + 3] << 16));
		}
		if (out8)
		{
			uint32_t *d = (uint32_t *)out8 + 3 * (size_t)q;
			for (int w = 0; w < 3; w++)
				d[w] = (uint32_t)(o[4 * w] >> 8) | ((uint32_t)(o[4 * w + 1] >> 8) << 8) |
					((uint32_t)(o[4 * w + 2] >> 8) << 16) | ((uint32_t)(o[4 * w + 3] >> 8) << 24);
		}
	}
	for (uint32_t i = 4 * quads + tid; i < n; i += stride)
	{
		const size_t k = 3 * (size_t)i;
		uint16_t o[3];
		calculate_pixel(rgb[k], rgb[k + 1], rgb[k + 2], gammaTable, brightness, contrast, saturation, hdr, o);
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
	const bool vec4 = ((uintptr_t)rgb % 16 == 0) && ((uintptr_t)out16 % 8 == 0) && ((uintptr_t)out8 % 4 == 0);
	const uint32_t items = vec4 ? (n / 4 + (n % 4)) : n;
	const uint32_t blocks = (items + 255u) / 256u;
	const uint32_t grid = blocks < 65536u ? (blocks ? blocks : 1u) : 65536u;
	hipLaunchKernelGGL(compile_image_kernel, dim3(grid), dim3(256), 0, stream, rgb, n, gammaTable, brightness,
		contrast, saturation, hdr, out16, out8, vec4 ? 1 : 0);
	return hipGetLastError();
}

} // namespace mbx

namespace mbx
{

// z-buffer range (file_image.cpp:631-651): the smallest depth and the largest depth below 1e19
// (misses are 1e20), as order-preserving bit patterns of the non-negative depths.  range[0] starts
// at +inf (the reference's float(1e50)), range[1] at 0.
__global__ __launch_bounds__(64) void zrange_init_kernel(uint32_t *range)
{
	if (threadIdx.x == 0)
	{
		range[0] = 0x7f800000u;
		range[1] = 0u;
	}
}

__global__ __launch_bounds__(256) void zrange_kernel(const float *__restrict__ depth, uint32_t n, uint32_t *range)
{
	__shared__ uint32_t sMin[256], sMax[256];
	uint32_t lo = 0x7f800000u, hi = 0u;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		const float z = depth[i];
		const uint32_t b = __float_as_uint(z);
		if (z < __uint_as_float(lo)) lo = b;
		if (z > __uint_as_float(hi) && z < 1e19f) hi = b;
	}
	sMin[threadIdx.x] = lo;
	sMax[threadIdx.x] = hi;
	__syncthreads();
	for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1)
	{
		if (threadIdx.x < s)
		{
			sMin[threadIdx.x] = min(sMin[threadIdx.x], sMin[threadIdx.x + s]);
			sMax[threadIdx.x] = max(sMax[threadIdx.x], sMax[threadIdx.x + s]);
		}
		__syncthreads();
	}
	if (threadIdx.x == 0)
	{
		atomicMin(&range[0], sMin[0]);
		atomicMax(&range[1], sMax[0]);
	}
}

// the grey value of each pixel (file_image.cpp:652-742)
__global__ __launch_bounds__(256) void zbuffer_kernel(const float *__restrict__ depth, uint32_t n,
	const uint32_t *__restrict__ range, float minDepth, float maxDepth, int useRange, int bits, int logarithmic,
	int invert, void *__restrict__ out)
{
#pragma clang fp contract(off)
	const float minZ = useRange ? __uint_as_float(range[0]) : minDepth;
	const float maxZ = useRange ? __uint_as_float(range[1]) : maxDepth;
	// logf as the host's: the FP64 log rounded to float (glibc's logf is correctly rounded in
	// practice; the device's differs in the last bit now and then)
	const float kZ = (float)log((double)(maxZ / minZ));
	const int top = bits == 16 ? 65534 : 254;
	const int miss = bits == 16 ? 65535 : 255;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		const float z = depth[i];
		float z1 = logarithmic ? (float)log((double)(z / minZ)) / kZ : (z - minZ) / (maxZ - minZ);
		if (z1 < 0) z1 = 0.0f;
		if (z1 > 1.0f) z1 = 1.0f;
		int intZ = (int)(z1 * (float)top);
		if (z > 1e19f) intZ = miss;
		if (invert) intZ = miss - intZ;
		if (bits == 16)
			((uint16_t *)out)[i] = (uint16_t)intZ;
		else
			((uint8_t *)out)[i] = (uint8_t)intZ;
	}
}

hipError_t launch_zbuffer(const float *depth, uint32_t n, uint32_t *range, int useRange, float minDepth,
	float maxDepth, int bits, int logarithmic, int invert, void *out, hipStream_t stream)
{
	const uint32_t blocks = (n + 255u) / 256u;
	const uint32_t grid = blocks < 4096u ? (blocks ? blocks : 1u) : 4096u;
	if (useRange)
	{
		hipLaunchKernelGGL(zrange_init_kernel, dim3(1), dim3(64), 0, stream, range);
		hipLaunchKernelGGL(zrange_kernel, dim3(grid), dim3(256), 0, stream, depth, n, range);
	}
	hipLaunchKernelGGL(zbuffer_kernel, dim3(grid), dim3(256), 0, stream, depth, n, (const uint32_t *)range, minDepth,
		maxDepth, useRange, bits, logarithmic, invert, out);
	return hipGetLastError();
}

} // namespace mbx

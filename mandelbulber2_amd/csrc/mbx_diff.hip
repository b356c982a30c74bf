// mbx_diff.hip - differentiable render and adjoint kernels for gfx950 (FP64).
//
// The differentiable render keeps the reference's march (RayMarching, render_worker.cpp:708-951,
// through the kind's kernel path) to find the hit, then makes the image a smooth function of the
// scene parameters:
//   * the hit depth t is refined by Newton iterations on G(t) = DE(camera + t dir) - distThresh(t)
//     = 0, the surface the march's termination test dist < distThresh approximates;
//   * the normal is the analytic gradient of the DE (the reference takes central differences,
//     shader_calculate_normals.cpp:44-78);
//   * the Lambert shading (shader_object.cpp:40-229, shader_light_shading.cpp:40-122) and the
//     background (shader_background.cpp:39-200) run in FP64.
// The DE is the Mandelbulb analytic logarithmic DE (fractal_mandelbulb.cpp:29-43,
// compute_fractal.cpp:561-568) of the rotated node (calculate_distance.cpp:296).
//
// Derivatives use hyper-dual numbers v + a e1 + b e2 + c e1 e2 (e1^2 = e2^2 = 0): one evaluation
// gives the value, two directional derivatives and the exact mixed second derivative.  The
// adjoint differentiates per pixel: dt/dtheta from the implicit function G(t, theta) = 0, the
// normal's derivative from the DE's Hessian and mixed parameter derivatives, then the shading.
#include <hip/hip_runtime.h>

#include "mbx_diff.h"
#include "mbx_pixel.h"

namespace mbx
{

struct HD
{
	double v, a, b, c;
};

__device__ __forceinline__ HD hd(double v) { return HD{v, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ HD operator+(HD x, HD y) { return HD{x.v + y.v, x.a + y.a, x.b + y.b, x.c + y.c}; }
__device__ __forceinline__ HD operator-(HD x, HD y) { return HD{x.v - y.v, x.a - y.a, x.b - y.b, x.c - y.c}; }
__device__ __forceinline__ HD operator+(HD x, double s) { return HD{x.v + s, x.a, x.b, x.c}; }
__device__ __forceinline__ HD operator-(double s, HD x) { return HD{s - x.v, -x.a, -x.b, -x.c}; }
__device__ __forceinline__ HD operator*(HD x, double s) { return HD{x.v * s, x.a * s, x.b * s, x.c * s}; }
__device__ __forceinline__ HD operator*(HD x, HD y)
{
	return HD{x.v * y.v, x.v * y.a + x.a * y.v, x.v * y.b + x.b * y.v,
		x.v * y.c + x.a * y.b + x.b * y.a + x.c * y.v};
}
// f(x) from f(v), f'(v), f''(v)
__device__ __forceinline__ HD chain(HD x, double f, double f1, double f2)
{
	return HD{f, f1 * x.a, f1 * x.b, f1 * x.c + f2 * x.a * x.b};
}
__device__ __forceinline__ HD hrcp(HD x)
{
	const double r = 1.0 / x.v;
	return chain(x, r, -r * r, 2.0 * r * r * r);
}
__device__ __forceinline__ HD hsqrt(HD x)
{
	const double s = sqrt(x.v);
	return chain(x, s, 0.5 / s, -0.25 / (s * x.v));
}
__device__ __forceinline__ HD hlog(HD x)
{
	const double r = 1.0 / x.v;
	return chain(x, log(x.v), r, -r * r);
}
__device__ __forceinline__ HD hsin(HD x)
{
	double s, c;
	sincos(x.v, &s, &c);
	return chain(x, s, c, -s);
}
__device__ __forceinline__ HD hcos(HD x)
{
	double s, c;
	sincos(x.v, &s, &c);
	return chain(x, c, -s, -c);
}
__device__ __forceinline__ HD hasin(HD x)
{
	const double q = 1.0 - x.v * x.v;
	const double rq = 1.0 / sqrt(q);
	return chain(x, asin(x.v), rq, x.v * rq * rq * rq);
}
// atan2(y, x) with its first and second partials
__device__ __forceinline__ HD hatan2(HD y, HD x)
{
	const double q = x.v * x.v + y.v * y.v;
	const double iq = 1.0 / q, iq2 = iq * iq;
	const double fy = x.v * iq, fx = -y.v * iq;
	const double fyy = -2.0 * x.v * y.v * iq2, fxx = 2.0 * x.v * y.v * iq2;
	const double fxy = (y.v * y.v - x.v * x.v) * iq2;
	return HD{atan2(y.v, x.v), fy * y.a + fx * x.a, fy * y.b + fx * x.b,
		fy * y.c + fx * x.c + fyy * y.a * y.b + fxx * x.a * x.b + fxy * (x.a * y.b + y.a * x.b)};
}
// x^p (x > 0) with both arguments differentiable
__device__ __forceinline__ HD hpow(HD x, HD p)
{
	const double f = pow(x.v, p.v);
	const double lx = log(x.v);
	const double ix = 1.0 / x.v;
	const double fx = p.v * f * ix, fp = f * lx;
	const double fxx = p.v * (p.v - 1.0) * f * ix * ix, fpp = f * lx * lx, fxp = f * ix * (1.0 + p.v * lx);
	return HD{f, fx * x.a + fp * p.a, fx * x.b + fp * p.b,
		fx * x.c + fp * p.c + fxx * x.a * x.b + fpp * p.a * p.b + fxp * (x.a * p.b + p.a * x.b)};
}
// x^p for a constant p (x >= 0)
__device__ __forceinline__ HD hpowc(HD x, double p)
{
	return chain(x, pow(x.v, p), p * pow(x.v, p - 1.0), p * (p - 1.0) * pow(x.v, p - 2.0));
}

struct HD3
{
	HD x, y, z;
};
__device__ __forceinline__ HD hdot(const HD3 &u, const HD3 &w) { return u.x * w.x + u.y * w.y + u.z * w.z; }

// Analytic logarithmic DE of the refined render at world point pw (seeds already in pw).  With
// rotSeed in 0..2 the node rotation's derivative per degree of that axis rides on e2, with
// powSeed the power's.  ok = false where the DE is not differentiable (non-finite orbit, r <= 1,
// DE <= 0).
__device__ HD diff_de(const DiffConst &c, const HD3 &pw, int rotSeed, bool powSeed, int &iters, bool &ok)
{
	HD Rm[9];
#pragma unroll
	for (int k = 0; k < 9; k++) Rm[k] = HD{c.R[k], 0.0, rotSeed >= 0 ? c.dR[rotSeed][k] : 0.0, 0.0};
	const HD3 p{Rm[0] * pw.x + Rm[1] * pw.y + Rm[2] * pw.z, Rm[3] * pw.x + Rm[4] * pw.y + Rm[5] * pw.z,
		Rm[6] * pw.x + Rm[7] * pw.y + Rm[8] * pw.z};
	const HD P{c.power, 0.0, powSeed ? 1.0 : 0.0, 0.0};
	const HD Pm1 = P + (-1.0);
	HD3 z = p;
	HD DE = hd(1.0);
	HD r = hsqrt(hdot(z, z));
	int i;
	for (i = 0; i < c.maxN; i++)
	{
		if (!(r.v > 0.0))
		{
			ok = false;
			break;
		}
		// FormulaCode (fractal_mandelbulb.cpp:32-42)
		const HD th0 = hasin(z.z * hrcp(r)) + c.beta;
		const HD ph0 = hatan2(z.y, z.x) + c.alpha;
		HD rp = hpow(r, Pm1);
		const HD th = th0 * P;
		const HD ph = ph0 * P;
		const HD cth = hcos(th);
		DE = (rp * DE) * P + 1.0;
		rp = rp * r;
		// z += c, c = p * constantMultiplier (compute_fractal.cpp:246-281)
		z = HD3{cth * hcos(ph) * rp + p.x * c.cm[0], cth * hsin(ph) * rp + p.y * c.cm[1],
			hsin(th) * rp + p.z * c.cm[2]};
		r = hsqrt(hdot(z, z));
		if (!isfinite(r.v))
		{
			ok = false;
			break;
		}
		if (r.v > c.bailout) break;
	}
	iters = i + 1;
	if (!(DE.v > 0.0) || !(r.v > 1.0))
	{
		ok = false;
		return hd(0.0);
	}
	return r * hlog(r) * hrcp(DE) * 0.5; // 0.5 r ln r / DE (compute_fractal.cpp:561-568)
}

__device__ __forceinline__ HD3 hd3_point(DV3 p) { return HD3{hd(p.x), hd(p.y), hd(p.z)}; }

// FP64 view vector (doWork image point, CalculateViewVector, normalised): before rotation `vv`,
// after rotation `d`.
__device__ __forceinline__ void diff_view(const DiffConst &c, int xs, int ys, DV3 &vv, DV3 &d)
{
	const double ix = (c.invW * (double)xs - 0.5) * c.aspect;
	const double iy = 0.5 - c.invH * (double)ys;
	vv = dv3(ix * c.fov, 1.0, iy * c.fov);
	vv = vv * (1.0 / sqrt(ddot(vv, vv)));
	const double *m = c.M;
	d = dv3(m[0] * vv.x + m[1] * vv.y + m[2] * vv.z, m[3] * vv.x + m[4] * vv.y + m[5] * vv.z,
		m[6] * vv.x + m[7] * vv.y + m[8] * vv.z);
	d = d * (1.0 / sqrt(ddot(d, d)));
}

// d(dir)/d(theta_k) for the camera parameters: dir = normalize(M vv)
__device__ __forceinline__ DV3 diff_view_tangent(const DiffConst &c, int k, DV3 vv, DV3 d)
{
	const double *m = c.dM[k];
	const DV3 u = dv3(m[0] * vv.x + m[1] * vv.y + m[2] * vv.z, m[3] * vv.x + m[4] * vv.y + m[5] * vv.z,
		m[6] * vv.x + m[7] * vv.y + m[8] * vv.z);
	return u - d * ddot(d, u); // |M vv| = 1
}

__device__ __forceinline__ double diff_thresh(const DiffConst &c, double t)
{
	return c.constThr ? c.DEThresh : t * c.threshK;
}

// Newton on G(t) = DE(o + t d) - distThresh(t) from the march's depth.  Returns false when an
// iteration leaves the threshold's neighbourhood, the ray does not approach the surface
// (dG/dt >= 0) or the DE is not differentiable.
__device__ bool refine_hit(const DiffConst &c, DV3 o, DV3 d, double &t, int &iters)
{
	const double dthr = c.constThr ? 0.0 : c.threshK;
	for (int it = 0; it < 24; it++)
	{
		const HD3 p{HD{o.x + t * d.x, d.x, 0.0, 0.0}, HD{o.y + t * d.y, d.y, 0.0, 0.0},
			HD{o.z + t * d.z, d.z, 0.0, 0.0}};
		bool ok = true;
		int n = 0;
		const HD de = diff_de(c, p, -1, false, n, ok);
		if (!ok) return false;
		const double thr = diff_thresh(c, t);
		const double G = de.v - thr;
		const double Gt = de.a - dthr;
		if (!(Gt < 0.0)) return false;
		const double dt = -G / Gt;
		if (!(fabs(dt) <= 4.0 * thr)) return false;
		t += dt;
		iters = n;
		if (fabs(dt) <= 1e-15 * t) return true;
	}
	return false;
}

// DE gradient (first-order passes along the axes)
__device__ bool de_gradient(const DiffConst &c, DV3 p, double g[3], int &iters)
{
	bool ok = true;
	for (int j = 0; j < 3; j++)
	{
		HD3 q = hd3_point(p);
		if (j == 0) q.x.a = 1.0;
		if (j == 1) q.y.a = 1.0;
		if (j == 2) q.z.a = 1.0;
		g[j] = diff_de(c, q, -1, false, iters, ok).a;
	}
	return ok;
}

// Lambert shade and colour (FP64), as object_shader with shadows and AO off
struct DiffShade
{
	double rgb[3];
	double ndotl, shadeRaw; // before the max(., 0) and the 500 clamp
};
__device__ DiffShade diff_shade(const DiffConst &c, const double n[3])
{
	DiffShade s;
	s.ndotl = n[0] * c.light[0] + n[1] * c.light[1] + n[2] * c.light[2];
	const double sh = s.ndotl > 0.0 ? s.ndotl : 0.0;
	s.shadeRaw = (1.0 - c.shading + sh * c.shading) * c.intensity;
	const double shade = s.shadeRaw < 500.0 ? s.shadeRaw : 500.0;
	for (int k = 0; k < 3; k++)
	{
		const double light = c.lightEnabled ? c.lightColor[k] * shade : 0.0;
		const double v = c.fill[k] * c.colour[k] + light * c.colour[k] + c.lum[k];
		s.rgb[k] = v > 0.0 ? v : 0.0;
	}
	return s;
}

// Background (3-colour gradient and the light's disk), generic in the direction's tangents
__device__ HD3 diff_background(const DiffConst &c, HD3 d)
{
	const HD il = hrcp(hsqrt(hdot(d, d)));
	const HD3 vn{d.x * il, d.y * il, d.z * il};
	HD col[3];
	if (c.bg3)
	{
		const HD grad = vn.z + 1.0;
		for (int k = 0; k < 3; k++)
			col[k] = (grad.v < 1.0 ? (1.0 - grad) * c.bgc3[k] + grad * c.bg2[k]
								   : (2.0 - grad) * c.bg2[k] + (grad + (-1.0)) * c.bg1[k]) * c.bgBrightness;
	}
	else
		for (int k = 0; k < 3; k++) col[k] = hd(c.bg1[k] * c.bgBrightness);
	if (c.lightEnabled)
	{
		const HD dl = vn.x * c.light[0] + vn.y * c.light[1] + vn.z * c.light[2];
		const HD a = (1.0 - dl) * (360.0 / c.lightSize);
		const HD inten = hrcp(hpowc(a, c.sunPow) + 1.0) * (c.visibility * c.intensity);
		for (int k = 0; k < 3; k++) col[k] = col[k] + inten * c.lightColor[k];
	}
	return HD3{col[0], col[1], col[2]};
}

// ------------------------------------------------------------------------------------------
// forward: march, refine, shade, checkpoint
// ------------------------------------------------------------------------------------------
// two waves per SIMD (256 VGPRs, 12 bytes of scratch): 4-5% faster than one wave at 260; the
// adjoint stays at one wave (at two it spills 624 bytes per lane and runs 8% slower)
template <int KIND>
__global__ __launch_bounds__(64, 2) void render_diff_kernel(const FrameConst *__restrict__ frame,
	const DiffConst *__restrict__ dcp, float *__restrict__ rgb, float *__restrict__ depth,
	int32_t *__restrict__ itersOut, uint8_t *__restrict__ mask, float *__restrict__ normal,
	RayCheckpoint *__restrict__ ckpt)
{
	const FrameConst &f = *frame;
	const DiffConst &c = *dcp;
	const int tilesX = (f.width + 7) >> 3;
	const int xs = (blockIdx.x % tilesX) * 8 + (threadIdx.x & 7);
	const int ys = (blockIdx.x / tilesX) * 8 + (threadIdx.x >> 3);
	if (xs >= f.width || ys >= f.height) return;
	const size_t i = (size_t)ys * f.width + xs;

	// the march of the kind, as render_pixel runs it for one sample
	using Q = PrecOf<KIND>;
	using R = typename Q::R;
	Work w = make_work(f);
	uint32_t seed = pixel_seed((uint32_t)i, f.randomSeed);
	const R ox = (Q::invWidth(f) * (R)xs - (R)0.5) * Q::aspect(f);
	const R oy = (R)0.5 - Q::invHeight(f) * (R)ys;
	MarchOut<KIND> m;
	ray_marching<KIND, 2>(f, Q::cam(f), Q::view(f, ox, oy), (R)f.viewDistanceMax, seed, m, w);

	DV3 vv, d;
	diff_view(c, xs, ys, vv, d);
	const DV3 o = dv3(c.cam[0], c.cam[1], c.cam[2]);
	RayCheckpoint ck{0.0, m.found ? 1 : 0, 0, m.iters, 0};
	double col[3], nrm[3] = {0.0, 0.0, 0.0};
	if (m.found)
	{
		double t = (double)m.depth;
		int n = m.iters;
		ck.refined = refine_hit(c, o, d, t, n) ? 1 : 0;
		if (!ck.refined) t = (double)m.depth;
		ck.t = t;
		ck.iters = n;
		double g[3];
		int n2;
		const bool okg = de_gradient(c, o + d * t, g, n2);
		const double gl = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
		if (okg && gl > 0.0)
			for (int k = 0; k < 3; k++) nrm[k] = g[k] / gl;
		else
		{
			nrm[0] = 1.0; // the reference's fallback normal (shader_calculate_normals.cpp:111-123)
			ck.refined = 0;
		}
		const DiffShade s = diff_shade(c, nrm);
		for (int k = 0; k < 3; k++) col[k] = s.rgb[k];
	}
	else
	{
		const HD3 bg = diff_background(c, hd3_point(d));
		col[0] = bg.x.v;
		col[1] = bg.y.v;
		col[2] = bg.z.v;
	}
	if (rgb)
		for (int k = 0; k < 3; k++) rgb[3 * i + k] = (float)col[k];
	if (depth) depth[i] = m.found ? (float)ck.t : 1e20f;
	if (itersOut) itersOut[i] = ck.iters;
	if (mask) mask[i] = (uint8_t)ck.found;
	if (normal)
		for (int k = 0; k < 3; k++) normal[3 * i + k] = (float)nrm[k];
	ckpt[i] = ck;
}

// ------------------------------------------------------------------------------------------
// adjoint
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
	return v;
}

__global__ void zero_f64_kernel(double *p, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 0.0;
}

__global__ __launch_bounds__(64) void adjoint_kernel(const FrameConst *__restrict__ frame,
	const DiffConst *__restrict__ dcp, const RayCheckpoint *__restrict__ ckpt,
	const float *__restrict__ dLdI, double *__restrict__ grad, double *__restrict__ pixelGrad)
{
	const FrameConst &f = *frame;
	const DiffConst &c = *dcp;
	const int tilesX = (f.width + 7) >> 3;
	const int xs = (blockIdx.x % tilesX) * 8 + (threadIdx.x & 7);
	const int ys = (blockIdx.x / tilesX) * 8 + (threadIdx.x >> 3);
	double gk[kDiffParams];
#pragma unroll
	for (int k = 0; k < kDiffParams; k++) gk[k] = 0.0;
	if (xs < f.width && ys < f.height)
	{
		const size_t i = (size_t)ys * f.width + xs;
		const RayCheckpoint ck = ckpt[i];
		const double L[3] = {(double)dLdI[3 * i], (double)dLdI[3 * i + 1], (double)dLdI[3 * i + 2]};
		DV3 vv, d;
		diff_view(c, xs, ys, vv, d);
		const DV3 o = dv3(c.cam[0], c.cam[1], c.cam[2]);
		if (ck.found && ck.refined)
		{
			const double t = ck.t;
			const DV3 p = o + d * t;
			// gradient and Hessian of the DE at the refined point: passes (e_i, e_j), i <= j
			double g[3], H[3][3];
			bool ok = true;
			int n;
			for (int a = 0; a < 3; a++)
				for (int b = a; b < 3; b++)
				{
					HD3 q = hd3_point(p);
					(a == 0 ? q.x : a == 1 ? q.y : q.z).a = 1.0;
					(b == 0 ? q.x : b == 1 ? q.y : q.z).b = 1.0;
					const HD de = diff_de(c, q, -1, false, n, ok);
					H[a][b] = H[b][a] = de.c;
					g[a] = de.a;
				}
			const double gl = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
			const double Gt = g[0] * d.x + g[1] * d.y + g[2] * d.z - (c.constThr ? 0.0 : c.threshK);
			if (ok && gl > 0.0 && Gt < 0.0)
			{
				const double nv[3] = {g[0] / gl, g[1] / gl, g[2] / gl};
				const DiffShade s = diff_shade(c, nv);
				const bool shadeLive = c.lightEnabled && s.ndotl > 0.0 && s.shadeRaw < 500.0;
				const double dd[3] = {d.x, d.y, d.z};
				double Hd[3];
				for (int a = 0; a < 3; a++) Hd[a] = H[a][0] * d.x + H[a][1] * d.y + H[a][2] * d.z;
				for (int k = 0; k < kDiffParams; k++)
				{
					// dg = total derivative of the DE gradient at the refined point
					double dg[3];
					if (k < MBX_DIFF_CAMERA)
					{
						// power / rotation: mixed passes (e_j, parameter)
						double Gk = 0.0, mix[3];
						for (int j = 0; j < 3; j++)
						{
							HD3 q = hd3_point(p);
							(j == 0 ? q.x : j == 1 ? q.y : q.z).a = 1.0;
							const HD de = diff_de(c, q, k == MBX_DIFF_POWER ? -1 : k - MBX_DIFF_ROTATION,
								k == MBX_DIFF_POWER, n, ok);
							mix[j] = de.c;
							Gk = de.b;
						}
						const double dt = -Gk / Gt;
						for (int j = 0; j < 3; j++) dg[j] = mix[j] + Hd[j] * dt;
					}
					else
					{
						// camera / target: the ray moves, dp = dray + d dt
						const int kc = k - MBX_DIFF_CAMERA;
						const DV3 ddk = diff_view_tangent(c, kc, vv, d);
						double pk[3] = {ddk.x * t, ddk.y * t, ddk.z * t};
						if (kc < 3) pk[kc] += 1.0;
						const double Gk = g[0] * pk[0] + g[1] * pk[1] + g[2] * pk[2];
						const double dt = -Gk / Gt;
						double dp[3];
						for (int j = 0; j < 3; j++) dp[j] = pk[j] + dd[j] * dt;
						for (int j = 0; j < 3; j++) dg[j] = H[j][0] * dp[0] + H[j][1] * dp[1] + H[j][2] * dp[2];
					}
					// normal n = g / |g|, Lambert shade
					const double ndg = nv[0] * dg[0] + nv[1] * dg[1] + nv[2] * dg[2];
					double ldn = 0.0;
					for (int j = 0; j < 3; j++) ldn += c.light[j] * (dg[j] - nv[j] * ndg) / gl;
					const double dshade = shadeLive ? c.intensity * c.shading * ldn : 0.0;
					double acc = 0.0;
					for (int ch = 0; ch < 3; ch++)
						if (s.rgb[ch] > 0.0) acc += L[ch] * c.colour[ch] * c.lightColor[ch] * dshade;
					gk[k] = ok ? acc : 0.0;
				}
				if (!ok)
					for (int k = 0; k < kDiffParams; k++) gk[k] = 0.0;
			}
		}
		else if (!ck.found)
		{
			// miss: the background moves with the view vector (camera parameters only)
			for (int kc = 0; kc < 6; kc++)
			{
				const DV3 ddk = diff_view_tangent(c, kc, vv, d);
				const HD3 q{HD{d.x, ddk.x, 0.0, 0.0}, HD{d.y, ddk.y, 0.0, 0.0}, HD{d.z, ddk.z, 0.0, 0.0}};
				const HD3 bg = diff_background(c, q);
				gk[MBX_DIFF_CAMERA + kc] = L[0] * bg.x.a + L[1] * bg.y.a + L[2] * bg.z.a;
			}
		}
		if (pixelGrad)
			for (int k = 0; k < kDiffParams; k++) pixelGrad[(size_t)kDiffParams * i + k] = gk[k];
	}
	for (int k = 0; k < kDiffParams; k++)
	{
		const double sum = wave_sum_f64(gk[k]);
		if (threadIdx.x == 0) atomicAdd(&grad[k], sum);
	}
}

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------
template <int KIND>
static hipError_t launch_render_diff_t(const FrameConst *frame, const DiffConst *dc, int width, int height,
	float *rgb, float *depth, int32_t *iters, uint8_t *mask, float *normal, RayCheckpoint *ckpt,
	hipStream_t stream)
{
	const int tiles = ((width + 7) / 8) * ((height + 7) / 8);
	hipLaunchKernelGGL(render_diff_kernel<KIND>, dim3(tiles), dim3(64), 0, stream, frame, dc, rgb, depth,
		iters, mask, normal, ckpt);
	return hipGetLastError();
}

hipError_t launch_render_diff(int deKind, const FrameConst *frame, const DiffConst *dc, int width,
	int height, float *rgb, float *depth, int32_t *iters, uint8_t *mask, float *normal,
	RayCheckpoint *ckpt, hipStream_t stream)
{
	switch (deKind)
	{
		case kBulb8:
			return launch_render_diff_t<kBulb8>(frame, dc, width, height, rgb, depth, iters, mask, normal,
				ckpt, stream);
		case kBulbTrig:
			return launch_render_diff_t<kBulbTrig>(frame, dc, width, height, rgb, depth, iters, mask, normal,
				ckpt, stream);
		default: return hipErrorInvalidValue;
	}
}

hipError_t launch_adjoint(const FrameConst *frame, const DiffConst *dc, int width, int height,
	const RayCheckpoint *ckpt, const float *dLdI, double *grad, double *pixelGrad, hipStream_t stream)
{
	hipLaunchKernelGGL(zero_f64_kernel, dim3(1), dim3(64), 0, stream, grad, kDiffParams);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	const int tiles = ((width + 7) / 8) * ((height + 7) / 8);
	hipLaunchKernelGGL(adjoint_kernel, dim3(tiles), dim3(64), 0, stream, frame, dc, ckpt, dLdI, grad,
		pixelGrad);
	return hipGetLastError();
}

} // namespace mbx

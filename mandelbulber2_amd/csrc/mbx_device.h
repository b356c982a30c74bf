// mbx_device.h - device-side building blocks of the raymarcher: vector helpers, the
// Park-Miller stream, the formula plugins (FormulaCode), the Compute loop and
// CalculateDistance, all in FP32 for gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mbx_frame.h"

namespace mbx
{

// ------------------------------------------------------------------------------------------
// small vector helpers
// ------------------------------------------------------------------------------------------
struct V3
{
	float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a)
{
	const float k = 1.0f / length(a);
	return a * k;
}
__device__ __forceinline__ bool finite3(V3 a)
{
	return isfinite(a.x) && isfinite(a.y) && isfinite(a.z);
}
__device__ __forceinline__ V3 rot(const float *m, V3 v)
{
	return V3{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
		m[6] * v.x + m[7] * v.y + m[8] * v.z};
}

// ------------------------------------------------------------------------------------------
// Park-Miller minimal standard generator (mandelbulber2/src/common_math.cpp:47-72), one
// stream per pixel.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int random_int(uint32_t &seed)
{
	uint32_t lo = 16807u * (seed & 0xFFFFu);
	uint32_t hi = 16807u * (seed >> 16);
	lo += (hi & 0x7FFFu) << 16;
	hi >>= 15;
	lo += hi;
	if (lo > 2147483647u) lo -= 2147483647u;
	seed = lo;
	return (int)lo;
}

// Random(max) = RandomInt() % (max + 1); max + 1 is a compile-time constant at every call site
// except the AO sample pick.
__device__ __forceinline__ int random_mod(uint32_t &seed, int max)
{
	return random_int(seed) % (max + 1);
}

// ------------------------------------------------------------------------------------------
// formula plugins.  Each keeps the reference's FormulaCode(z, fractal, aux) contract: z and
// the running derivative are updated in place.  w stays 0 on the device path (the host
// rejects a non-zero initial_waxis), so z is carried as three floats.
// ------------------------------------------------------------------------------------------

// sign() of mandelbulber2/src/math_utilities.h:65 (-1, 0 or +1)
__device__ __forceinline__ float sgn(float v) { return (float)((0.0f < v) - (v < 0.0f)); }

// (c + i s)^8 by three complex squarings
__device__ __forceinline__ void unit_pow8(float c, float s, float &c8, float &s8)
{
	float c2 = c * c - s * s, s2 = 2.0f * c * s;
	float c4 = c2 * c2 - s2 * s2, s4 = 2.0f * c2 * s2;
	c8 = c4 * c4 - s4 * s4;
	s8 = 2.0f * c4 * s4;
}

// Mandelbulb power 8 with zero angle offsets (fractal_mandelbulb.cpp:29-43).  With
// theta0 = asin(z/r), phi0 = atan2(y, x) and rho = sqrt(x^2 + y^2) >= 0:
//   (cos theta0, sin theta0) = (rho, z) / r      and   (cos phi0, sin phi0) = (x, y) / rho,
// so cos/sin of 8*theta0 and 8*phi0 are the 8th powers of these unit complex numbers.
// atan2(0, 0) = 0 in the reference, i.e. (cos 8phi0, sin 8phi0) = (1, 0) on the z axis.
__device__ __forceinline__ void MandelbulbPow8Iteration(V3 &z, float r2, float &DE)
{
	const float rho2 = z.x * z.x + z.y * z.y;
	const float inv_r = __builtin_amdgcn_rsqf(r2);
	const float inv_rho = __builtin_amdgcn_rsqf(rho2);
	const float rho = rho2 * inv_rho;
	float ct8, st8, cp8, sp8;
	unit_pow8(rho * inv_r, z.z * inv_r, ct8, st8);
	const bool axis = rho2 == 0.0f;
	unit_pow8(axis ? 1.0f : z.x * inv_rho, axis ? 0.0f : z.y * inv_rho, cp8, sp8);
	const float r4 = r2 * r2;
	const float r8 = r4 * r4;
	DE = (r8 * inv_r * DE) * 8.0f + 1.0f; // r^7 * DE * power + 1
	const float a = ct8 * r8;
	z.x = a * cp8;
	z.y = a * sp8;
	z.z = st8 * r8;
}

// Mandelbulb, any power and angle offsets: the reference's trigonometric form.
__device__ __forceinline__ void MandelbulbIteration(V3 &z, float r, float &DE, const FrameConst &f)
{
	const float th0 = asinf(z.z / r) + f.beta;
	const float ph0 = atan2f(z.y, z.x) + f.alpha;
	float rp = __expf(__logf(r) * (f.power - 1.0f));
	const float th = th0 * f.power;
	const float ph = ph0 * f.power;
	float sth, cth, sph, cph;
	__sincosf(th, &sth, &cth);
	__sincosf(ph, &sph, &cph);
	DE = (rp * DE) * f.power + 1.0f;
	rp *= r;
	z.x = cth * cph * rp;
	z.y = cth * sph * rp;
	z.z = sth * rp;
}

// Mandelbox without rotated folds (fractal_mandelbox.cpp:31-107).  The three box folds and
// the two sphere-fold regimes are select-based, so a wave does not diverge on them.
// With TANGENT, the three columns of dz/dp are carried along (forward mode) for the delta DE.
template <bool TANGENT>
__device__ __forceinline__ void MandelboxIteration(V3 &z, float &DE, V3 *J, const FrameConst &f)
{
	const float L = f.mbFoldingLimit, V = f.mbFoldingValue;
	const bool fx = fabsf(z.x) > L, fy = fabsf(z.y) > L, fz = fabsf(z.z) > L;
	z.x = fx ? sgn(z.x) * V - z.x : z.x;
	z.y = fy ? sgn(z.y) * V - z.y : z.y;
	z.z = fz ? sgn(z.z) * V - z.z : z.z;
	if (TANGENT)
	{
#pragma unroll
		for (int k = 0; k < 3; k++)
		{
			J[k].x = fx ? -J[k].x : J[k].x;
			J[k].y = fy ? -J[k].y : J[k].y;
			J[k].z = fz ? -J[k].z : J[k].z;
		}
	}
	const float r2 = dot(z, z);
	const V3 o = v3(f.mbOffset[0], f.mbOffset[1], f.mbOffset[2]);
	const V3 zo = z + o;
	const bool inner = r2 < f.mbMR2;
	const bool outer = !inner && r2 < f.mbFR2;
	const float t2 = f.mbFR2 / r2;
	const float k = inner ? f.mbFactor1 : (outer ? t2 : 1.0f);
	if (TANGENT)
	{
		// d[(z + o) t(z)] = t dz - 2 t (z.dz) / r2 (z + o) in the outer regime, k dz otherwise
		const float g = outer ? -2.0f * t2 / r2 : 0.0f;
#pragma unroll
		for (int c = 0; c < 3; c++)
		{
			const float zd = dot(z, J[c]) * g;
			J[c] = (J[c] * k + zo * zd) * f.mbScale;
		}
	}
	z = (zo * k - o) * f.mbScale;
	DE = DE * k * f.mbAbsScale + 1.0f;
}

// Kaleidoscopic IFS (fractal_kaleidoscopic_ifs.cpp:30-74)
__device__ __forceinline__ void KaleidoscopicIfsIteration(V3 &z, float &DE, const FrameConst &f)
{
	if (f.ifsAbs[0]) z.x = fabsf(z.x);
	if (f.ifsAbs[1]) z.y = fabsf(z.y);
	if (f.ifsAbs[2]) z.z = fabsf(z.z);
	if (f.ifsEnabledCount > 0)
	{
		for (int i = 0; i < MBX_IFS_VECTOR_COUNT; i++)
		{
			if (f.ifsEnabled[i])
			{
				z = rot(f.ifsRot[i], z);
				const V3 d = v3(f.ifsDirection[i][0], f.ifsDirection[i][1], f.ifsDirection[i][2]);
				const float len = dot(z, d);
				if (len < f.ifsDistance[i])
					z = z - d * (2.0f * (len - f.ifsDistance[i]) * f.ifsIntensity[i]);
			}
		}
	}
	const V3 off = v3(f.ifsOffset[0], f.ifsOffset[1], f.ifsOffset[2]);
	if (f.ifsRotationEnabled) z = rot(f.ifsMainRot, z - off) + off;
	if (f.ifsEdgeEnabled)
	{
		if (f.ifsEdge[0] > 0.0f) z.x = f.ifsEdge[0] - fabsf(f.ifsEdge[0] - z.x);
		if (f.ifsEdge[1] > 0.0f) z.y = f.ifsEdge[1] - fabsf(f.ifsEdge[1] - z.y);
		if (f.ifsEdge[2] > 0.0f) z.z = f.ifsEdge[2] - fabsf(f.ifsEdge[2] - z.z);
	}
	z = z * f.ifsScale;
	const float sm1 = f.ifsScale - 1.0f;
	if (f.ifsMenger)
	{
		z.x -= off.x * sm1;
		z.y -= off.y * sm1;
		if (z.z > 0.5f * off.z * sm1) z.z -= off.z * sm1;
	}
	else
	{
		z = z - off * sm1;
	}
	DE *= f.ifsAbsScale;
}

// ------------------------------------------------------------------------------------------
// Compute<calcModeNormal> + CalculateDistanceSimple (compute_fractal.cpp:67-654,
// calculate_distance.cpp:533-763) for one single-formula node.
// ------------------------------------------------------------------------------------------
struct Work
{
	uint32_t de_evals;
	uint32_t iterations;
};

__device__ __forceinline__ V3 add_c(V3 z, V3 p, const FrameConst &f)
{
	if (!f.addC) return z;
	if (f.julia) return z + v3(f.juliaC[0], f.juliaC[1], f.juliaC[2]);
	return z + v3(p.x * f.constMul[0], p.y * f.constMul[1], p.z * f.constMul[2]);
}

// Analytic-DE orbit.  Returns the distance and the iteration count (i + 1, maxN + 1 when the
// orbit never escapes).
template <int KIND>
__device__ __forceinline__ float compute_analytic(V3 p, const FrameConst &f, int &iters)
{
	V3 z = p;
	float r2 = dot(z, z);
	float DE = 1.0f;
	const int maxN = f.formulaMaxiter;
	const float bailout2 = f.bailout * f.bailout;
	int i;
	for (i = 0; i < maxN; i++)
	{
		const V3 lastZ = z;
		if (KIND == kBulb8)
			MandelbulbPow8Iteration(z, r2, DE);
		else if (KIND == kBulbTrig)
			MandelbulbIteration(z, sqrtf(r2), DE, f);
		else if (KIND == kBoxAnalytic)
			MandelboxIteration<false>(z, DE, nullptr, f);
		else
			KaleidoscopicIfsIteration(z, DE, f);
		z = add_c(z, p, f);
		r2 = dot(z, z);
		if (!finite3(z))
		{
			z = lastZ;
			r2 = dot(z, z);
			break;
		}
		if (r2 > bailout2) break;
	}
	iters = i + 1;
	const float r = sqrtf(r2);
	float distance;
	if (DE > 0.0f)
	{
		if (f.analyticFunction == MBX_ANALYTIC_LOGARITHMIC)
			distance = (r > 1.0f) ? 0.5f * r * __logf(r) / DE : 0.0f;
		else if (f.analyticFunction == MBX_ANALYTIC_IFS)
			distance = (r - 2.0f) / DE;
		else
			distance = r / DE;
	}
	else
		distance = r;
	return distance;
}

// Delta-DE orbit of the Mandelbox.  The reference differences |z_n| over three axis offsets
// of delta = max(|p| 1e-14, 1e-5 detailSize), below FP32 resolution at screen-space detail
// sizes; the device carries dz_n/dp instead and takes dr_k = |z_n . dz_n/dp_k| / |z_n|, the
// limit of the reference's quotient.  A non-escaping orbit reproduces the reference's
// quirk: its offset orbits run one iteration further (forcedMaxiter = maxN + 1), so the quotient
// is ||z_{N+1}| - |z_N|| / delta on every axis.
__device__ __forceinline__ float compute_delta_box(V3 p, const FrameConst &f, float detailSize,
	int &iters)
{
	V3 z = p;
	V3 J[3] = {v3(1.0f, 0.0f, 0.0f), v3(0.0f, 1.0f, 0.0f), v3(0.0f, 0.0f, 1.0f)};
	float DE = 1.0f;
	const int maxN = f.formulaMaxiter;
	const float bailout2 = f.bailout * f.bailout;
	const V3 cm = f.julia ? v3(0.0f, 0.0f, 0.0f) : v3(f.constMul[0], f.constMul[1], f.constMul[2]);
	int i;
	bool escaped = false;
	for (i = 0; i < maxN; i++)
	{
		const V3 lastZ = z;
		MandelboxIteration<true>(z, DE, J, f);
		z = add_c(z, p, f);
		if (f.addC)
		{
			J[0].x += cm.x;
			J[1].y += cm.y;
			J[2].z += cm.z;
		}
		if (!finite3(z))
		{
			z = lastZ;
			break;
		}
		if (dot(z, z) > bailout2)
		{
			escaped = true;
			break;
		}
	}
	iters = i + 1;
	const float r = length(z);
	float dr;
	if (escaped)
	{
		const float inv_r = 1.0f / r;
		const float d1 = fabsf(dot(z, J[0])) * inv_r;
		const float d2 = fabsf(dot(z, J[1])) * inv_r;
		const float d3 = fabsf(dot(z, J[2])) * inv_r;
		dr = sqrtf(d1 * d1 + d2 * d2 + d3 * d3);
	}
	else
	{
		V3 z2 = z;
		float DE2 = DE;
		MandelboxIteration<false>(z2, DE2, nullptr, f);
		z2 = add_c(z2, p, f);
		const float delta = fmaxf(length(p) * 1e-14f, 1e-5f * detailSize);
		dr = 1.7320508f * fabsf(length(z2) - r) / delta;
	}
	float distance;
	if (dr > 0.0f)
		distance = (f.deFunction == MBX_DE_FUNCTION_LOGARITHMIC) ? 0.5f * r * __logf(r) / dr
		                                                        : 0.5f * r / dr;
	else
		distance = r;
	return distance;
}

// CalculateDistance for the single fractal node (calculate_distance.cpp:225-516 reduces to
// CalculateDistanceSimple for one node with identity transform and scale 1).  iterThresh and
// interior modes are off on the device path.
template <int KIND>
__device__ __forceinline__ float calculate_distance(V3 p, float detailSize, const FrameConst &f,
	int &iters, Work &w)
{
	float d;
	if (KIND == kBoxDelta)
	{
		d = compute_delta_box(p, f, detailSize, iters);
		w.iterations += 4u * (uint32_t)iters;
	}
	else
	{
		d = compute_analytic<KIND>(p, f, iters);
		w.iterations += (uint32_t)iters;
	}
	w.de_evals++;
	if (iters < f.minN && d < detailSize) d = detailSize;
	// clamp as the reference does; a NaN passes the clamp and the boolean-add merge of the
	// objects tree then keeps its initial 1e20
	d = (d < 0.0f) ? 0.0f : d;
	d = (d > 10.0f) ? 10.0f : d;
	if (!(d < 1e20f)) d = 1e20f;
	return d;
}

} // namespace mbx

namespace mbx
{

// ------------------------------------------------------------------------------------------
// Resumable DE evaluation for the persistent kernels: the Compute loop split into begin /
// one-iteration step / finish, so that every lane of a wave advances its own evaluation by
// one formula iteration per loop trip.
// ------------------------------------------------------------------------------------------
template <int KIND>
struct DEState
{
	V3 p;       // evaluation point
	V3 z;
	float r2;   // |z|^2 (aux.r squared)
	float DE;   // running derivative
	int it;     // formula iterations done
	int status; // 0 running, 1 escaped, 2 iteration limit, 3 non-finite orbit (reverted)
	V3 J[KIND == kBoxDelta ? 3 : 1];
};

template <int KIND>
__device__ __forceinline__ void de_begin(DEState<KIND> &s, V3 p)
{
	s.p = p;
	s.z = p;
	s.r2 = dot(p, p);
	s.DE = 1.0f;
	s.it = 0;
	s.status = 0;
	if (KIND == kBoxDelta)
	{
		s.J[0] = v3(1.0f, 0.0f, 0.0f);
		s.J[1] = v3(0.0f, 1.0f, 0.0f);
		s.J[2] = v3(0.0f, 0.0f, 1.0f);
	}
}

template <int KIND>
__device__ __forceinline__ void de_step(DEState<KIND> &s, const FrameConst &f)
{
	const V3 lastZ = s.z;
	if (KIND == kBulb8)
		MandelbulbPow8Iteration(s.z, s.r2, s.DE);
	else if (KIND == kBulbTrig)
		MandelbulbIteration(s.z, sqrtf(s.r2), s.DE, f);
	else if (KIND == kBoxAnalytic)
		MandelboxIteration<false>(s.z, s.DE, nullptr, f);
	else if (KIND == kBoxDelta)
		MandelboxIteration<true>(s.z, s.DE, s.J, f);
	else
		KaleidoscopicIfsIteration(s.z, s.DE, f);
	s.z = add_c(s.z, s.p, f);
	if (KIND == kBoxDelta && f.addC && !f.julia)
	{
		s.J[0].x += f.constMul[0];
		s.J[1].y += f.constMul[1];
		s.J[2].z += f.constMul[2];
	}
	s.r2 = dot(s.z, s.z);
	s.it++;
	if (!finite3(s.z))
	{
		s.z = lastZ;
		s.r2 = dot(lastZ, lastZ);
		s.status = 3;
	}
	else if (s.r2 > f.bailout * f.bailout)
		s.status = 1;
	else if (s.it >= f.formulaMaxiter)
		s.status = 2;
}

// Distance of a finished evaluation after CalculateDistanceSimple's clamps; `iters` gets the
// reference's iteration count (i + 1, maxN + 1 for a non-escaping orbit).
template <int KIND>
__device__ __forceinline__ float de_finish(DEState<KIND> &s, const FrameConst &f, float detailSize,
	int &iters)
{
	iters = s.status == 2 ? s.it + 1 : s.it;
	const float r = sqrtf(s.r2);
	float d;
	if (KIND == kBoxDelta)
	{
		float dr;
		if (s.status == 1)
		{
			const float inv_r = 1.0f / r;
			const float d1 = fabsf(dot(s.z, s.J[0])) * inv_r;
			const float d2 = fabsf(dot(s.z, s.J[1])) * inv_r;
			const float d3 = fabsf(dot(s.z, s.J[2])) * inv_r;
			dr = sqrtf(d1 * d1 + d2 * d2 + d3 * d3);
		}
		else
		{
			V3 z2 = s.z;
			float DE2 = s.DE;
			MandelboxIteration<false>(z2, DE2, nullptr, f);
			z2 = add_c(z2, s.p, f);
			const float delta = fmaxf(length(s.p) * 1e-14f, 1e-5f * detailSize);
			dr = 1.7320508f * fabsf(length(z2) - r) / delta;
		}
		if (dr > 0.0f)
			d = (f.deFunction == MBX_DE_FUNCTION_LOGARITHMIC) ? 0.5f * r * __logf(r) / dr : 0.5f * r / dr;
		else
			d = r;
	}
	else if (s.DE > 0.0f)
	{
		if (f.analyticFunction == MBX_ANALYTIC_LOGARITHMIC)
			d = (r > 1.0f) ? 0.5f * r * __logf(r) / s.DE : 0.0f;
		else if (f.analyticFunction == MBX_ANALYTIC_IFS)
			d = (r - 2.0f) / s.DE;
		else
			d = r / s.DE;
	}
	else
		d = r;
	if (iters < f.minN && d < detailSize) d = detailSize;
	d = (d < 0.0f) ? 0.0f : d;
	d = (d > 10.0f) ? 10.0f : d;
	if (!(d < 1e20f)) d = 1e20f;
	return d;
}

} // namespace mbx

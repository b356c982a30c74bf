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

// Hardware square root / reciprocal / reciprocal square root (v_sqrt_f32, v_rcp_f32,
// v_rsq_f32: about 1 ulp) instead of the correctly rounded multi-instruction sequences; the
// kernels are compared against an FP64 reference, where these errors are far below tolerance.
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * frcp(b); }

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(V3 a) { return fsqrt(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return a * frsq(dot(a, a)); }
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
	// RandomInt() is in [1, 2^31 - 2]: the unsigned remainder equals the reference's signed one
	return (int)((uint32_t)random_int(seed) % (uint32_t)(max + 1));
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
	const float inv_r = frsq(r2);
	const float inv_rho = frsq(rho2);
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
	const float th0 = asinf(z.z * frcp(r)) + f.beta;
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
	const float t2 = f.mbFR2 * frcp(r2);
	const float k = inner ? f.mbFactor1 : (outer ? t2 : 1.0f);
	if (TANGENT)
	{
		// d[(z + o) t(z)] = t dz - 2 t (z.dz) / r2 (z + o) in the outer regime, k dz otherwise
		const float g = outer ? -2.0f * t2 * frcp(r2) : 0.0f;
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

// ------------------------------------------------------------------------------------------
// FP64 Mandelbox.  Each Mandelbox iteration scales the orbit by |scale| (2 in the benchmark
// scenes), so FP32 rounding doubles every iteration and after ~20 iterations escape counts no
// longer follow the reference's FP64 orbit.  gfx950 runs FP64 FMA at the same rate as
// (unpacked) FP32, so the orbit runs in FP64 and only the final distance returns to FP32.
// ------------------------------------------------------------------------------------------
struct DV3
{
	double x, y, z;
};
__device__ __forceinline__ DV3 dv3(double x, double y, double z) { return DV3{x, y, z}; }
__device__ __forceinline__ DV3 operator+(DV3 a, DV3 b) { return DV3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ DV3 operator-(DV3 a, DV3 b) { return DV3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ DV3 operator*(DV3 a, double s) { return DV3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double ddot(DV3 a, DV3 b)
{
#pragma clang fp contract(off)
	return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ double dsgn(double v) { return (double)((0.0 < v) - (v < 0.0)); }

// With COLOR, aux.color gains the fold colour factors in the reference's order: x, y, z box
// folds, then the inner (Sp1) or outer (Sp2) sphere fold.
// mandelbox.mainRot.RotateVector (algebra.cpp:320-344) in the reference's operation order
__device__ __forceinline__ DV3 rot_main_d(const FrameConst &f, DV3 v)
{
#pragma clang fp contract(off)
	const double *m = f.mbMainRotD;
	return dv3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
		m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

template <bool TANGENT, bool COLOR = false>
__device__ __forceinline__ void MandelboxIterationD(DV3 &z, double &DE, DV3 *J, const FrameConst &f,
	double *color = nullptr)
{
	// no FMA contraction: the orbit expands rounding differences by up to (4 |scale|)^n, so the
	// escape counts follow the reference only with the reference's separately rounded products
#pragma clang fp contract(off)
	const double L = f.mbFoldingLimitD, V = f.mbFoldingValueD;
	const bool fx = fabs(z.x) > L, fy = fabs(z.y) > L, fz = fabs(z.z) > L;
	z.x = fx ? dsgn(z.x) * V - z.x : z.x;
	z.y = fy ? dsgn(z.y) * V - z.y : z.y;
	z.z = fz ? dsgn(z.z) * V - z.z : z.z;
	if (COLOR)
	{
		if (fx) *color += f.mbColorD[0];
		if (fy) *color += f.mbColorD[1];
		if (fz) *color += f.mbColorD[2];
	}
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
	const double r2 = ddot(z, z);
	const DV3 o = dv3(f.mbOffsetD[0], f.mbOffsetD[1], f.mbOffsetD[2]);
	const DV3 zo = z + o;
	const bool inner = r2 < f.mbMR2D;
	const bool outer = !inner && r2 < f.mbFR2D;
	const double t2 = f.mbFR2D / r2;
	const double k = inner ? f.mbFactor1D : (outer ? t2 : 1.0);
	if (COLOR)
	{
		if (inner)
			*color += f.mbSp1D;
		else if (outer)
			*color += f.mbSp2D;
	}
	if (TANGENT)
	{
		const double g = outer ? -2.0 * t2 / r2 : 0.0;
#pragma unroll
		for (int c = 0; c < 3; c++)
		{
			const double zd = ddot(z, J[c]) * g;
			J[c] = J[c] * k + zo * zd;
			if (f.mbMainRotOn) J[c] = rot_main_d(f, J[c]);
			J[c] = J[c] * f.mbScaleD;
		}
	}
	// fractal_mandelbox.cpp:101-105: offset back, main rotation, scale
	DV3 zr = zo * k - o;
	if (f.mbMainRotOn) zr = rot_main_d(f, zr);
	z = zr * f.mbScaleD;
	DE = DE * k * f.mbAbsScaleD + 1.0;
}

// aux.r > bailout with aux.r = |z| as the reference compares it (compute_fractal.cpp:299,317):
// the squared compare decides all but the orbits within rounding of the bailout radius, which
// are exactly the ones the binary search converges to (the DE jumps where the escape count
// changes), so those take the square root.
__device__ __forceinline__ bool escaped_d(double r2, double bailout)
{
	const double b2 = bailout * bailout;
	if (r2 > b2 * (1.0 + 1e-12)) return true;
	if (r2 < b2 * (1.0 - 1e-12)) return false;
	return sqrt(r2) > bailout;
}

__device__ __forceinline__ DV3 add_c_d(DV3 z, DV3 p, const FrameConst &f)
{
#pragma clang fp contract(off)
	if (!f.addC) return z;
	if (f.julia) return z + dv3(f.juliaCD[0], f.juliaCD[1], f.juliaCD[2]);
	return z + dv3(p.x * f.constMulD[0], p.y * f.constMulD[1], p.z * f.constMulD[2]);
}

// Mandelbox orbit in FP64 with the analytic DE (DELTA = false) or the delta DE (DELTA = true).
// The reference's delta DE differences |z_n| over three axis offsets of
// delta = max(|p| 1e-14, 1e-5 detailSize) (calculate_distance.cpp:613-672); the device carries
// the three columns of dz_n/dp along the orbit instead (forward mode) and takes
// dr_k = |z_n . dz_n/dp_k| / |z_n|, the limit of the reference's quotient, so no cancellation
// at screen-space detail sizes.
template <bool DELTA>
__device__ __forceinline__ double compute_box_d(DV3 p, const FrameConst &f, double detailSize, int &iters)
{
	DV3 z = p;
	DV3 J[3] = {dv3(1.0, 0.0, 0.0), dv3(0.0, 1.0, 0.0), dv3(0.0, 0.0, 1.0)};
	double DE = 1.0;
	double r2 = ddot(z, z);
	const int maxN = f.formulaMaxiter;
	const bool addJ = f.addC && !f.julia;
	int i;
	bool escaped = false;
	for (i = 0; i < maxN; i++)
	{
		const DV3 lastZ = z;
		MandelboxIterationD<DELTA>(z, DE, J, f);
		z = add_c_d(z, p, f);
		if (DELTA && addJ)
		{
			J[0].x += f.constMulD[0];
			J[1].y += f.constMulD[1];
			J[2].z += f.constMulD[2];
		}
		r2 = ddot(z, z);
		if (!(r2 <= 1.0e300))
		{
			if (!(isfinite(z.x) && isfinite(z.y) && isfinite(z.z)))
			{
				z = lastZ;
				r2 = ddot(z, z);
			}
			else
				escaped = true;
			break;
		}
		if (escaped_d(r2, f.bailoutD))
		{
			escaped = true;
			break;
		}
	}
	iters = i + 1;
	const double r = sqrt(r2);
	double d;
	if (!DELTA)
	{
		if (DE > 0.0)
		{
			if (f.analyticFunction == MBX_ANALYTIC_LOGARITHMIC)
				d = (r > 1.0) ? 0.5 * r * log(r) / DE : 0.0;
			else if (f.analyticFunction == MBX_ANALYTIC_IFS)
				d = (r - 2.0) / DE;
			else
				d = r / DE;
		}
		else
			d = r;
		return d;
	}
	double dr;
	if (escaped)
	{
		const double d1 = fabs(ddot(z, J[0])) / r;
		const double d2 = fabs(ddot(z, J[1])) / r;
		const double d3 = fabs(ddot(z, J[2])) / r;
		dr = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
	}
	else
	{
		// the reference's offset orbits run one iteration past a non-escaping orbit
		// (forcedMaxiter = maxN + 1): the quotient is ||z_{N+1}| - |z_N|| / delta per axis
		DV3 z2 = z;
		double DE2 = DE;
		MandelboxIterationD<false>(z2, DE2, nullptr, f);
		z2 = add_c_d(z2, p, f);
		const double delta = fmax(sqrt(ddot(p, p)) * 1e-14, f.deltaRel * detailSize);
		dr = 1.7320508075688772 * fabs(sqrt(ddot(z2, z2)) - r) / delta;
	}
	if (dr > 0.0)
		d = (f.deFunction == MBX_DE_FUNCTION_LOGARITHMIC) ? 0.5 * r * log(r) / dr : 0.5 * r / dr;
	else
		d = r;
	return d;
}

// Compute<calcModeDeltaDE1> of the Mandelbox in FP64 (compute_fractal.cpp:67-654): the final z
// of the orbit of p, with bailout, at most maxN formula calls.  Returns the iteration count.
__device__ __forceinline__ int box_orbit_d(DV3 p, const FrameConst &f, int maxN, DV3 &zOut)
{
#pragma clang fp contract(off)
	DV3 z = p;
	double DE = 1.0;
	int i;
	for (i = 0; i < maxN; i++)
	{
		const DV3 lastZ = z;
		MandelboxIterationD<false>(z, DE, nullptr, f);
		z = add_c_d(z, p, f);
		const double r2 = ddot(z, z);
		if (!(r2 <= 1.0e300))
		{
			if (!(isfinite(z.x) && isfinite(z.y) && isfinite(z.z))) z = lastZ;
			break;
		}
		if (escaped_d(r2, f.bailoutD)) break;
	}
	zOut = z;
	return i + 1;
}

// The reference's delta DE of the Mandelbox (calculate_distance.cpp:609-711): one orbit at p,
// then three orbits at p + delta e_k capped at the first orbit's iteration count, dr_k =
// | |z(p + delta e_k)| - |z(p)| | / delta, dr = |(dr_1, dr_2, dr_3)|, distance 0.5 r / dr (or
// 0.5 r log r / dr).  FP64 throughout: the quotient's offset is below FP32 resolution.
__device__ __forceinline__ double compute_box_fd(DV3 p, const FrameConst &f, double detailSize, int &iters,
	int &itersLast, uint32_t &work)
{
#pragma clang fp contract(off)
	const double deltaDE = fmax(sqrt(ddot(p, p)) * 1e-14, f.deltaRel * detailSize);
	DV3 z;
	iters = box_orbit_d(p, f, f.formulaMaxiter, z);
	const double r = sqrt(ddot(z, z));
	const int forced = iters;
	DV3 z1, z2, z3;
	const int i1 = box_orbit_d(dv3(p.x + deltaDE, p.y, p.z), f, forced, z1);
	const int i2 = box_orbit_d(dv3(p.x, p.y + deltaDE, p.z), f, forced, z2);
	const int i3 = box_orbit_d(dv3(p.x, p.y, p.z + deltaDE), f, forced, z3);
	itersLast = i3;
	work += (uint32_t)(iters + i1 + i2 + i3);
	const double dr1 = fabs(sqrt(ddot(z1, z1)) - r) / deltaDE;
	const double dr2 = fabs(sqrt(ddot(z2, z2)) - r) / deltaDE;
	const double dr3 = fabs(sqrt(ddot(z3, z3)) - r) / deltaDE;
	const double dr = sqrt(dr1 * dr1 + dr2 * dr2 + dr3 * dr3);
	if (dr > 0.0)
		return (f.deFunction == MBX_DE_FUNCTION_LOGARITHMIC) ? 0.5 * r * log(r) / dr : 0.5 * r / dr;
	return r;
}

// Mandelbulb orbit in FP64 with the analytic DE, in the reference's operation order:
// cFractalMandelbulb::FormulaCode (fractal_mandelbulb.cpp:29-43) with libm-class asin, atan2,
// pow, sin and cos, then z += c, aux.r = |z|, the NaN revert and the r > bailout test of
// Compute (compute_fractal.cpp:246-321) and the DE finish (compute_fractal.cpp:561-617).
__device__ __forceinline__ void MandelbulbIterationD(DV3 &z, double r, double &DE, const FrameConst &f)
{
#pragma clang fp contract(off)
	const double power = f.powerD;
	const double th0 = asin(z.z / r) + f.betaD;
	const double ph0 = atan2(z.y, z.x) + f.alphaD;
	double rp = pow(r, power - 1.0);
	const double th = th0 * power;
	const double ph = ph0 * power;
	const double cth = cos(th);
	DE = (rp * DE) * power + 1.0;
	rp *= r;
	z = dv3(cth * cos(ph) * rp, cth * sin(ph) * rp, sin(th) * rp);
}

__device__ __forceinline__ double compute_bulb_d(DV3 p, const FrameConst &f, int &iters)
{
#pragma clang fp contract(off)
	DV3 z = p;
	double r = sqrt(ddot(z, z));
	double DE = 1.0;
	const int maxN = f.formulaMaxiter;
	int i;
	for (i = 0; i < maxN; i++)
	{
		const DV3 lastZ = z;
		MandelbulbIterationD(z, r, DE, f);
		z = add_c_d(z, p, f);
		r = sqrt(ddot(z, z));
		if (!(isfinite(z.x) && isfinite(z.y) && isfinite(z.z)))
		{
			z = lastZ;
			r = sqrt(ddot(z, z));
			break;
		}
		if (r > f.bailoutD) break;
	}
	iters = i + 1;
	if (!(DE > 0.0)) return r;
	if (f.analyticFunction == MBX_ANALYTIC_LOGARITHMIC) return (r > 1.0) ? 0.5 * r * log(r) / DE : 0.0;
	if (f.analyticFunction == MBX_ANALYTIC_IFS) return (r - 2.0) / DE;
	return r / DE;
}

} // namespace mbx

#include "mbx_generic.h"

namespace mbx
{

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
// Per-lane state of the DE evaluator: the work counters, and the orbit constants of the
// current frame copied out of FrameConst once per tile.  The constants are wave-uniform, so
// they live in scalar registers and the march loops issue no constant loads.
struct Work
{
	uint32_t de_evals;
	uint32_t iterations;
	uint32_t far_steps;
	int maxN, minN, addC, julia, analyticFunction, nodeRotated;
	float bailout2;
	float cx, cy, cz; // Julia constant, or the constant multiplier of c = p
};

__device__ __forceinline__ void load_orbit_constants(Work &w, const FrameConst &f)
{
	w.maxN = f.formulaMaxiter;
	w.minN = f.minN;
	w.addC = f.addC;
	w.julia = f.julia;
	w.analyticFunction = f.analyticFunction;
	w.nodeRotated = f.nodeRotated;
	w.bailout2 = f.bailout * f.bailout;
	w.cx = f.julia ? f.juliaC[0] : f.constMul[0];
	w.cy = f.julia ? f.juliaC[1] : f.constMul[1];
	w.cz = f.julia ? f.juliaC[2] : f.constMul[2];
}

__device__ __forceinline__ Work make_work(const FrameConst &f)
{
	Work w;
	w.de_evals = 0u;
	w.iterations = 0u;
	w.far_steps = 0u;
	load_orbit_constants(w, f);
	return w;
}

// z + c with c = p * constantMultiplier, or the Julia constant (compute_fractal.cpp:246-281)
__device__ __forceinline__ V3 add_c(V3 z, V3 p, const Work &w)
{
	if (!w.addC) return z;
	if (w.julia) return z + v3(w.cx, w.cy, w.cz);
	return v3(z.x + p.x * w.cx, z.y + p.y * w.cy, z.z + p.z * w.cz);
}

// ln(x) for x > 1 from the hardware log2 (v_log_f32)
__device__ __forceinline__ float flog(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }

// Analytic-DE orbit.  Returns the distance and the iteration count (i + 1, maxN + 1 when the
// orbit never escapes).  kBulb8 is only chosen for c = p * constantMultiplier.
template <int KIND>
__device__ __forceinline__ float compute_analytic(V3 p, const FrameConst &f, const Work &w, int &iters)
{
	V3 z = p;
	float r2 = dot(z, z);
	float DE = 1.0f;
	int i;
	for (i = 0; i < w.maxN; i++)
	{
		const float lastR2 = r2;
		if (KIND == kBulb8)
			MandelbulbPow8Iteration(z, r2, DE);
		else if (KIND == kBulbTrig)
			MandelbulbIteration(z, fsqrt(r2), DE, f);
		else if (KIND == kBoxAnalytic)
			MandelboxIteration<false>(z, DE, nullptr, f);
		else
			KaleidoscopicIfsIteration(z, DE, f);
		if (KIND == kBulb8)
			z = v3(z.x + p.x * w.cx, z.y + p.y * w.cy, z.z + p.z * w.cz);
		else
			z = add_c(z, p, w);
		r2 = dot(z, z);
		// one compare ends the orbit: escape (r2 > bailout^2), or a NaN / infinite component
		// (r2 NaN or +inf).  The rare branch tells an overflowing finite orbit (an escape) from
		// a non-finite one, which reverts to the last z (compute_fractal.cpp:303-309); only
		// |z| is used after the loop.
		if (!(r2 <= w.bailout2))
		{
			if (!(r2 <= 3.0e38f) && !finite3(z)) r2 = lastR2;
			break;
		}
	}
	iters = i + 1;
	const float r = fsqrt(r2);
	// DE finish (compute_fractal.cpp:561-617) as selects rather than branches: every lane of a
	// wave takes it, and the log and reciprocal are needed by almost all of them
	const float invDE = frcp(DE);
	float distance;
	if (KIND == kBulb8 || w.analyticFunction == MBX_ANALYTIC_LOGARITHMIC)
		distance = (r > 1.0f) ? 0.5f * r * flog(r) * invDE : 0.0f;
	else if (w.analyticFunction == MBX_ANALYTIC_IFS)
		distance = (r - 2.0f) * invDE;
	else
		distance = r * invDE;
	return (DE > 0.0f) ? distance : r;
}

// Position precision of a DE kind.  Mandelbox kinds march in FP64 (ray position, scan, normal
// taps, secondary-ray points): the reference's delta DE with its tiny offset is chaotic at FP32
// position resolution (rounding a hit point to FP32 changes the iteration count of ~0.7% of
// Mandelbox hits in the reference's own FP64 arithmetic), and the deep-zoom scenes need
// sub-1e-7 positions.  kBulbD is the Mandelbulb for views whose distance threshold is within a
// few hundred FP32 rounding steps of the coordinates (the host picks it).  FP64 add/mul/FMA run
// at the unpacked FP32 rate on gfx950.
template <int KIND>
struct PosPrecision
{
	static constexpr bool kFP64 = (KIND == kBoxAnalytic || KIND == kBoxDelta || KIND == kBulbD || KIND == kGeneric);
};

template <bool D>
struct Prec;

template <>
struct Prec<false>
{
	using R = float;
	using P = V3;
	static __device__ __forceinline__ P cam(const FrameConst &f) { return v3(f.camera[0], f.camera[1], f.camera[2]); }
	static __device__ __forceinline__ R threshK(const FrameConst &f) { return f.threshK; }
	static __device__ __forceinline__ R deltaK(const FrameConst &f) { return f.deltaK; }
	static __device__ __forceinline__ R DEThresh(const FrameConst &f) { return f.DEThresh; }
	static __device__ __forceinline__ R DEFactor(const FrameConst &f) { return f.DEFactor; }
	static __device__ __forceinline__ R smoothness(const FrameConst &f) { return f.smoothness; }
	static __device__ __forceinline__ R len(P a) { return fsqrt(dot(a, a)); }
	static __device__ __forceinline__ R rcp(R x) { return frcp(x); }
	static __device__ __forceinline__ R thresh(const FrameConst &f, P p) { return len(cam(f) - p) * f.threshK; }
	static __device__ __forceinline__ P along(P start, P dir, R scan) { return start + dir * scan; }
	static __device__ __forceinline__ P from(V3 a) { return a; }
	static __device__ __forceinline__ V3 tof(P a) { return a; }
	static __device__ __forceinline__ P fromD(double x, double y, double z) { return v3((float)x, (float)y, (float)z); }
	static __device__ __forceinline__ P lightDir(const FrameConst &f) { return v3(f.lightDir[0], f.lightDir[1], f.lightDir[2]); }
	static __device__ __forceinline__ P aoVector(const FrameConst &f, int i)
	{
		return v3(f.aoVectors[i][0], f.aoVectors[i][1], f.aoVectors[i][2]);
	}
	// world -> local point of the rotated fractal node (calculate_distance.cpp:296)
	static __device__ __forceinline__ P node_local(const FrameConst &f, P p) { return rot(f.nodeRot, p); }
	static __device__ __forceinline__ R invWidth(const FrameConst &f) { return f.invWidth; }
	static __device__ __forceinline__ R invHeight(const FrameConst &f) { return f.invHeight; }
	static __device__ __forceinline__ R aspect(const FrameConst &f) { return f.aspectRatio; }
	// CalculateViewVector three-point (projection_3d.cpp:77-89), normalised again by doWork
	static __device__ __forceinline__ P view(const FrameConst &f, R ix, R iy)
	{
		const V3 vv = normalize(v3(ix * f.fov, 1.0f, iy * f.fov));
		return normalize(rot(f.mRot, vv));
	}
	// CalculateViewVector before the march normalises it again, and the camera rotation of a vector
	static __device__ __forceinline__ P view_raw(const FrameConst &f, R ix, R iy)
	{
		return rot(f.mRot, normalize(v3(ix * f.fov, 1.0f, iy * f.fov)));
	}
	static __device__ __forceinline__ P cam_rot(const FrameConst &f, P v) { return rot(f.mRot, v); }
	static __device__ __forceinline__ P normalize_p(P v) { return normalize(v); }
	static __device__ __forceinline__ R softRange(const FrameConst &f) { return f.softRange; }
};

template <>
struct Prec<true>
{
	using R = double;
	using P = DV3;
	static __device__ __forceinline__ P cam(const FrameConst &f) { return dv3(f.cameraD[0], f.cameraD[1], f.cameraD[2]); }
	static __device__ __forceinline__ R threshK(const FrameConst &f) { return f.threshKD; }
	static __device__ __forceinline__ R deltaK(const FrameConst &f) { return f.deltaKD; }
	static __device__ __forceinline__ R DEThresh(const FrameConst &f) { return f.DEThreshD; }
	static __device__ __forceinline__ R DEFactor(const FrameConst &f) { return f.DEFactorD; }
	static __device__ __forceinline__ R smoothness(const FrameConst &f) { return f.smoothnessD; }
	static __device__ __forceinline__ R len(P a) { return sqrt(ddot(a, a)); }
	static __device__ __forceinline__ R rcp(R x) { return 1.0 / x; }
	// CalcDistThresh in the reference's operation order: |camera - p| * resolution * fov / detail,
	// without the detail level in iteration-threshold mode (render_worker.cpp:656-661)
	static __device__ __forceinline__ R thresh(const FrameConst &f, P p)
	{
#pragma clang fp contract(off)
		const R t = len(cam(f) - p) * f.resolutionD * f.fovD;
		return (f.iterThreshMode ? t : t / f.detailLevelD) * f.threshScaleD;
	}
	// start + dir * scan without FMA contraction (CVector3 operator* then operator+)
	static __device__ __forceinline__ P along(P start, P dir, R scan)
	{
#pragma clang fp contract(off)
		return dv3(start.x + dir.x * scan, start.y + dir.y * scan, start.z + dir.z * scan);
	}
	static __device__ __forceinline__ P from(V3 a) { return dv3(a.x, a.y, a.z); }
	static __device__ __forceinline__ V3 tof(P a) { return v3((float)a.x, (float)a.y, (float)a.z); }
	static __device__ __forceinline__ P fromD(double x, double y, double z) { return dv3(x, y, z); }
	// secondary-ray directions in FP64, as the reference holds them: sample positions along them
	// must match where the DE is discontinuous (interior and iteration-threshold modes)
	static __device__ __forceinline__ P lightDir(const FrameConst &f) { return dv3(f.lightDirD[0], f.lightDirD[1], f.lightDirD[2]); }
	static __device__ __forceinline__ P aoVector(const FrameConst &f, int i)
	{
		return dv3(f.aoVectorsD[i][0], f.aoVectorsD[i][1], f.aoVectorsD[i][2]);
	}
	// world -> local point of the rotated fractal node (calculate_distance.cpp:296; CMatrix44
	// TransformPoint sums left to right)
	static __device__ __forceinline__ P node_local(const FrameConst &f, P p)
	{
#pragma clang fp contract(off)
		const double *m = f.nodeRotD;
		return dv3(m[0] * p.x + m[1] * p.y + m[2] * p.z, m[3] * p.x + m[4] * p.y + m[5] * p.z,
			m[6] * p.x + m[7] * p.y + m[8] * p.z);
	}
	static __device__ __forceinline__ R invWidth(const FrameConst &f) { return f.invWidthD; }
	static __device__ __forceinline__ R invHeight(const FrameConst &f) { return f.invHeightD; }
	static __device__ __forceinline__ R aspect(const FrameConst &f) { return f.aspectRatioD; }
	static __device__ __forceinline__ P view(const FrameConst &f, R ix, R iy)
	{
#pragma clang fp contract(off)
		DV3 vv = dv3(ix * f.fovD, 1.0, iy * f.fovD);
		vv = vv * (1.0 / sqrt(ddot(vv, vv)));
		const double *m = f.mRotD;
		DV3 d = dv3(m[0] * vv.x + m[1] * vv.y + m[2] * vv.z, m[3] * vv.x + m[4] * vv.y + m[5] * vv.z,
			m[6] * vv.x + m[7] * vv.y + m[8] * vv.z);
		return d * (1.0 / sqrt(ddot(d, d)));
	}
	static __device__ __forceinline__ P view_raw(const FrameConst &f, R ix, R iy)
	{
#pragma clang fp contract(off)
		DV3 vv = dv3(ix * f.fovD, 1.0, iy * f.fovD);
		vv = vv * (1.0 / sqrt(ddot(vv, vv)));
		return cam_rot(f, vv);
	}
	// mRot.RotateVector (CMatrix33 times the vector, each row summed left to right)
	static __device__ __forceinline__ P cam_rot(const FrameConst &f, P v)
	{
#pragma clang fp contract(off)
		const double *m = f.mRotD;
		return dv3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
			m[6] * v.x + m[7] * v.y + m[8] * v.z);
	}
	static __device__ __forceinline__ P normalize_p(P v)
	{
#pragma clang fp contract(off)
		return v * (1.0 / sqrt(ddot(v, v)));
	}
	static __device__ __forceinline__ R softRange(const FrameConst &f) { return f.softRangeD; }
};

// The tail of CalculateDistanceSimple with interior or iteration-threshold mode on
// (calculate_distance.cpp:561-605, 722-759): maxiter is the first orbit's Compute flag.  The
// analytic branch tests that raw flag for the interior; the delta branch a copy kept only in
// iteration-threshold mode ("limits disabled").
__device__ __forceinline__ double distance_modes(const FrameConst &f, double d, double detailSize, int itersMinN,
	int minN, bool maxiter, bool normalCalc)
{
#pragma clang fp contract(off)
	if (f.iterThreshMode && !normalCalc && maxiter) d = 0.0; // out->maxiter
	if (itersMinN < minN && d < detailSize) d = detailSize;
	const bool interiorMaxiter = (f.genDEType == MBX_DE_ANALYTIC) ? maxiter : (f.iterThreshMode && maxiter);
	if (f.interiorMode && !normalCalc)
	{
		if (d < 0.5 * detailSize || interiorMaxiter) d = detailSize;
	}
	else if (f.interiorMode && normalCalc)
	{
		if (d < 0.9 * detailSize) d = detailSize - d;
	}
	if (f.iterThreshMode && !normalCalc && !maxiter && d < detailSize) d = detailSize * 1.01;
	d = (d < 0.0) ? 0.0 : d;
	d = (d > 10.0) ? 10.0 : d;
	if (!(d < 1e20)) d = 1e20;
	return d;
}

// CalculateDistance for the single fractal node (calculate_distance.cpp:225-516 reduces to
// CalculateDistanceSimple for one node with identity transform and scale 1).  normalCalc marks the
// normal-vector taps (sDistanceIn's normalCalculationMode).  Interior and iteration-threshold
// modes run on the general kernel only (the host routes such scenes there).
template <int KIND>
__device__ __forceinline__ typename Prec<PosPrecision<KIND>::kFP64>::R calculate_distance(
	typename Prec<PosPrecision<KIND>::kFP64>::P p, typename Prec<PosPrecision<KIND>::kFP64>::R detailSize,
	const FrameConst &f, int &iters, Work &w, bool normalCalc = false)
{
	using R = typename Prec<PosPrecision<KIND>::kFP64>::R;
	if (w.nodeRotated) p = Prec<PosPrecision<KIND>::kFP64>::node_local(f, p);
	R d;
	int itersMinN = 0;
	if constexpr (KIND == kBoxDelta)
	{
		if (f.deltaTangent)
		{
			d = compute_box_d<true>(p, f, detailSize, iters);
			w.iterations += 4u * (uint32_t)iters;
			itersMinN = iters;
		}
		else
			d = compute_box_fd(p, f, detailSize, iters, itersMinN, w.iterations);
	}
	else if constexpr (KIND == kBoxAnalytic)
	{
		d = compute_box_d<false>(p, f, detailSize, iters);
		w.iterations += (uint32_t)iters;
		itersMinN = iters;
	}
	else if constexpr (KIND == kBulbD)
	{
		d = compute_bulb_d(p, f, iters);
		w.iterations += (uint32_t)iters;
		itersMinN = iters;
	}
	else if constexpr (KIND == kGeneric)
	{
		const GenDistance g = compute_generic(p, f, detailSize, normalCalc ? 1 : 0);
		d = g.d;
		iters = g.iters;
		itersMinN = g.itersLast;
		w.iterations += g.work;
		if (f.interiorMode || f.iterThreshMode)
		{
			w.de_evals++;
			return distance_modes(f, d, detailSize, itersMinN, w.minN, g.maxiter != 0, normalCalc);
		}
	}
	else
	{
		d = compute_analytic<KIND>(p, f, w, iters);
		w.iterations += (uint32_t)iters;
		itersMinN = iters;
	}
	w.de_evals++;
	if (w.minN > 1 && itersMinN < w.minN && d < detailSize) d = detailSize; // iters >= 1 always
	// clamp to [0, 10] as the reference does; a NaN passes the clamp and the boolean-add merge
	// of the objects tree then keeps its initial 1e20
	if constexpr (PosPrecision<KIND>::kFP64)
	{
		d = (d < (R)0) ? (R)0 : d;
		d = (d > (R)10) ? (R)10 : d;
		if (!(d < (R)1e20)) d = (R)1e20;
	}
	else
		d = (d != d) ? 1e20f : fminf(fmaxf(d, 0.0f), 10.0f);
	return d;
}

// Iteration count of Compute<calcModeNormal> at p (compute_fractal.cpp:67-654, iters = i + 1):
// the orbit alone, without CalculateDistance's node transform (slow-shading normals call Compute
// on the grid points directly, shader_calculate_normals.cpp:96-101)
template <int KIND>
__device__ __forceinline__ int orbit_iterations(typename Prec<PosPrecision<KIND>::kFP64>::P p, const FrameConst &f,
	Work &w)
{
	int iters;
	if constexpr (KIND == kBoxDelta)
	{
		DV3 z;
		iters = box_orbit_d(p, f, f.formulaMaxiter, z);
	}
	else if constexpr (KIND == kBoxAnalytic)
		(void)compute_box_d<false>(p, f, 0.0, iters);
	else if constexpr (KIND == kBulbD)
		(void)compute_bulb_d(p, f, iters);
	else if constexpr (KIND == kGeneric)
	{
		ExtendedAux a;
		DV4 z;
		iters = gen_orbit(f, p, f.formulaMaxiter, z, a);
	}
	else
		(void)compute_analytic<KIND>(p, f, w, iters);
	w.iterations += (uint32_t)iters;
	return iters;
}

// ------------------------------------------------------------------------------------------
// Compute<calcModeColouring> (compute_fractal.cpp:67-654, colouring branch :343-436) and
// CalculateColorIndex for a single formula (fractal_coloring.cpp:318-340).  SurfaceColour runs it
// with maxiterMultiplier 4 (shader_surface_color.cpp:51-55).  The orbit's w stays 0 on the device,
// so the 4D measures reduce to the 3D ones and the 4D cross measure to 0.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float colour_measure(const FrameConst &f, V3 z, V3 p, float r)
{
	switch (f.colorAlgorithm)
	{
		case 1: return fabsf(dot(z, p));
		case 2: return fabsf(length(z - p) - f.colorSphereRadius);
		case 3:
		{
			const float ax = fabsf(z.x), ay = fabsf(z.y), az = fabsf(z.z);
			const float len = (ax < ay) ? (ax < az ? ax : az) : (ay < az ? ay : az); // dMin
			return (f.color4D && 0.0f < len) ? 0.0f : len;                          // min(len, |w|)
		}
		case 4: return fabsf(dot(z, v3(f.colorLineDir[0], f.colorLineDir[1], f.colorLineDir[2])));
		default: return r;
	}
}

__device__ __forceinline__ double colour_measure_d(const FrameConst &f, DV3 z, DV3 p, double r)
{
#pragma clang fp contract(off)
	switch (f.colorAlgorithm)
	{
		case 1: return fabs(ddot(z, p));
		case 2: return fabs(sqrt(ddot(z - p, z - p)) - f.colorSphereRadiusD);
		case 3:
		{
			const double ax = fabs(z.x), ay = fabs(z.y), az = fabs(z.z);
			const double len = (ax < ay) ? (ax < az ? ax : az) : (ay < az ? ay : az);
			return (f.color4D && 0.0 < len) ? 0.0 : len;
		}
		case 4: return fabs(ddot(z, dv3(f.colorLineDirD[0], f.colorLineDirD[1], f.colorLineDirD[2])));
		default: return r;
	}
}

// FP32 kinds (Mandelbulb, IFS): colorMin of the orbit, bailout at the formula's radius (or
// 1e15 and the stall test before v2.15); colour index colorMin * 5000 (Mandelbulb) or * 1000 (IFS)
template <int KIND>
__device__ double colour_index_f(V3 p, const FrameConst &f, const Work &w)
{
	V3 z = p;
	float r2 = dot(z, z);
	float DE = 1.0f;
	float colorMin = 1000.0f;
	const int maxN = 4 * w.maxN;
	const bool pre = f.colorPreV215 != 0;
	for (int i = 0; i < maxN; i++)
	{
		const V3 lastZ = z;
		if (KIND == kBulb8)
			MandelbulbPow8Iteration(z, r2, DE);
		else if (KIND == kBulbTrig)
			MandelbulbIteration(z, fsqrt(r2), DE, f);
		else
			KaleidoscopicIfsIteration(z, DE, f);
		z = add_c(z, p, w);
		r2 = dot(z, z);
		if (!finite3(z))
		{
			z = lastZ;
			break;
		}
		const float r = fsqrt(r2);
		const float len = colour_measure(f, z, p, r);
		if (len < colorMin) colorMin = len;
		if (!pre)
		{
			if (r > f.bailout) break;
		}
		else if (r > 1e15f || length(z - lastZ) / r < 1e-15f)
			break;
	}
	return (KIND == kIfsAnalytic) ? (double)colorMin * 1000.0 : (double)colorMin * 5000.0;
}

// FP64 kinds (Mandelbox, FP64 Mandelbulb) in the reference's operation order.  The Mandelbox
// index is aux.color * 100 + r * factorR / 1e13 (+ colorMin * 1000 for a non-standard measure);
// with the standard measure its orbit runs to |z| > 1e15 or until it stalls.
template <int KIND>
__device__ double colour_index_d(DV3 p, const FrameConst &f)
{
#pragma clang fp contract(off)
	constexpr bool box = KIND == kBoxAnalytic || KIND == kBoxDelta;
	DV3 z = p;
	double r = sqrt(ddot(z, z));
	double DE = 1.0, color = 1.0, colorMin = 1000.0;
	const int maxN = 4 * f.formulaMaxiter;
	const int alg = f.colorAlgorithm;
	const bool pre = f.colorPreV215 != 0;
	for (int i = 0; i < maxN; i++)
	{
		const DV3 lastZ = z;
		if constexpr (box)
			MandelboxIterationD<false, true>(z, DE, nullptr, f, &color);
		else
			MandelbulbIterationD(z, r, DE, f);
		z = add_c_d(z, p, f);
		r = sqrt(ddot(z, z));
		if (!(isfinite(z.x) && isfinite(z.y) && isfinite(z.z)))
		{
			z = lastZ;
			r = sqrt(ddot(z, z));
			break;
		}
		const double len = colour_measure_d(f, z, p, r);
		const DV3 dz = z - lastZ;
		const double change = sqrt(ddot(dz, dz)) / r;
		if (!pre)
		{
			if (!box)
			{
				if (len < colorMin) colorMin = len;
				if (r > f.bailoutD) break;
			}
			else if (alg == 0)
			{
				if (r > 1e15 || change < 1e-15) break;
			}
			else
			{
				if (len < colorMin) colorMin = len;
				if (r > f.bailoutD || change < 1e-15) break;
			}
		}
		else if (!box || alg != 0)
		{
			if (len < colorMin) colorMin = len;
			if (r > 1e15 || change < 1e-15) break;
		}
		else if (r > 1e15 || change < 1e-15)
			break;
	}
	if (box) return color * 100.0 + r * f.mbColorRD / 1e13 + ((alg != 0) ? colorMin * 1000.0 : 0.0);
	return colorMin * 5000.0;
}

// Colour index of a fractal-local point in the kind's position precision
template <int KIND>
__device__ __forceinline__ double colour_index(typename Prec<PosPrecision<KIND>::kFP64>::P p,
	const FrameConst &f, const Work &w)
{
	if constexpr (KIND == kGeneric)
		return colour_index_generic(p, f);
	else if constexpr (PosPrecision<KIND>::kFP64)
		return colour_index_d<KIND>(p, f);
	else
		return colour_index_f<KIND>(p, f, w);
}

// cColorGradient::GetColorFloat without smoothing (color_gradient.cpp:128-136, 180-229): the
// stop below the position, then linear interpolation of the 8-bit channels / 256; the last
// stop and a zero-width interval give the 8-bit values unscaled, as the reference does
__device__ __forceinline__ V3 gradient_colour(const FrameConst &f, float pos)
{
#pragma clang fp contract(off)
	const int n = f.gradientCount;
	int idx = 0;
	while (idx < n - 1 && pos > f.gradientPos[idx + 1]) idx++;
	if (idx == n - 1)
		return v3(f.gradientColor[idx - 1][0], f.gradientColor[idx - 1][1], f.gradientColor[idx - 1][2]);
	const float p1 = f.gradientPos[idx], p2 = f.gradientPos[idx + 1];
	const V3 c1 = v3(f.gradientColor[idx][0], f.gradientColor[idx][1], f.gradientColor[idx][2]);
	const V3 c2 = v3(f.gradientColor[idx + 1][0], f.gradientColor[idx + 1][1], f.gradientColor[idx + 1][2]);
	if (!(p2 - p1 > 0.0f)) return c1;
	const float d = (pos - p1) / (p2 - p1);
	const float nd = 1.0f - d;
	return v3((c1.x * nd + c2.x * d) / 256.0f, (c1.y * nd + c2.y * d) / 256.0f, (c1.z * nd + c2.z * d) / 256.0f);
}

} // namespace mbx

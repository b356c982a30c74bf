// mbx_generic.h - the general formula path (DE kind kGeneric): every formula of the table, alone
// or in a hybrid sequence, in FP64 with a 4D orbit and the reference's operation order
// (no FMA contraction).  The formula of iteration i is FrameConst::genSeq[i]; its container
// (parameters, weight, bailout, c addition) is FrameConst::gen[].  Included by mbx_device.h.
#pragma once

#include <hip/hip_runtime.h>

#include "mbx_frame.h"

namespace mbx
{

struct DV4
{
	double x, y, z, w;
};
__device__ __forceinline__ DV4 dv4(double x, double y, double z, double w) { return DV4{x, y, z, w}; }
__device__ __forceinline__ double dot4(DV4 a, DV4 b)
{
#pragma clang fp contract(off)
	return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
}
// CVector4::Length (algebra.hpp:401)
__device__ __forceinline__ double len4(DV4 a) { return sqrt(dot4(a, a)); }
__device__ __forceinline__ bool finite4(DV4 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z) && isfinite(a.w); }

// sExtendedAux (fractal.h:77-104) with the reference's field names: the state one formula
// iteration reads and updates.  Fields that no formula of the table reads cost nothing: the
// compiler drops the stores to them.
struct ExtendedAux
{
	DV4 c;       // variable c
	DV4 const_c; // constant c
	DV4 old_z;   // hybrid colour 2
	double pos_neg;
	double r;    // |z|, refreshed after every iteration
	double DE;   // running derivative
	double DE0;
	double dist;
	double pseudoKleinianDE;
	double actualScale; // Mandelbox vary scale
	double actualScaleA;
	double color;
	double colorHybrid;
	double temp1000;
	int i;       // iteration index
};

// Compute's initial aux for the orbit of z (compute_fractal.cpp:107-132); actualScale is the
// Mandelbox scale of the sequence's first container
__device__ __forceinline__ void init_aux(ExtendedAux &aux, DV4 z, double actualScale)
{
	aux.c = z;
	aux.const_c = z;
	aux.old_z = z;
	aux.pos_neg = 1.0;
	aux.r = len4(z);
	aux.DE = 1.0;
	aux.DE0 = 0.0;
	aux.dist = 1000.0;
	aux.pseudoKleinianDE = 1.0;
	aux.actualScale = actualScale;
	aux.actualScaleA = 0.0;
	aux.color = 1.0;
	aux.colorHybrid = 0.0;
	aux.temp1000 = 1000.0;
	aux.i = 0;
}

// 3x3 rotation of xyz, w kept; identity matrices skip the multiply (algebra.cpp:320-344)
__device__ __forceinline__ DV4 rot4(const double *m, int identity, DV4 v)
{
#pragma clang fp contract(off)
	if (identity) return v;
	return dv4(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
		m[6] * v.x + m[7] * v.y + m[8] * v.z, v.w);
}

// ---- formulas -------------------------------------------------------------------------------

// cFractalMandelbulb::FormulaCode (fractal_mandelbulb.cpp:29-43)
__device__ __forceinline__ void MandelbulbIteration(DV4 &z, const FractalParams *fractal, ExtendedAux &aux)
{
#pragma clang fp contract(off)
	const double th0 = asin(z.z / aux.r) + fractal->beta;
	const double ph0 = atan2(z.y, z.x) + fractal->alpha;
	double rp = pow(aux.r, fractal->power - 1.0);
	const double th = th0 * fractal->power;
	const double ph = ph0 * fractal->power;
	const double cth = cos(th);
	aux.DE = (rp * aux.DE) * fractal->power + 1.0;
	rp *= aux.r;
	z.x = cth * cos(ph) * rp;
	z.y = cth * sin(ph) * rp;
	z.z = sin(th) * rp;
}

// cFractalMandelbulbPower2::FormulaCode (fractal_mandelbulb_power2.cpp:29-44)
__device__ __forceinline__ void MandelbulbPower2Iteration(DV4 &z, const FractalParams *fractal, ExtendedAux &aux)
{
#pragma clang fp contract(off)
	(void)fractal; // no parameters
	aux.DE = aux.DE * 2.0 * aux.r;
	const double x2 = z.x * z.x, y2 = z.y * z.y, z2 = z.z * z.z;
	const double temp = 1.0 - z2 / (x2 + y2);
	const double newx = (x2 - y2) * temp;
	const double newy = 2.0 * z.x * z.y * temp;
	const double newz = -2.0 * z.z * sqrt(x2 + y2);
	z.x = newx;
	z.y = newy;
	z.z = newz;
}

// cFractalHypercomplex::FormulaCode (fractal_hypercomplex.cpp:31-44)
__device__ __forceinline__ void HypercomplexIteration(DV4 &z, const FractalParams *fractal, ExtendedAux &aux)
{
#pragma clang fp contract(off)
	(void)fractal; // no parameters
	aux.DE = aux.DE * 2.0 * aux.r;
	const double newx = z.x * z.x - z.y * z.y - z.z * z.z - z.w * z.w;
	const double newy = 2.0 * z.x * z.y - 2.0 * z.w * z.z;
	const double newz = 2.0 * z.x * z.z - 2.0 * z.y * z.w;
	const double neww = 2.0 * z.x * z.w - 2.0 * z.y * z.z;
	z = dv4(newx, newy, newz, neww);
}

// cFractalQuaternion::FormulaCode (fractal_quaternion.cpp:29-40)
__device__ __forceinline__ void QuaternionIteration(DV4 &z, const FractalParams *fractal, ExtendedAux &aux)
{
#pragma clang fp contract(off)
	(void)fractal; // no parameters
	aux.DE = aux.DE * 2.0 * aux.r;
	const double newx = z.x * z.x - z.y * z.y - z.z * z.z;
	const double newy = 2.0 * z.x * z.y;
	const double newz = 2.0 * z.x * z.z;
	z.x = newx;
	z.y = newy;
	z.z = newz;
}

// cFractalMengerSponge::FormulaCode (fractal_menger_sponge.cpp:30-71), fold colour off
__device__ __forceinline__ void MengerSpongeIteration(DV4 &z, const FractalParams *fractal, ExtendedAux &aux)
{
#pragma clang fp contract(off)
	z = dv4(fabs(z.x), fabs(z.y), fabs(z.z), fabs(z.w));
	double t;
	if (z.x < z.y)
	{
		t = z.y;
		z.y = z.x;
		z.x = t;
	}
	if (z.x < z.z)
	{
		t = z.z;
		z.z = z.x;
		z.x = t;
	}
	if (z.y < z.z)
	{
		t = z.z;
		z.z = z.y;
		z.y = t;
	}
	z = dv4(z.x * fractal->scale3, z.y * fractal->scale3, z.z * fractal->scale3, z.w * fractal->scale3);
	z.x -= 2.0;
	z.y -= 2.0;
	if (z.z > 1.0) z.z -= 2.0;
	aux.DE *= fractal->scale3;
}

// cFractalMandelbox::FormulaCode (fractal_mandelbox.cpp:31-107) with the fold colour of aux.color
__device__ __forceinline__ void MandelboxIteration(DV4 &z, const FractalParams *fractal, ExtendedAux &aux)
{
#pragma clang fp contract(off)
	const double L = fractal->mbFoldingLimit, V = fractal->mbFoldingValue;
	if (fractal->mbRotOn)
	{
		// rotated folds (fractal_mandelbox.cpp:33-63): for each axis, the fold of the axis rotated
		// by rot[0] (positive side) or else rot[1] (negative side), rotated back by its transpose
		for (int dim = 0; dim < 3; dim++)
		{
			DV4 zRot = rot4(fractal->mbRot[0][dim], fractal->mbRotId[0][dim], z);
			double zr = dim == 0 ? zRot.x : (dim == 1 ? zRot.y : zRot.z);
			if (zr > L)
			{
				zr = V - zr;
				if (dim == 0) zRot.x = zr;
				else if (dim == 1) zRot.y = zr;
				else zRot.z = zr;
				z = rot4(fractal->mbRotInv[0][dim], 0, zRot);
				aux.color += fractal->mbColor[dim];
			}
			else
			{
				zRot = rot4(fractal->mbRot[1][dim], fractal->mbRotId[1][dim], z);
				zr = dim == 0 ? zRot.x : (dim == 1 ? zRot.y : zRot.z);
				if (zr < -L)
				{
					zr = -V - zr;
					if (dim == 0) zRot.x = zr;
					else if (dim == 1) zRot.y = zr;
					else zRot.z = zr;
					z = rot4(fractal->mbRotInv[1][dim], 0, zRot);
					aux.color += fractal->mbColor[dim];
				}
			}
		}
	}
	else
	{
	if (fabs(z.x) > L)
	{
		z.x = dsgn(z.x) * V - z.x;
		aux.color += fractal->mbColor[0];
	}
	if (fabs(z.y) > L)
	{
		z.y = dsgn(z.y) * V - z.y;
		aux.color += fractal->mbColor[1];
	}
	if (fabs(z.z) > L)
	{
		z.z = dsgn(z.z) * V - z.z;
		aux.color += fractal->mbColor[2];
	}
	}
	const double r2 = dot4(z, z);
	z = dv4(z.x + fractal->mbOffset[0], z.y + fractal->mbOffset[1], z.z + fractal->mbOffset[2], z.w + 0.0);
	if (r2 < fractal->mbMR2)
	{
		z = dv4(z.x * fractal->mbFactor1, z.y * fractal->mbFactor1, z.z * fractal->mbFactor1, z.w * fractal->mbFactor1);
		aux.DE *= fractal->mbFactor1;
		aux.color += fractal->mbSp1;
	}
	else if (r2 < fractal->mbFR2)
	{
		const double t = fractal->mbFR2 / r2;
		z = dv4(z.x * t, z.y * t, z.z * t, z.w * t);
		aux.DE *= t;
		aux.color += fractal->mbSp2;
	}
	z = dv4(z.x - fractal->mbOffset[0], z.y - fractal->mbOffset[1], z.z - fractal->mbOffset[2], z.w - 0.0);
	if (fractal->mbMainRotOn) z = rot4(fractal->mbMainRot, 0, z);
	z = dv4(z.x * fractal->mbScale, z.y * fractal->mbScale, z.z * fractal->mbScale, z.w * fractal->mbScale);
	aux.DE = aux.DE * fractal->mbAbsScale + 1.0;
}

// cFractalKaleidoscopicIfs::FormulaCode (fractal_kaleidoscopic_ifs.cpp:30-74)
__device__ __forceinline__ void KaleidoscopicIfsIteration(DV4 &z, const FractalParams *fractal, ExtendedAux &aux)
{
#pragma clang fp contract(off)
	if (fractal->ifsAbs[0]) z.x = fabs(z.x);
	if (fractal->ifsAbs[1]) z.y = fabs(z.y);
	if (fractal->ifsAbs[2]) z.z = fabs(z.z);
	for (int i = 0; i < MBX_IFS_VECTOR_COUNT; i++)
	{
		if (!fractal->ifsEnabled[i]) continue;
		z = rot4(fractal->ifsRot[i], fractal->ifsRotId[i], z);
		const DV4 d = dv4(fractal->ifsDirection[i][0], fractal->ifsDirection[i][1], fractal->ifsDirection[i][2], fractal->ifsDirection[i][3]);
		const double length = dot4(z, d);
		if (length < fractal->ifsDistance[i])
		{
			const double k = 2.0 * (length - fractal->ifsDistance[i]) * fractal->ifsIntensity[i];
			z = dv4(z.x - d.x * k, z.y - d.y * k, z.z - d.z * k, z.w - d.w * k);
		}
	}
	const DV4 off = dv4(fractal->ifsOffset[0], fractal->ifsOffset[1], fractal->ifsOffset[2], 0.0);
	if (fractal->ifsRotationEnabled)
	{
		DV4 t = rot4(fractal->ifsMainRot, fractal->ifsMainRotId, dv4(z.x - off.x, z.y - off.y, z.z - off.z, z.w - off.w));
		z = dv4(t.x + off.x, t.y + off.y, t.z + off.z, t.w + off.w);
	}
	if (fractal->ifsEdgeEnabled)
	{
		if (fractal->ifsEdge[0] > 0.0) z.x = fractal->ifsEdge[0] - fabs(fractal->ifsEdge[0] - z.x);
		if (fractal->ifsEdge[1] > 0.0) z.y = fractal->ifsEdge[1] - fabs(fractal->ifsEdge[1] - z.y);
		if (fractal->ifsEdge[2] > 0.0) z.z = fractal->ifsEdge[2] - fabs(fractal->ifsEdge[2] - z.z);
	}
	z = dv4(z.x * fractal->ifsScale, z.y * fractal->ifsScale, z.z * fractal->ifsScale, z.w * fractal->ifsScale);
	const double sm1 = fractal->ifsScale - 1.0;
	if (fractal->ifsMenger)
	{
		z.x -= off.x * sm1;
		z.y -= off.y * sm1;
		if (z.z > 0.5 * off.z * sm1) z.z -= off.z * sm1;
	}
	else
		z = dv4(z.x - off.x * sm1, z.y - off.y * sm1, z.z - off.z * sm1, z.w - off.w * sm1);
	aux.DE *= fabs(fractal->ifsScale);
}

// cFractalXenodreambuie::FormulaCode (fractal_xenodreambuie.cpp:30-47)
__device__ __forceinline__ void XenodreambuieIteration(DV4 &z, const FractalParams *fractal, ExtendedAux &aux)
{
#pragma clang fp contract(off)
	double rp = pow(aux.r, fractal->power - 1.0);
	aux.DE = rp * aux.DE * fractal->power + 1.0;
	rp *= aux.r;
	const double th = (atan2(z.y, z.x) + fractal->beta) * fractal->power;
	double ph = acos(z.z / aux.r) + fractal->alpha;
	if (fabs(ph) > 0.5 * M_PI) ph = dsgn(ph) * M_PI - ph;
	ph *= fractal->power;
	const double sph = sin(ph);
	z.x = rp * cos(th) * sph;
	z.y = rp * sin(th) * sph;
	z.z = rp * cos(ph);
}

// cFractalBoxFoldBulbPow2::FormulaCode (fractal_box_fold_bulb_pow2.cpp:28-77)
__device__ __forceinline__ void BoxFoldBulbPow2Iteration(DV4 &z, const FractalParams *fractal, ExtendedAux &aux)
{
#pragma clang fp contract(off)
	const double ff = fractal->fipFoldFactor;
	if (fabs(z.x) > ff) z.x = dsgn(z.x) * ff * 2.0 - z.x;
	if (fabs(z.y) > ff) z.y = dsgn(z.y) * ff * 2.0 - z.y;
	if (fabs(z.z) > ff) z.z = dsgn(z.z) * ff * 2.0 - z.z;
	const double fR2 = 1.0, mR2 = 0.25;
	const double r2 = dot4(z, z);
	const double factor1 = fR2 / mR2;
	if (r2 < mR2)
	{
		z = dv4(z.x * factor1, z.y * factor1, z.z * factor1, z.w * factor1);
		aux.DE *= factor1;
	}
	else if (r2 < fR2)
	{
		const double factor2 = fR2 / r2;
		z = dv4(z.x * factor2, z.y * factor2, z.z * factor2, z.w * factor2);
		aux.DE *= factor2;
	}
	z = dv4(z.x * 2.0, z.y * 2.0, z.z * 2.0, z.w * 2.0);
	const double x2 = z.x * z.x, y2 = z.y * z.y, z2 = z.z * z.z;
	const double temp = 1.0 - z2 / (x2 + y2);
	DV4 t;
	t.x = (x2 - y2) * temp;
	t.y = 2.0 * z.x * z.y * temp;
	t.z = -2.0 * z.z * sqrt(x2 + y2);
	t.w = z.w;
	z = t;
	z.z *= fractal->fipZFactor;
	if (fractal->adeEnabledFalse)
		aux.DE = (aux.DE + 1.0) * 5.0 * aux.r * fractal->adeScale1 * sqrt(fractal->fipZFactor * fractal->fipZFactor + 2.0 + fractal->adeOffset2) +
			fractal->adeOffset1;
}

// cFractalAexion::FormulaCode (fractal_aexion.cpp:29-60): the first iteration replaces the
// constant (aux.const_c) with a folded copy of the point
__device__ __forceinline__ void AexionIteration(DV4 &z, const FractalParams *fractal, ExtendedAux &aux)
{
#pragma clang fp contract(off)
	DV4 temp = z;
	double t;
	if (aux.i == 0)
	{
		t = fractal->cadd;
		temp.x = z.x + z.y + z.z;
		temp.y = -z.x - z.y + z.z;
		temp.z = -z.x + z.y - z.z;
		temp.w = z.x - z.y - z.z;
		z = dv4(fabs(temp.x) + t, fabs(temp.y) + t, fabs(temp.z) + t, fabs(temp.w) + t);
		aux.const_c = z;
	}
	t = 2.0 * z.w * z.z;
	temp.x = z.x * z.x - z.y * z.y;
	temp.y = t - temp.x;
	temp.x += t;
	t = 2.0 * z.x * z.y;
	temp.z = z.z * z.z - z.w * z.w;
	temp.w = t - temp.z;
	temp.z += t;
	z = dv4(temp.x + aux.const_c.x, temp.y + aux.const_c.y, temp.z + aux.const_c.z, temp.w + aux.const_c.w);
	if (fractal->adeEnabled)
	{
		double de1 = 1.1 * aux.r;
		de1 = de1 + (len4(z) / aux.r - de1) * fractal->adeScale1;
		aux.DE = de1 * 2.0 * aux.DE + fractal->adeOffset1;
	}
}

// The formulas with FP64 libm calls (pow, asin / acos, atan2, sin, cos) run out of line: inlined,
// their temporaries would set the register count of every caller of the orbit loop.  The orbit
// state goes in and out by value, in registers.
struct GenZ
{
	DV4 z;
	double DE;
};
__device__ __attribute__((noinline)) GenZ gen_mandelbulb_call(const FractalParams &slot, DV4 z, double r, double DE)
{
	const FractalParams &s = uniform_ref(slot);
	ExtendedAux a;
	a.r = r;
	a.DE = DE;
	MandelbulbIteration(z, &s, a);
	return GenZ{z, a.DE};
}
__device__ __attribute__((noinline)) GenZ gen_xenodreambuie_call(const FractalParams &slot, DV4 z, double r, double DE)
{
	const FractalParams &s = uniform_ref(slot);
	ExtendedAux a;
	a.r = r;
	a.DE = DE;
	XenodreambuieIteration(z, &s, a);
	return GenZ{z, a.DE};
}

// cAbstractFractal::FormulaCode of the container's formula (abstract_fractal.h:82): the formula
// table of the device.  A formula added here also needs its metadata row in the host's table
// (mbx_formula_get) and its parameters in FractalParams.
__device__ __forceinline__ void formula_iteration(const FractalParams &s, DV4 &z, ExtendedAux &a)
{
	switch (s.formula)
	{
		case MBX_FORMULA_MANDELBULB:
		{
			const GenZ o = gen_mandelbulb_call(s, z, a.r, a.DE);
			z = o.z;
			a.DE = o.DE;
			break;
		}
		case MBX_FORMULA_MANDELBULB_POWER2: MandelbulbPower2Iteration(z, &s, a); break;
		case MBX_FORMULA_HYPERCOMPLEX: HypercomplexIteration(z, &s, a); break;
		case MBX_FORMULA_QUATERNION: QuaternionIteration(z, &s, a); break;
		case MBX_FORMULA_MENGER_SPONGE: MengerSpongeIteration(z, &s, a); break;
		case MBX_FORMULA_MANDELBOX: MandelboxIteration(z, &s, a); break;
		case MBX_FORMULA_KALEIDOSCOPIC_IFS: KaleidoscopicIfsIteration(z, &s, a); break;
		case MBX_FORMULA_XENODREAMBUIE:
		{
			const GenZ o = gen_xenodreambuie_call(s, z, a.r, a.DE);
			z = o.z;
			a.DE = o.DE;
			break;
		}
		case MBX_FORMULA_BOX_FOLD_BULB_POW2: BoxFoldBulbPow2Iteration(z, &s, a); break;
		case MBX_FORMULA_AEXION: AexionIteration(z, &s, a); break;
		default: break;
	}
}

// ---- one iteration of the sequence (compute_fractal.cpp:170-321) ------------------------------

// BoxFolding and SphericalFolding (legacy_fractal_transforrms.cpp:43-101) before the formula, each
// followed by aux.r = |z| (compute_fractal.cpp:192-202)
__device__ __forceinline__ void gen_global_foldings(const FrameConst &f, DV4 &z, ExtendedAux &a)
{
#pragma clang fp contract(off)
	if (f.boxFold)
	{
		const double L = f.boxFoldLimitD, V = f.boxFoldValueD;
		if (z.x > L) { z.x = V - z.x; a.color *= 0.9; }
		else if (z.x < -L) { z.x = -V - z.x; a.color *= 0.9; }
		if (z.y > L) { z.y = V - z.y; a.color *= 0.9; }
		else if (z.y < -L) { z.y = -V - z.y; a.color *= 0.9; }
		if (z.z > L) { z.z = V - z.z; a.color *= 0.9; }
		else if (z.z < -L) { z.z = -V - z.z; a.color *= 0.9; }
		a.r = len4(z);
	}
	if (f.sphFold)
	{
		const double r2 = a.r * a.r;
		if (r2 < f.sphMR2D)
		{
			const double k = f.sphFactor1D;
			z = dv4(z.x * k, z.y * k, z.z * k, z.w * k);
			a.DE *= k;
			a.color *= 0.9;
		}
		else if (r2 < f.sphFR2D)
		{
			const double k = f.sphFR2D / r2;
			z = dv4(z.x * k, z.y * k, z.z * k, z.w * k);
			a.DE *= k;
			a.color *= 0.9;
		}
		a.r = len4(z);
	}
}

// formula, z + c, hybrid weight; returns the iteration's container
__device__ __forceinline__ const FractalParams &gen_step(const FrameConst &f, int i, DV4 &z, ExtendedAux &a)
{
#pragma clang fp contract(off)
	const FractalParams &s = f.gen[i < f.genSeqLength ? f.genSeq[i] : f.genSeq[0]];
	if (f.boxFold || f.sphFold) gen_global_foldings(f, z, a);
	const DV4 tempZ = z;
	const double tempDE = a.DE, tempColor = a.color;
	a.i = i;
	if (s.weight > 0.0) formula_iteration(s, z, a);
	if (s.addC)
	{
		if (f.julia)
			z = dv4(z.x + f.juliaCD[0], z.y + f.juliaCD[1], z.z + f.juliaCD[2], z.w + 0.0);
		else
			z = dv4(z.x + a.const_c.x * f.constMulD[0], z.y + a.const_c.y * f.constMulD[1],
				z.z + a.const_c.z * f.constMulD[2], z.w + a.const_c.w);
	}
	if (f.genHybrid && s.weight < 1.0)
	{
		// SmoothCVector (common_math.cpp:132-163) and the linear DE / colour blend
		const double k = s.weight, kn = 1.0 - k;
		if (k <= 0.0)
			z = tempZ;
		else
		{
			const double lenInterp = len4(tempZ) * kn + len4(z) * k;
			const DV4 v = dv4(tempZ.x * kn + z.x * k, tempZ.y * kn + z.y * k, tempZ.z * kn + z.z * k,
				tempZ.w * kn + z.w * k);
			const double lv = len4(v);
			z = (lv > 0.0) ? dv4(v.x / lv * lenInterp, v.y / lv * lenInterp, v.z / lv * lenInterp, v.w / lv * lenInterp)
							: tempZ;
		}
		a.DE = a.DE * k + tempDE * kn;
		a.color = a.color * k + tempColor * kn;
	}
	a.r = len4(z);
	return s;
}

// Compute<calcModeNormal / calcModeDeltaDE1> (compute_fractal.cpp:67-654): the orbit of p with
// at most maxN iterations; returns the iteration count (i + 1) and leaves z and aux.  maxiter is
// Compute's flag: set unless the orbit bailed out (a non-finite z also leaves it set).
__device__ __forceinline__ int gen_orbit(const FrameConst &f, DV3 p, int maxN, DV4 &zOut, ExtendedAux &a,
	bool &maxiter)
{
	DV4 z = dv4(p.x, p.y, p.z, f.genInitialW);
	init_aux(a, z, f.gen[f.genSeq[0]].mbScale);
	maxiter = true;
	int i;
	for (i = 0; i < maxN; i++)
	{
		const DV4 lastZ = z;
		const FractalParams &s = gen_step(f, i, z, a);
		if (!finite4(z))
		{
			z = lastZ;
			a.r = len4(z);
			break;
		}
		if (s.checkBailout && a.r > s.bailout)
		{
			maxiter = false;
			break;
		}
	}
	zOut = z;
	return i + 1;
}
__device__ __forceinline__ int gen_orbit(const FrameConst &f, DV3 p, int maxN, DV4 &zOut, ExtendedAux &a)
{
	bool maxiter;
	return gen_orbit(f, p, maxN, zOut, a, maxiter);
}

// What CalculateDistanceSimple needs of the general kernel's orbits: the DE, the first orbit's
// iteration count and maxiter flag, the last orbit's count (the delta DE's minN rule reads the
// z-offset orbit, calculate_distance.cpp:727) and the iterations done
struct GenDistance
{
	double d;
	int iters, itersLast, maxiter;
	uint32_t work;
};

// The distance of the sequence at the fractal-local point p: the analytic finish
// (compute_fractal.cpp:493-617) or the reference's 4-orbit delta DE (calculate_distance.cpp:609-711).
// One out-of-line copy serves every call site of the kernels, and the delta DE's four orbits
// (p, then p + delta along x, y, z with the first orbit's iteration count as their limit) run
// through one copy of the orbit loop: the code and the registers of one orbit.  normalCalc: the
// normal taps of iteration-threshold mode run 5 x maxiter (calculate_distance.cpp:541).
__device__ __attribute__((noinline)) GenDistance compute_generic(DV3 p, const FrameConst &frame, double detailSize,
	int normalCalc)
{
#pragma clang fp contract(off)
	const FrameConst &f = uniform_ref(frame);
	const bool analytic = f.genDEType == MBX_DE_ANALYTIC;
	const double deltaDE = analytic ? 0.0 : fmax(sqrt(ddot(p, p)) * 1e-14, f.deltaRel * detailSize);
	const int orbits = analytic ? 1 : 4;
	const int maxN = f.formulaMaxiter * ((normalCalc && f.iterThreshMode) ? 5 : 1);
	int iters = 0, itersLast = 0;
	bool maxiterFirst = true;
	uint32_t total = 0;
	double r = 0.0, DE = 0.0, r3 = 0.0, drSq = 0.0;
#pragma unroll 1
	for (int k = 0; k < orbits; k++)
	{
		DV3 q = p;
		if (k == 1) q.x = p.x + deltaDE;
		if (k == 2) q.y = p.y + deltaDE;
		if (k == 3) q.z = p.z + deltaDE;
		ExtendedAux a;
		DV4 z;
		bool maxiter;
		const int it = gen_orbit(f, q, k == 0 ? maxN : iters, z, a, maxiter);
		total += (uint32_t)it;
		itersLast = it;
		const double len3 = sqrt(z.x * z.x + z.y * z.y + z.z * z.z);
		if (k == 0)
		{
			iters = it;
			maxiterFirst = maxiter;
			r = a.r;
			DE = a.DE;
			r3 = len3;
		}
		else
		{
			const double dr = fabs(len3 - r3) / deltaDE;
			drSq = drSq + dr * dr; // dr1^2 + dr2^2 + dr3^2, summed left to right
		}
	}
	GenDistance o;
	o.iters = iters;
	o.itersLast = analytic ? iters : itersLast;
	o.maxiter = maxiterFirst ? 1 : 0;
	o.work = total;
	if (analytic)
	{
		if (!(DE > 0.0))
			o.d = r;
		else if (f.genHybrid)
		{
			if (f.genDEFunction == MBX_DE_FUNCTION_LINEAR)
				o.d = (r - f.genLinearDEOffset) / DE;
			else
				o.d = (r > 1.0) ? 0.5 * r * log(r) / DE : 0.0;
		}
		else
		{
			switch (f.genAnalytic)
			{
				case MBX_ANALYTIC_LOGARITHMIC: o.d = (r > 1.0) ? 0.5 * r * log(r) / DE : 0.0; break;
				case MBX_ANALYTIC_LINEAR: o.d = r / DE; break;
				case MBX_ANALYTIC_IFS: o.d = (r - 2.0) / DE; break;
				default: o.d = r; break;
			}
		}
		return o;
	}
	const double dr = sqrt(drSq);
	if (dr > 0.0)
		o.d = (f.genDEFunction == MBX_DE_FUNCTION_LOGARITHMIC) ? 0.5 * r3 * log(r3) / dr : 0.5 * r3 / dr;
	else
		o.d = r3;
	return o;
}

// Compute<calcModeColouring> + CalculateColorIndex of the sequence (compute_fractal.cpp:343-436,
// fractal_coloring.cpp:292-340) at the fractal-local point p
__device__ __attribute__((noinline)) double colour_index_generic(DV3 p, const FrameConst &frame)
{
#pragma clang fp contract(off)
	const FrameConst &f = uniform_ref(frame);
	DV4 z = dv4(p.x, p.y, p.z, f.genInitialW);
	const DV4 start = z;
	ExtendedAux a;
	init_aux(a, z, f.gen[f.genSeq[0]].mbScale);
	double colorMin = 1000.0;
	const int alg = f.colorAlgorithm;
	const bool pre = f.colorPreV215 != 0;
	const int maxN = 4 * f.formulaMaxiter;
	for (int i = 0; i < maxN; i++)
	{
		const DV4 lastZ = z;
		const FractalParams &s = gen_step(f, i, z, a);
		if (!finite4(z))
		{
			z = lastZ;
			a.r = len4(z);
			break;
		}
		DV4 cz = z;
		if (!f.color4D) cz.w = 0.0;
		double len;
		switch (alg)
		{
			case 0: len = len4(cz); break;
			case 1: len = fabs(dot4(cz, start)); break;
			case 2: len = fabs(len4(dv4(cz.x - start.x, cz.y - start.y, cz.z - start.z, cz.w - start.w)) - f.colorSphereRadiusD); break;
			case 3:
			{
				const double ax = fabs(cz.x), ay = fabs(cz.y), az = fabs(cz.z);
				len = (ax < ay) ? (ax < az ? ax : az) : (ay < az ? ay : az);
				if (f.color4D && fabs(cz.w) < len) len = fabs(cz.w);
				break;
			}
			default: len = fabs(cz.x * f.colorLineDirD[0] + cz.y * f.colorLineDirD[1] + cz.z * f.colorLineDirD[2]); break;
		}
		const DV4 dz = dv4(z.x - lastZ.x, z.y - lastZ.y, z.z - lastZ.z, z.w - lastZ.w);
		const double change = len4(dz) / a.r;
		if (!s.checkBailout) continue;
		const bool box = s.formula == MBX_FORMULA_MANDELBOX;
		if (!pre)
		{
			if (!box)
			{
				if (len < colorMin) colorMin = len;
				if (a.r > s.bailout) break;
			}
			else if (alg == 0)
			{
				if (a.r > 1e15 || change < 1e-15) break;
			}
			else
			{
				if (len < colorMin) colorMin = len;
				if (a.r > s.bailout || change < 1e-15) break;
			}
		}
		else if (!box || alg != 0)
		{
			if (len < colorMin) colorMin = len;
			if (a.r > 1e15 || change < 1e-15) break;
		}
		else if (a.r > 1e15 || change < 1e-15)
			break;
	}
	if (f.genHybrid)
	{
		colorMin = fmin(100.0, colorMin);
		const double r2 = fmin(a.r / fabs(a.DE), 20.0);
		return colorMin * 1000.0 + a.color * 100.0 + r2 * 5000.0;
	}
	if (f.genColoring == 1) return a.color * 100.0 + a.r * f.mbColorRD / 1e13 + ((alg != 0) ? colorMin * 1000.0 : 0.0);
	if (f.genColoring == 2) return colorMin * 1000.0;
	return colorMin * 5000.0;
}

} // namespace mbx

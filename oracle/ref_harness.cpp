// ref_harness.cpp - drives the reference's own formula and RNG translation units to produce
// golden vectors.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/formula_ref from:
//   mandelbulber2/formula/definition/fractal_mandelbulb.cpp
//   mandelbulber2/formula/definition/fractal_mandelbox.cpp
//   mandelbulber2/formula/definition/fractal_kaleidoscopic_ifs.cpp
//   mandelbulber2/formula/definition/fractal_{mandelbulb_power2,hypercomplex,quaternion,
//     menger_sponge,xenodreambuie,box_fold_bulb_pow2,aexion}.cpp
//   mandelbulber2/formula/definition/legacy_fractal_transforrms.cpp (BoxFolding, SphericalFolding)
//   mandelbulber2/src/common_math.cpp
// compiled where they lie under the reference checkout.  abstract_fractal.cpp and algebra.cpp
// need Qt and are not built: the link leaves their symbols (the cAbstractFractal base
// constructor/destructor/vtable slots and CRotationMatrix::RotateVector) unresolved, and this
// harness never executes code that reaches them (no formula object is constructed, virtual
// dispatch is bypassed with a qualified call, and the rotation branches of the Mandelbox and
// IFS formulas are disabled).
//
// Output: one record per line, every double printed as a C99 hex float so the fixture is exact.
#include <cstdio>
#include <cstring>
#include <cmath>
#include <cstdint>

#include "formula/definition/all_fractal_definitions.h"
#include "formula/definition/legacy_fractal_transforms.hpp"
#include "src/common_params.hpp"
#include "src/fractal.h"
#include "src/common_math.h"

extern unsigned int gRandomSeed;

namespace
{
// Splitmix64: deterministic inputs for the vectors.
uint64_t g_state = 0x2545F4914F6CDD1DULL;
uint64_t next_u64()
{
	uint64_t z = (g_state += 0x9E3779B97F4A7C15ULL);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
	return z ^ (z >> 31);
}
double uniform(double a, double b)
{
	return a + (b - a) * (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0);
}

alignas(sFractal) unsigned char g_fractal_storage[sizeof(sFractal)];

sFractal *fresh_fractal()
{
	std::memset(g_fractal_storage, 0, sizeof(g_fractal_storage));
	return reinterpret_cast<sFractal *>(g_fractal_storage);
}

template <class F>
void call_formula(CVector4 &z, const sFractal *f, sExtendedAux &aux)
{
	alignas(F) unsigned char obj[sizeof(F)];
	std::memset(obj, 0, sizeof(obj));
	reinterpret_cast<F *>(obj)->F::FormulaCode(z, f, aux);
}

void print_record(const char *tag, const double *params, int np, const CVector4 &zin,
	const sExtendedAux &ain, const CVector4 &zout, const sExtendedAux &aout)
{
	std::printf("%s", tag);
	for (int i = 0; i < np; i++) std::printf(" %a", params[i]);
	std::printf(" | %a %a %a %a | %a %a %a", zin.x, zin.y, zin.z, zin.w, ain.r, ain.DE, ain.color);
	std::printf(" | %a %a %a %a | %a %a %a\n", zout.x, zout.y, zout.z, zout.w, aout.r, aout.DE,
		aout.color);
}

void fresh_aux(sExtendedAux &aux, const CVector4 &z)
{
	std::memset(&aux, 0, sizeof(aux));
	aux.r = z.Length();
	aux.DE = uniform(1.0, 200.0);
	aux.color = uniform(0.0, 3.0);
}

// Mandelbulb params: power, alpha (deg), beta (deg)
void gen_mandelbulb()
{
	const double cases[][3] = {{8.0, 0.0, 0.0}, {2.0, 0.0, 0.0}, {3.5, 0.0, 0.0}, {9.0, 0.0, 0.0},
		{12.0, 0.0, 0.0}, {8.0, 20.0, -35.0}};
	for (const auto &c : cases)
	{
		sFractal *f = fresh_fractal();
		f->bulb.power = c[0];
		f->bulb.alphaAngleOffset = c[1] * M_PI_180;
		f->bulb.betaAngleOffset = c[2] * M_PI_180;
		for (int k = 0; k < 48; k++)
		{
			CVector4 z(uniform(-1.4, 1.4), uniform(-1.4, 1.4), uniform(-1.4, 1.4), 0.0);
			if (k == 0) z = CVector4(0.0, 0.0, 0.7, 0.0);   // on the axis: atan2(0,0)
			if (k == 1) z = CVector4(0.6, -0.3, 0.0, 0.0);  // in the z=0 plane
			if (k == 2) z = CVector4(-0.2, 0.0, -0.9, 0.0); // y = 0, x < 0
			sExtendedAux aux;
			fresh_aux(aux, z);
			CVector4 zin = z;
			sExtendedAux ain = aux;
			call_formula<cFractalMandelbulb>(z, f, aux);
			print_record("mandelbulb", c, 3, zin, ain, z, aux);
		}
	}
}

// Mandelbox params: scale, foldingLimit, foldingValue, minR, fixedR, offset xyz, color xyz, sp1,
// sp2
void gen_mandelbox()
{
	const double cases[][13] = {
		{2.0, 1.0, 2.0, 0.5, 1.0, 0.0, 0.0, 0.0, 0.03, 0.05, 0.07, 0.2, 0.1},
		{-1.5, 1.0, 2.0, 0.5, 1.0, 0.0, 0.0, 0.0, 0.03, 0.05, 0.07, 0.2, 0.1},
		{3.0, 0.8, 1.7, 0.4, 1.2, 0.1, -0.2, 0.05, 0.01, 0.02, 0.03, 0.5, 0.25}};
	for (const auto &c : cases)
	{
		sFractal *f = fresh_fractal();
		f->mandelbox.scale = c[0];
		f->mandelbox.foldingLimit = c[1];
		f->mandelbox.foldingValue = c[2];
		f->mandelbox.foldingSphericalMin = c[3];
		f->mandelbox.foldingSphericalFixed = c[4];
		f->mandelbox.offset = CVector4(c[5], c[6], c[7], 0.0);
		f->mandelbox.color.factor = CVector3(c[8], c[9], c[10]);
		f->mandelbox.color.factorSp1 = c[11];
		f->mandelbox.color.factorSp2 = c[12];
		// RecalculateFractalParams (fractal.cpp:673-675)
		f->mandelbox.fR2 = f->mandelbox.foldingSphericalFixed * f->mandelbox.foldingSphericalFixed;
		f->mandelbox.mR2 = f->mandelbox.foldingSphericalMin * f->mandelbox.foldingSphericalMin;
		f->mandelbox.mboxFactor1 = f->mandelbox.fR2 / f->mandelbox.mR2;
		f->mandelbox.rotationsEnabled = false;
		f->mandelbox.mainRotationEnabled = false;
		for (int k = 0; k < 64; k++)
		{
			double range = (k < 16) ? 0.6 : 3.0; // small |z| exercises both sphere folds
			CVector4 z(uniform(-range, range), uniform(-range, range), uniform(-range, range), 0.0);
			sExtendedAux aux;
			fresh_aux(aux, z);
			CVector4 zin = z;
			sExtendedAux ain = aux;
			call_formula<cFractalMandelbox>(z, f, aux);
			print_record("mandelbox", c, 13, zin, ain, z, aux);
		}
	}
}

// IFS params: scale, offset xyz, edge xyz, absX, absY, absZ, menger, edgeEnabled
void gen_kifs()
{
	const double cases[][12] = {{3.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1, 1, 1, 1, 0},
		{2.0, 1.0, 0.5, 0.25, 0.0, 0.0, 0.0, 1, 1, 1, 0, 0},
		{3.0, 1.0, 1.0, 1.0, 0.5, 0.7, 0.0, 1, 0, 1, 1, 1}};
	for (const auto &c : cases)
	{
		sFractal *f = fresh_fractal();
		f->IFS.scale = c[0];
		f->IFS.offset = CVector4(c[1], c[2], c[3], 0.0);
		f->IFS.edge = CVector3(c[4], c[5], c[6]);
		f->IFS.absX = c[7] != 0.0;
		f->IFS.absY = c[8] != 0.0;
		f->IFS.absZ = c[9] != 0.0;
		f->IFS.mengerSpongeMode = c[10] != 0.0;
		f->IFS.edgeEnabled = c[11] != 0.0;
		f->IFS.rotationEnabled = false;
		for (int i = 0; i < IFS_VECTOR_COUNT; i++) f->IFS.enabled[i] = false;
		for (int k = 0; k < 48; k++)
		{
			CVector4 z(uniform(-2.0, 2.0), uniform(-2.0, 2.0), uniform(-2.0, 2.0), 0.0);
			sExtendedAux aux;
			fresh_aux(aux, z);
			CVector4 zin = z;
			sExtendedAux ain = aux;
			call_formula<cFractalKaleidoscopicIfs>(z, f, aux);
			print_record("kifs", c, 12, zin, ain, z, aux);
		}
	}
}

// 4D orbit points for the formulas that use w
CVector4 point4(double r, bool withW)
{
	return CVector4(uniform(-r, r), uniform(-r, r), uniform(-r, r), withW ? uniform(-r, r) : 0.0);
}

// The parameter-free power-2 formulas and the Menger sponge: params are the formula's own
// (Menger: transf_scale_3)
template <class F>
void gen_simple(const char *tag, const double *params, int np, bool withW, double range)
{
	sFractal *f = fresh_fractal();
	if (np > 0) f->transformCommon.scale3 = params[0];
	for (int k = 0; k < 48; k++)
	{
		CVector4 z = point4(range, withW);
		if (k == 0) z = CVector4(0.0, 0.0, 0.7, withW ? 0.3 : 0.0); // on the z axis
		sExtendedAux aux;
		fresh_aux(aux, z);
		CVector4 zin = z;
		sExtendedAux ain = aux;
		call_formula<F>(z, f, aux);
		print_record(tag, params, np, zin, ain, z, aux);
	}
}

// Xenodreambuie params: power, alpha (deg), beta (deg)
void gen_xenodreambuie()
{
	const double cases[][3] = {{8.0, 0.0, 0.0}, {3.0, 0.0, 0.0}, {5.5, 25.0, -40.0}};
	for (const auto &c : cases)
	{
		sFractal *f = fresh_fractal();
		f->bulb.power = c[0];
		f->bulb.alphaAngleOffset = c[1] * M_PI_180;
		f->bulb.betaAngleOffset = c[2] * M_PI_180;
		for (int k = 0; k < 48; k++)
		{
			CVector4 z = point4(1.4, false);
			if (k == 0) z = CVector4(0.0, 0.0, -0.7, 0.0);
			sExtendedAux aux;
			fresh_aux(aux, z);
			CVector4 zin = z;
			sExtendedAux ain = aux;
			call_formula<cFractalXenodreambuie>(z, f, aux);
			print_record("xenodreambuie", c, 3, zin, ain, z, aux);
		}
	}
}

// Box fold bulb pow 2 params: foldFactor, zFactor, analyticDE enabledFalse, scale1, offset1, offset2
void gen_box_fold_bulb_pow2()
{
	const double cases[][6] = {{2.0, 5.0, 0, 1.0, 1.0, 1.0}, {1.0, 3.0, 1, 1.2, 0.5, 0.7}};
	for (const auto &c : cases)
	{
		sFractal *f = fresh_fractal();
		f->foldingIntPow.foldFactor = c[0];
		f->foldingIntPow.zFactor = c[1];
		f->analyticDE.enabledFalse = c[2] != 0.0;
		f->analyticDE.scale1 = c[3];
		f->analyticDE.offset1 = c[4];
		f->analyticDE.offset2 = c[5];
		for (int k = 0; k < 48; k++)
		{
			CVector4 z = point4(3.0, true);
			sExtendedAux aux;
			fresh_aux(aux, z);
			CVector4 zin = z;
			sExtendedAux ain = aux;
			call_formula<cFractalBoxFoldBulbPow2>(z, f, aux);
			print_record("boxfoldbulbpow2", c, 6, zin, ain, z, aux);
		}
	}
}

// Aexion params: cadd, iteration index, analyticDE enabled, scale1, offset1, then the input
// const_c (xyzw); the record ends with the output const_c
void gen_aexion()
{
	const double cases[][5] = {{-1.3, 0, 1, 1.0, 1.0}, {-1.3, 3, 1, 1.0, 1.0}, {0.4, 0, 0, 1.0, 1.0},
		{-0.7, 5, 1, 0.8, 0.3}};
	for (const auto &c : cases)
	{
		sFractal *f = fresh_fractal();
		f->aexion.cadd = c[0];
		f->analyticDE.enabled = c[2] != 0.0;
		f->analyticDE.scale1 = c[3];
		f->analyticDE.offset1 = c[4];
		for (int k = 0; k < 48; k++)
		{
			CVector4 z = point4(1.5, true);
			sExtendedAux aux;
			fresh_aux(aux, z);
			aux.i = (int)c[1];
			aux.const_c = point4(1.5, true);
			double params[9] = {c[0], c[1], c[2], c[3], c[4], aux.const_c.x, aux.const_c.y, aux.const_c.z,
				aux.const_c.w};
			CVector4 zin = z;
			sExtendedAux ain = aux;
			call_formula<cFractalAexion>(z, f, aux);
			print_record("aexion", params, 9, zin, ain, z, aux);
			std::printf("aexion_c %a %a %a %a\n", aux.const_c.x, aux.const_c.y, aux.const_c.z, aux.const_c.w);
		}
	}
}

// Global foldings (compute_fractal.cpp:192-202).  Params: boxLimit, boxValue, sphericalOuter,
// sphericalInner.  Each record is one call on a fresh point (aux.r = |z| as Compute leaves it).
void gen_foldings()
{
	const double cases[][4] = {{1.0, 2.0, 1.0, 0.5}, {0.7, 1.5, 1.3, 0.4}, {1.2, 2.6, 0.9, 0.6}};
	for (const auto &c : cases)
	{
		sFractalFoldings fold;
		std::memset(&fold, 0, sizeof(fold));
		fold.boxLimit = c[0];
		fold.boxValue = c[1];
		fold.sphericalOuter = c[2];
		fold.sphericalInner = c[3];
		for (int k = 0; k < 48; k++)
		{
			const double range = (k < 24) ? 2.0 : 1.2;
			CVector4 z(uniform(-range, range), uniform(-range, range), uniform(-range, range),
				(k % 3 == 0) ? uniform(-0.5, 0.5) : 0.0);
			sExtendedAux aux;
			fresh_aux(aux, z);
			CVector4 zin = z;
			sExtendedAux ain = aux;
			BoxFolding(z, &fold, aux);
			print_record("boxfolding", c, 4, zin, ain, z, aux);
			z = zin;
			aux = ain;
			SphericalFolding(z, &fold, aux);
			print_record("sphericalfolding", c, 4, zin, ain, z, aux);
		}
	}
}

void gen_random()
{
	gRandomSeed = 1;
	std::printf("randomint");
	for (int i = 0; i < 64; i++) std::printf(" %d", RandomInt());
	std::printf("\n");
	std::printf("random1000");
	for (int i = 0; i < 64; i++) std::printf(" %d", Random(1000));
	std::printf("\n");
	std::printf("random20000");
	for (int i = 0; i < 64; i++) std::printf(" %d", Random(20000));
	std::printf("\n");
	std::printf("final_seed %u\n", gRandomSeed);
}
} // namespace

int main()
{
	gen_mandelbulb();
	gen_mandelbox();
	gen_kifs();
	const double none[1] = {0.0};
	gen_simple<cFractalMandelbulbPower2>("mandelbulbpower2", none, 0, false, 1.4);
	gen_simple<cFractalHypercomplex>("hypercomplex", none, 0, true, 1.4);
	gen_simple<cFractalQuaternion>("quaternion", none, 0, true, 1.4);
	const double scales[2][1] = {{3.0}, {2.6}};
	for (const auto &sc : scales)
		gen_simple<cFractalMengerSponge>("mengersponge", sc, 1, true, 1.6);
	gen_xenodreambuie();
	gen_box_fold_bulb_pow2();
	gen_aexion();
	gen_foldings();
	gen_random();
	return 0;
}

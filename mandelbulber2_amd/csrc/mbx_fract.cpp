// mbx_fract.cpp - reference parameter defaults and the .fract scene reader (include/mbx.h).
//
// The reader follows what cSettings::LoadFromString + Decode do with a settings file
// (mandelbulber2/src/settings.cpp) for the parameters the raymarch path consumes:
//   header           DecodeHeader            settings.cpp:540-601
//   sections, lines  Decode / CheckSection   settings.cpp:647-905, 1247-1256
//   one line         DecodeOneLine           settings.cpp:1015-1245
//   renames          Compatibility           settings.cpp:1262-1492
//   migrations       Compatibility2          settings.cpp:1494-1754
//   legacy fractals  TryResolveLegacyFractalParam, MigrateToObjectsTree
//                                            settings.cpp:2268-2322, 2551-2760
//   objects tree     cObjectsTree::CreateNodeDataFromParameters  objects_tree.cpp:36-96
// Values are typed as cMultiVal stores them (multi_val.cpp:87-123): booleans are "true" or not,
// decimals take ',' or '.', colours are 16-bit hexadecimal triples, and numeric parameters are
// clamped to their registered limits (one_parameter.cpp:112-150).  The file is parsed into
// per-section key/value maps (the last line for a key wins, as repeated Set calls do); the typed
// getters apply each parameter's default, type and limits.  Unknown parameter names are not
// errors here (the reference rejects names it has not registered; this reader knows only the
// names on the path), so only malformed lines count toward the three-error limit.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "../../include/mbx.h"

namespace
{

// ------------------------------------------------------------------------------------------
// string helpers (QString semantics where they matter)
// ------------------------------------------------------------------------------------------
bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

std::string trim(const std::string &s)
{
	size_t a = 0, b = s.size();
	while (a < b && is_space(s[a])) a++;
	while (b > a && is_space(s[b - 1])) b--;
	return s.substr(a, b - a);
}

bool starts_with(const std::string &s, const char *p) { return s.compare(0, std::strlen(p), p) == 0; }
bool ends_with(const std::string &s, const char *p)
{
	const size_t n = std::strlen(p);
	return s.size() >= n && s.compare(s.size() - n, n, p) == 0;
}
bool contains(const std::string &s, const char *p) { return s.find(p) != std::string::npos; }

// QString::split(c) with empty parts kept
std::vector<std::string> split(const std::string &s, char c)
{
	std::vector<std::string> out;
	size_t a = 0;
	for (;;)
	{
		const size_t b = s.find(c, a);
		if (b == std::string::npos)
		{
			out.push_back(s.substr(a));
			return out;
		}
		out.push_back(s.substr(a, b - a));
		a = b + 1;
	}
}

void replace_all(std::string &s, const char *from, const char *to)
{
	const size_t n = std::strlen(from), m = std::strlen(to);
	if (n == 0) return;
	for (size_t pos = s.find(from); pos != std::string::npos; pos = s.find(from, pos + m))
		s.replace(pos, n, to);
}

// QString::toInt(&ok) in base 10 or 16: the whole (trimmed) string must be a number
bool parse_int(const std::string &t, int base, long long &out)
{
	const std::string s = trim(t);
	if (s.empty()) return false;
	char *end = nullptr;
	const long long v = std::strtoll(s.c_str(), &end, base);
	if (end == s.c_str() || *end != '\0') return false;
	if (v < -2147483648LL || v > 2147483647LL) return false;
	out = v;
	return true;
}
long long to_int(const std::string &s, int base = 10)
{
	long long v = 0;
	return parse_int(s, base, v) ? v : 0;
}

// QLocale::toDouble after everyLocaleDouble (settings.cpp:2063-2072): ',' and '.' both decimal
// points; the whole string must be a number, otherwise 0
double to_double(std::string s)
{
	s = trim(s);
	for (char &c : s)
		if (c == ',') c = '.';
	if (s.empty()) return 0.0;
	char *end = nullptr;
	const double v = std::strtod(s.c_str(), &end);
	if (end == s.c_str() || *end != '\0') return 0.0;
	return v;
}

// ------------------------------------------------------------------------------------------
// the decoded file
// ------------------------------------------------------------------------------------------
using KV = std::map<std::string, std::string>;

struct Doc
{
	double version = 0.0;
	int format = 0;
	KV main;                    // [main_parameters], after renames
	std::map<int, KV> fractal;  // [fractal_N] and legacy name_N lines, keyed by N (1-based)
	int lines = 0, errors = 0;
	bool animation = false;
	std::string firstError;
	std::set<int> materialsSeen; // matN_* names decoded (lazy InitMaterialParams)
	std::set<int> lightsSeen;    // lightN_* names decoded (lazy InitLightParams)
	bool primitives = false;
	std::set<int> nodeIds;       // node_XXXX_definition names
};

// typed views of one container with defaults
struct View
{
	const KV *kv;
	bool has(const std::string &k) const { return kv && kv->count(k); }
	const std::string *raw(const std::string &k) const
	{
		if (!kv) return nullptr;
		auto it = kv->find(k);
		return it == kv->end() ? nullptr : &it->second;
	}
	bool b(const std::string &k, bool def) const
	{
		const std::string *v = raw(k);
		return v ? *v == "true" : def; // DecodeOneLine: value == "true" ? 1 : 0
	}
	long long i(const std::string &k, long long def, long long lo = -2147483648LL, long long hi = 2147483647LL) const
	{
		const std::string *v = raw(k);
		if (!v) return def;
		return std::min(std::max(to_int(*v), lo), hi);
	}
	double d(const std::string &k, double def, double lo = -INFINITY, double hi = INFINITY) const
	{
		const std::string *v = raw(k);
		if (!v) return def;
		return std::min(std::max(to_double(*v), lo), hi);
	}
	// vectors: components past the last one given keep their default (cMultiVal::Store(QString)
	// assigns only the parts present)
	void v3(const std::string &k, double out[3], double d0, double d1, double d2) const
	{
		out[0] = d0;
		out[1] = d1;
		out[2] = d2;
		const std::string *v = raw(k);
		if (!v) return;
		std::vector<std::string> parts = split(*v, ' ');
		for (size_t j = 0; j < parts.size() && j < 3; j++) out[j] = to_double(parts[j]);
	}
	// colours: 16-bit hexadecimal components, clamped to [0, 65535], returned as c / 65536
	void rgb(const std::string &k, double out[3], int r, int g, int bl) const
	{
		long long c[3] = {r, g, bl};
		const std::string *v = raw(k);
		if (v)
		{
			std::vector<std::string> parts = split(*v, ' ');
			for (size_t j = 0; j < parts.size() && j < 3; j++)
				c[j] = std::min(std::max(to_int(parts[j], 16), 0LL), 65535LL);
		}
		for (int j = 0; j < 3; j++) out[j] = (double)(float)(c[j] / 65536.0);
	}
	// enumeration parameters: a name from the list or its index (cOneParameter::GetIndexByEnumeration)
	long long e(const std::string &k, long long def, const std::vector<const char *> &names) const
	{
		const std::string *v = raw(k);
		if (!v) return def;
		for (size_t j = 0; j < names.size(); j++)
			if (*v == names[j]) return (long long)j;
		return to_int(*v);
	}
};

// Compatibility (settings.cpp:1262-1492): renames and value rewrites by file version.  Only the
// branches that reach parameters this reader uses or reports are restated.
void compatibility(double ver, std::string &name, std::string &value)
{
	if (ver <= 2.01)
	{
		if (name == "volumetric_light_enabled_0") name = "main_light_volumetric_enabled";
		else if (name == "volumetric_light_intensity_0") name = "main_light_volumetric_intensity";
		else if (contains(name, "volumetric_light")) replace_all(name, "volumetric_light", "aux_light_volumetric");
		if (contains(name, "aux_light_predefined")) replace_all(name, "aux_light_predefined", "aux_light");
	}
	if (ver <= 2.04 && name == "fractal_constant_factor") value = value + " " + value + " " + value;
	if (ver <= 2.06 && name == "linear_DE_mode")
	{
		name = "delta_DE_function";
		value = "1";
	}
	if (ver <= 2.071)
	{
		static const char *const materialNames[][2] = {{"shading", "mat1_shading"},
			{"specular", "mat1_specular"}, {"reflect", "mat1_reflectance"},
			{"transparency_of_surface", "mat1_transparency_of_surface"},
			{"transparency_of_interior", "mat1_transparency_of_interior"},
			{"transparency_index_of_refraction", "mat1_transparency_index_of_refraction"},
			{"transparency_interior_color", "mat1_transparency_interior_color"},
			{"fresnel_reflectance", "mat1_fresnel_reflectance"}, {"coloring_random_seed", "mat1_coloring_random_seed"},
			{"coloring_saturation", "mat1_coloring_saturation"}, {"coloring_speed", "mat1_coloring_speed"},
			{"coloring_palette_size", "mat1_coloring_palette_size"},
			{"coloring_palette_offset", "mat1_coloring_palette_offset"},
			{"fractal_color", "mat1_use_colors_from_palette"},
			{"surface_color_palette", "mat1_surface_color_palette"},
			{"fractal_coloring_algorithm", "mat1_fractal_coloring_algorithm"},
			{"fractal_coloring_sphere_radius", "mat1_fractal_coloring_sphere_radius"},
			{"fractal_coloring_line_direction", "mat1_fractal_coloring_line_direction"}};
		for (const auto &m : materialNames)
			if (name == m[0])
			{
				name = m[1];
				break;
			}
	}
	if (ver < 2.09 && name == "delta_DE_function")
	{
		if (value == "0") value = "2";
		else if (value == "2") value = "0";
	}
	if (ver < 2.12) replace_all(name, "gpu_", "opencl_");
	if (ver < 2.19)
	{
		if (contains(name, "surface_color_palette"))
		{
			// a list of colours becomes "position colour" pairs spread evenly over 0..10000
			replace_all(name, "surface_color_palette", "surface_color_gradient");
			std::vector<std::string> parts = split(value, ' ');
			size_t numberOfColors = parts.size();
			if (!parts.empty() && parts.back().size() < 6) numberOfColors--;
			const double step = 1.0 / (double)numberOfColors;
			std::string v;
			for (size_t i = 0; i < numberOfColors; i++)
			{
				if (i > 0) v += " ";
				v += std::to_string((int)(i * step * 10000.0)) + " " + parts[i];
			}
			value = v;
		}
		if (contains(name, "luminosity_color_thesame"))
			replace_all(name, "luminosity_color_thesame", "luminosity_gradient_enable");
		if (contains(name, "coloring_palette_size") || contains(name, "coloring_random_seed") ||
			contains(name, "coloring_saturation"))
			name = "skip";
	}
	if (ver < 2.25)
	{
		static const char *const lightNames[][2] = {{"main_light_intensity", "light1_intensity"},
			{"main_light_visibility_size", "light1_size"},
			{"main_light_contour_sharpness", "light1_contour_sharpness"}, {"main_light_alpha", "light1_alpha"},
			{"main_light_beta", "light1_beta"}, {"main_light_colour", "light1_color"},
			{"penetrating_lights", "light1_penetrating"}, {"shadows_enabled", "light1_cast_shadows"},
			{"shadows_cone_angle", "light1_soft_shadow_cone"}, {"main_light_enable", "light1_enabled"},
			{"main_light_position_relative", "light1_relative_position"},
			{"main_light_volumetric_intensity", "light1_volumetric_visibility"},
			{"main_light_volumetric_enabled", "light1_volumetric"}};
		if (name == "main_light_visibility") name = "light1_visibility";
		for (const auto &m : lightNames)
			if (contains(name, m[0])) replace_all(name, m[0], m[1]);
		if (contains(name, "aux_light"))
		{
			if (contains(name, "visibility_size")) replace_all(name, "aux_light_visibility_size", "light2_size");
			else if (contains(name, "visibility")) replace_all(name, "aux_light_visibility", "light2_visibility");
			else if (!contains(name, "aux_light_place_behind"))
			{
				// aux_light_<what>_<i> -> light<i+1>_<what>
				std::vector<std::string> parts = split(name, '_');
				if (parts.size() >= 4)
				{
					const std::string prefix = "light" + std::to_string(to_int(parts.back()) + 1);
					if (parts[2] == "intensity") name = prefix + "_intensity";
					else if (parts[2] == "position") name = prefix + "_position";
					else if (parts[2] == "enabled") name = prefix + "_enabled";
					else if (parts[2] == "colour") name = prefix + "_color";
					else if (parts[2] == "volumetric" && parts.size() >= 5)
					{
						if (parts[3] == "intensity") name = prefix + "_volumetric_visibility";
						else if (parts[3] == "enabled") name = prefix + "_volumetric";
					}
				}
			}
		}
	}
}

// Parameters of the per-fractal containers that this reader reads (InitFractalParams,
// initparameters.cpp:675-768); a legacy main-section line "name_N" with one of these names goes
// to fractal N (TryResolveLegacyFractalParam, settings.cpp:2268-2322).
bool is_fractal_param(const std::string &base)
{
	static const std::set<std::string> names = {"formula", "formula_iterations", "formula_weight",
		"formula_start_iteration", "formula_stop_iteration", "fractal_enable", "dont_add_c_constant",
		"check_for_bailout", "power", "alpha_angle_offset", "beta_angle_offset", "gamma_angle_offset",
		"IFS_scale", "IFS_rotation", "IFS_rotation_enabled", "IFS_offset", "IFS_edge", "IFS_edge_enabled",
		"IFS_abs_x", "IFS_abs_y", "IFS_abs_z", "IFS_menger_sponge_mode", "mandelbox_scale",
		"mandelbox_folding_limit", "mandelbox_folding_value", "mandelbox_folding_min_radius",
		"mandelbox_folding_fixed_radius", "mandelbox_sharpness", "mandelbox_offset", "mandelbox_rotation_main",
		"mandelbox_color", "mandelbox_color_Sp1", "mandelbox_color_Sp2", "mandelbox_color_R", "mandelbox_main_rotation_enabled",
		"mandelbox_rotations_enabled", "formula_position", "formula_rotation", "formula_repeat", "formula_scale",
		"formula_material_id", "transf_scale_3", "cadd", "boxfold_bulbpow2_folding_factor",
		"boxfold_bulbpow2_z_factor", "analyticDE_enabled", "analyticDE_enabled_false", "analyticDE_scale_1",
		"analyticDE_offset_1", "analyticDE_offset_2", "fold_color_aux_color_enabled_false"};
	return names.count(base) != 0;
}

// main-section names of the form base_N that are main parameters, not legacy fractal lines:
// the indexed temporaries InjectTemporaryLegacyFractalParams and InjectTemporaryLegacyBooleanParams
// add (settings.cpp:2128-2172; dont_add_c_constant_N for N = 1..8)
bool is_main_indexed(const std::string &base, long long n)
{
	return base == "julia_mode" || base == "julia_c" || base == "fractal_constant_factor" ||
		base == "formula_maxiter" || base == "initial_waxis" || (base == "dont_add_c_constant" && n <= 8);
}

bool decode_one_line(Doc &doc, const std::string &lineIn, KV *container, bool mainSection, int fractalIndex)
{
	const std::string line = trim(lineIn);
	if (line.empty()) return true;
	const size_t firstSpace = line.find(' ');
	const size_t semicolon = line.find(';');
	if (firstSpace == std::string::npos || semicolon == std::string::npos || semicolon <= firstSpace)
		return false;
	std::string name = trim(line.substr(0, firstSpace));
	std::string value = trim(line.substr(firstSpace + 1, semicolon - firstSpace - 1));
	auto strip = [](const std::string &t) -> std::string {
		if (t.size() >= 2 && ((t.front() == '"' && t.back() == '"') || (t.front() == '\'' && t.back() == '\'')))
			return trim(t.substr(1, t.size() - 2));
		return t;
	};
	name = strip(name);
	value = strip(value);
	doc.lines++;

	if (mainSection && doc.version < 2.35)
	{
		// legacy per-fractal line "name_N" in the main section
		const size_t us = name.rfind('_');
		long long idx = 0;
		if (us != std::string::npos && us > 0 && parse_int(name.substr(us + 1), 10, idx) && idx >= 1)
		{
			const std::string base = name.substr(0, us);
			if (!is_main_indexed(base, idx) && is_fractal_param(base))
			{
				std::string n = base, v = value;
				compatibility(doc.version, n, v);
				if (n != "skip") doc.fractal[(int)idx][n] = v;
				return true;
			}
		}
	}

	compatibility(doc.version, name, value);
	if (name == "skip") return true;

	if (mainSection)
	{
		if (starts_with(name, "primitive_")) doc.primitives = true;
		if (starts_with(name, "mat"))
		{
			const size_t us = name.find('_');
			long long idx = 0;
			if (us != std::string::npos && parse_int(name.substr(3, us - 3), 10, idx))
			{
				if (!doc.materialsSeen.count((int)idx) && doc.version < 2.15)
				{
					// PreCompatibilityMaterials (settings.cpp:2107-2119): pre-2.15 material defaults,
					// set when the material is first seen; later lines override them
					const std::string p = "mat" + std::to_string(idx) + "_";
					container->emplace(p + "metallic", "false");
					container->emplace(p + "specular", "1");
					container->emplace(p + "specular_width", "1");
				}
				doc.materialsSeen.insert((int)idx);
			}
		}
		if (starts_with(name, "light"))
		{
			const size_t us = name.find('_');
			long long idx = 0;
			if (us != std::string::npos && parse_int(name.substr(5, us - 5), 10, idx)) doc.lightsSeen.insert((int)idx);
		}
		if (starts_with(name, "node_") && ends_with(name, "_definition"))
		{
			std::vector<std::string> parts = split(name, '_');
			long long id = 0;
			if (parts.size() >= 3 && parse_int(parts[1], 10, id) && id > 0) doc.nodeIds.insert((int)id);
		}
	}
	(void)fractalIndex;
	(*container)[name] = value;
	return true;
}

int fail(mbx_fract_info *info, const std::string &msg)
{
	if (info) std::snprintf(info->message, sizeof(info->message), "%s", msg.c_str());
	return MBX_ERR_INVALID;
}

// cColorGradient::SetColorsFromString + SortGradient (color_gradient.cpp:271-282, 305-354):
// "position colour" pairs, positions in 1/10000, colours 24-bit hexadecimal; the first colour is
// also placed at 1.0.  Fewer than two parts give white at 0 and 1.  Sorting keeps the order of
// equal positions (the reference's std::sort does too for gradients of up to 16 stops, where it
// runs an insertion sort).  Stops past MBX_MAX_GRADIENT_STOPS are dropped.
int parse_gradient(const std::string &text, double pos[], double col[][3])
{
	struct Stop
	{
		float position;
		int rgb[3];
	};
	std::vector<Stop> stops;
	std::vector<std::string> parts = split(text, ' ');
	if (parts.size() < 2)
	{
		stops.push_back({0.0f, {255, 255, 255}});
		stops.push_back({1.0f, {255, 255, 255}});
	}
	else
	{
		float position = 0.0f;
		for (size_t i = 0; i < parts.size(); i++)
		{
			if (parts[i].empty()) continue;
			if (i % 2 == 0)
				position = (float)to_int(parts[i]) / 10000.0f;
			else
			{
				const long long hex = to_int(parts[i], 16);
				stops.push_back({position, {(int)(hex / 65536), (int)((hex / 256) % 256), (int)(hex % 256)}});
				if (i == 1) stops.push_back({1.0f, {stops[0].rgb[0], stops[0].rgb[1], stops[0].rgb[2]}});
			}
		}
	}
	std::stable_sort(stops.begin(), stops.end(), [](const Stop &a, const Stop &b) { return a.position < b.position; });
	const int n = (int)std::min(stops.size(), (size_t)MBX_MAX_GRADIENT_STOPS);
	for (int i = 0; i < MBX_MAX_GRADIENT_STOPS; i++)
	{
		pos[i] = i < n ? stops[i].position : 0.0;
		for (int c = 0; c < 3; c++) col[i][c] = i < n ? stops[i].rgb[c] : 0.0;
	}
	return n;
}

// material surface gradient default (initparameters.cpp, InitMaterialParams)
const char *const kDefaultSurfaceGradient =
	"0 fd6029 1000 698403 2000 fff59b 3000 f5bd22 4000 0b5e87 5000 c68876 6000 a51c64 7000 3b9fee "
	"8000 d4ffd4 9000 aba53c";

// ------------------------------------------------------------------------------------------
// defaults (initparameters.cpp line numbers)
// ------------------------------------------------------------------------------------------
void set3(double *o, double a, double b, double c)
{
	o[0] = a;
	o[1] = b;
	o[2] = c;
}
void rgb16(double *o, int r, int g, int b)
{
	o[0] = (double)(float)(r / 65536.0);
	o[1] = (double)(float)(g / 65536.0);
	o[2] = (double)(float)(b / 65536.0);
}

} // namespace

// One fractal container's defaults (InitFractalParams, initparameters.cpp:695-1010) with the
// sequence-wide values of InitParams (:128-156) and the node defaults (:2102-2125)
static void default_fractal(mbx_fractal &f, int formula)
{
	f.formula = formula;                   // :695 (container fractal0 Mandelbulb, the others none)
	f.formulaMaxiter = 250;                // node formula_maxiter :2123, N :130
	f.deltaDEMethod = 0;                   // :154
	f.deltaDEFunction = 0;                 // :152
	f.useDefaultBailout = 1;               // :155
	f.bailout = 100.0;                     // :150
	f.juliaMode = 0;                       // :128
	set3(f.juliaC, 0.0, 0.0, 0.0);         // :129
	set3(f.constantFactor, 1.0, 1.0, 1.0); // :132
	f.initialWAxis = 0.0;                  // :156
	f.formulaIterations = 1;               // :701
	f.formulaWeight = 1.0;                 // :702
	f.formulaStartIteration = 0;           // :703
	f.formulaStopIteration = 250;          // :704
	f.dontAddCConstant = 0;                // :713
	f.checkForBailout = 1;                 // :714
	f.bulb.power = 9.0;                    // :717
	mbx_fractal_mandelbox &mb = f.mandelbox;
	mb.scale = 2.0;                        // :746
	mb.foldingLimit = 1.0;
	mb.foldingValue = 2.0;
	mb.foldingSphericalMin = 0.5;
	mb.foldingSphericalFixed = 1.0;        // :750
	set3(mb.colorFactor, 0.03, 0.05, 0.07); // :761
	mb.colorFactorSp1 = 0.2;               // :765
	mb.colorFactorSp2 = 0.2;
	mb.colorFactorR = 0.0;                 // :764
	mbx_fractal_ifs &ifs = f.IFS;
	ifs.scale = 2.0;                       // :725
	set3(ifs.offset, 1.0, 0.0, 0.0);       // :728
	for (int i = 0; i < MBX_IFS_VECTOR_COUNT; i++)
	{
		set3(ifs.direction[i], 1.0, 0.0, 0.0); // :738-742
		ifs.intensity[i] = 1.0;
	}
	mbx_fractal_transform &t = f.transform;
	t.aexionCadd = -1.3;                   // :722
	t.foldingIntPowFoldFactor = 2.0;       // :784
	t.foldingIntPowZFactor = 5.0;          // :785
	t.analyticDEEnabled = 1;               // :904
	t.analyticDEEnabledFalse = 0;          // :905
	t.analyticDEScale1 = 1.0;              // :906
	t.analyticDEOffset1 = 1.0;             // :909
	t.analyticDEOffset2 = 1.0;             // :910
	t.scale3 = 3.0;                        // :1010
}

extern "C" void mbx_params_default(mbx_params *p)
{
	if (!p) return;
	std::memset(p, 0, sizeof(*p));
	p->imageWidth = 800;                   // :64
	p->imageHeight = 600;                  // :65
	set3(p->camera, 3.0, -6.0, 2.0);       // :113
	set3(p->target, 0.0, 0.0, 0.0);        // :114
	set3(p->topVector, -0.1277753, 0.2555506, 0.958314); // :116
	p->fov = 53.13;                        // :120
	p->perspectiveType = 0;                // :121 three_point
	p->sweetSpotHAngle = 0.0;              // :124
	p->sweetSpotVAngle = 0.0;              // :125
	p->minN = 1;                           // :131
	p->detailLevel = 1.0;                  // :133
	p->DEThresh = 0.01;                    // :135
	p->smoothness = 1.0;                   // :136
	p->iterThreshMode = 0;                 // :137
	p->DEFactor = 1.0;                     // :139
	p->viewDistanceMax = 50.0;             // :141
	p->interiorMode = 0;                   // :147
	p->constantDEThreshold = 0;            // :148
	p->maxRaymarchingSteps = 10000;        // :167
	// ambient occlusion :213-219
	p->ambientOcclusion = 1.0;
	p->ambientOcclusionQuality = 4;
	p->ambientOcclusionFastTune = 1.0;
	p->ambientOcclusionEnabled = 0;
	p->ambientOcclusionMode = MBX_AO_SCREEN_SPACE;
	rgb16(p->ambientOcclusionColor, 65535, 65535, 65535);
	p->aoVectorsCount = 0;
	// background :227-245
	p->backgroundBrightness = 1.0;
	p->backgroundGamma = 1.0;
	rgb16(p->backgroundColor1, 0, 38306, 65535);
	rgb16(p->backgroundColor2, 65535, 65535, 65535);
	rgb16(p->backgroundColor3, 0, 10000, 500);
	p->background3ColorsEnable = 1;
	p->reflectionsMax = 5;                 // :237
	rgb16(p->fillLightColor, 0, 0, 0);     // :309
	// Monte Carlo :329-340
	p->DOFMonteCarlo = 0;
	p->DOFSamples = 100;
	p->DOFMonteCarloGlobalIllumination = 0;
	p->monteCarloGIRadianceLimit = 10.0;
	p->rngMode = MBX_RNG_PER_PIXEL;
	p->randomSeed = 0;
	// material #1 (InitMaterialParams, :1600-1810)
	rgb16(p->surfaceColor, 50000, 50000, 50000);
	p->shading = 1.0;
	p->luminosity = 0.0;
	rgb16(p->luminosityColor, 65535, 65535, 65535);
	// light #1 (InitLightParams, :1841-1960)
	mbx_light &L = p->light;
	L.enabled = 1;
	L.castShadows = 1;
	L.penetrating = 1;
	L.relativePosition = 1;
	set3(L.rotation, -45.0, 45.0, 0.0);
	L.intensity = 1.0;
	L.visibility = 1.0;
	L.size = 0.5;
	L.softShadowCone = 1.0;
	L.contourSharpness = 1.0;
	rgb16(L.color, 65535, 65535, 65535);
	// fractal containers; a single formula (hybrid_fractal_enable :149, repeat_from :151,
	// linear_DE_offset :157)
	default_fractal(p->fractal, MBX_FORMULA_MANDELBULB);
	for (int k = 0; k < MBX_MAX_FRACTALS - 1; k++) default_fractal(p->hybridFractals[k], MBX_FORMULA_NONE);
	p->hybridCount = 0;
	p->repeatFrom = 1;
	p->linearDEOffset = 0.0;
	// advanced quality march limits :159-166
	p->advancedQuality = 0;
	p->absMaxMarchingStep = 3.0;
	p->absMinMarchingStep = 1e-15;
	p->relMaxMarchingStep = 1e4;
	p->relMinMarchingStep = 1e-2;
	p->detailSizeMax = 1.0;
	p->detailSizeMin = 1e-12;
	p->deltaDERelativeDelta = 0.01;
	p->image.brightness = 1.0f;
	p->image.contrast = 1.0f;
	p->image.gamma = 1.0f;
	p->image.saturation = 1.0f;
	p->image.hdr = 0;
	// global foldings (initparameters.cpp:199-204) and slow shading (:140)
	p->boxFolding = 0;
	p->boxFoldingLimit = 1.0;
	p->boxFoldingValue = 2.0;
	p->sphericalFolding = 0;
	p->sphericalFoldingOuter = 1.0;
	p->sphericalFoldingInner = 0.5;
	p->slowShading = 0;
	p->N = 250;
	p->DOFMinSamples = 10;
	p->DOFMaxNoise = 1.0;
	p->aoLightMapRotation = 0.0;
	p->DOFFocus = 6.0;
	p->DOFRadius = 10.0;
	p->DOFMonteCarloCACameraDispersion = 1.0;
	// palette colouring of material #1 (InitMaterialParams)
	p->useColorsFromPalette = 1;
	p->surfaceGradientEnable = 1;
	p->coloringAlgorithm = 0;
	p->coloringPreV215 = 0;
	p->coloring4D = 0;
	p->coloringSpeed = 1.0;
	p->coloringPaletteOffset = 0.0;
	p->coloringSphereRadius = 1.0;
	set3(p->coloringLineDirection, 1.0, 0.0, 0.0);
	p->surfaceGradientCount = parse_gradient(kDefaultSurfaceGradient, p->surfaceGradientPosition, p->surfaceGradientColor);
	// specular of material #1 (InitMaterialParams)
	p->specularPlasticEnable = 1;
	p->metallic = 1;
	p->specular = 5.0;
	p->specularWidth = 0.05;
	p->specularMetallic = 2.0;
	p->specularMetallicWidth = 1.0;
	p->specularMetallicRoughness = 0.1;
	rgb16(p->specularColor, 65535, 65535, 65535);
	p->luminosityEmissive = 1.0;
	// glow :222-223, :240-241
	p->glowEnabled = 1;
	p->glowIntensity = 0.2;
	rgb16(p->glowColor1, 40984, 44713, 49490);
	rgb16(p->glowColor2, 57192, 60888, 62408);
	// screen-space AO :221; fog :247-272, 315-319; volumetric light DE factor :350
	p->SSAORandomMode = 0;
	p->basicFogEnabled = 0;
	p->basicFogCastShadows = 0;
	p->basicFogVisibility = 20.0;
	rgb16(p->basicFogColor, 59399, 61202, 65535);
	p->volFogEnabled = 0;
	p->distanceFogShadows = 0;
	p->volFogDensity = 0.5;
	p->volFogColour1Distance = 1.0;
	p->volFogColour2Distance = 2.0;
	p->volFogDistanceFactor = 1.0;
	p->volFogDistanceFromSurface = 1e-15;
	rgb16(p->volFogColour[0], 30000, 30000, 30000);
	rgb16(p->volFogColour[1], 0, 30000, 65535);
	rgb16(p->volFogColour[2], 65535, 65535, 65535);
	p->iterFogEnabled = 0;
	p->cloudsEnabled = 0;
	p->volumetricLightDEFactor = 1.0;
	p->antialiasingEnabled = 0;            // :69
	p->antialiasingSize = 2;               // :70
}

extern "C" int mbx_load_fract(const char *text, size_t len, mbx_params *params, mbx_fract_info *info)
{
	if (info)
	{
		std::memset(info, 0, sizeof(*info));
		info->formula = MBX_FORMULA_NONE;
	}
	if (!text || !params) return fail(info, "null argument");

	// split into lines (settings.cpp:656-666: trimmed text split on \n, \r\n or \r)
	const std::string all = trim(std::string(text, len));
	std::vector<std::string> lines;
	{
		size_t a = 0;
		for (size_t i = 0; i <= all.size(); i++)
		{
			if (i == all.size() || all[i] == '\n' || all[i] == '\r')
			{
				lines.push_back(all.substr(a, i - a));
				if (i < all.size() && all[i] == '\r' && i + 1 < all.size() && all[i + 1] == '\n') i++;
				a = i + 1;
			}
		}
	}

	// header (DecodeHeader, settings.cpp:540-601)
	Doc doc;
	if (lines.size() < 3) return fail(info, "not a Mandelbulber settings file: no header");
	if (!contains(lines[0], "Mandelbulber settings file"))
		return fail(info, "not a Mandelbulber settings file: wrong header");
	{
		const size_t pos = lines[1].rfind("version");
		if (pos == std::string::npos) return fail(info, "not a Mandelbulber settings file: no version");
		std::string num = trim(lines[1].substr(pos + 7));
		char *end = nullptr;
		const double v = std::strtod(num.c_str(), &end);
		if (num.empty() || end == num.c_str() || *end != '\0')
			return fail(info, "not a Mandelbulber settings file: wrong version number");
		doc.version = v;
	}
	if (doc.version > 2.35) return fail(info, "file saved by a newer Mandelbulber (version > 2.35)");
	if (contains(lines[2], "all parameters")) doc.format = 0;
	else if (contains(lines[2], "only modified parameters")) doc.format = 1;
	else if (contains(lines[2], "application settings")) return fail(info, "application settings, not a scene");
	else return fail(info, "not a Mandelbulber settings file: format not given in the header");
	if (info)
	{
		info->fileVersion = doc.version;
		info->format = doc.format;
	}

	// lines (Decode, settings.cpp:742-905)
	std::string section;
	for (size_t l = 3; l < lines.size(); l++)
	{
		const std::string &line = lines[l];
		if (line.size() >= 2 && line.front() == '[' && line.back() == ']')
		{
			section = line.substr(1, line.size() - 2);
			continue;
		}
		if (section == "description" || line.empty()) continue;
		bool ok = true;
		if (section == "main_parameters") ok = decode_one_line(doc, line, &doc.main, true, 0);
		else if (contains(section, "fractal"))
		{
			const int idx = (int)to_int(section.size() > 8 ? section.substr(8) : std::string());
			ok = decode_one_line(doc, line, &doc.fractal[idx], false, idx);
		}
		else if (section == "frames" || section == "keyframes") doc.animation = true;
		if (!ok)
		{
			doc.errors++;
			if (doc.firstError.empty()) doc.firstError = "malformed line " + std::to_string(l) + ": " + line;
			if (doc.errors > 3)
			{
				if (info)
				{
					info->lines = doc.lines;
					info->errors = doc.errors;
				}
				return fail(info, "too many errors in the settings file; first: " + doc.firstError);
			}
		}
	}

	const View M{&doc.main};
	auto F = [&](int n) { auto it = doc.fractal.find(n); return View{it == doc.fractal.end() ? nullptr : &it->second}; };
	const double ver = doc.version;
	uint32_t flags = 0;

	// ---- which fractal node is rendered -------------------------------------------------------
	// fractal container formulas: container 1 defaults to Mandelbulb, the others to none
	auto formula_of = [&](int n) { return (int)F(n).i("formula", n == 1 ? MBX_FORMULA_MANDELBULB : 0); };
	int objectId = 1;      // fractal container the node renders
	bool nodeEnabled = true;
	double rotation[3] = {0, 0, 0}, position[3] = {0, 0, 0}, repeat[3] = {0, 0, 0};
	double nodeScale = 1.0;
	int material = 1;
	bool juliaMode = false;
	double juliaC[3] = {0, 0, 0}, constFactor[3] = {1, 1, 1};
	double initialW = 0.0;
	int maxiter = 250;
	std::vector<int> hybridIds; // containers of a hybrid node, in sequence order
	bool hybridNode = false;
	if (ver < 2.35)
	{
		// MigrateToObjectsTree (settings.cpp:2551-2760) for the cases with one fractal node
		const bool hybrid = M.b("hybrid_fractal_enable", false);
		const bool boolean = M.b("boolean_operators", false);
		std::vector<int> enabled;
		for (int n = 1; n <= 9; n++)
			if (formula_of(n) != 0) enabled.push_back(n);
		// legacy global transform (MigrateLegacyParamsToFractal, settings.cpp:2368-2402)
		double globalRot[3], globalPos[3], globalRepeat[3];
		M.v3("fractal_rotation", globalRot, 0, 0, 0);
		M.v3("fractal_position", globalPos, 0, 0, 0);
		M.v3("repeat", globalRepeat, 0, 0, 0);
		auto transform_of = [&](int n) {
			const View fv = F(n);
			fv.v3("formula_rotation", rotation, 0, 0, 0);
			fv.v3("formula_position", position, 0, 0, 0);
			fv.v3("formula_repeat", repeat, 0, 0, 0);
			nodeScale = fv.d("formula_scale", 1.0);
			material = (int)fv.i("formula_material_id", 1);
			if (M.has("formula_material_id") && M.i("formula_material_id", 1) != 1)
				material = (int)M.i("formula_material_id", 1);
			auto nonzero = [](const double *v) { return v[0] != 0.0 || v[1] != 0.0 || v[2] != 0.0; };
			if (nonzero(globalRot)) std::memcpy(rotation, globalRot, sizeof(rotation));
			if (nonzero(globalPos)) std::memcpy(position, globalPos, sizeof(position));
			if (nonzero(globalRepeat)) std::memcpy(repeat, globalRepeat, sizeof(repeat));
		};
		if (boolean)
		{
			if (enabled.size() == 1)
			{
				objectId = enabled[0];
				transform_of(objectId);
				const std::string k = "_" + std::to_string(objectId);
				juliaMode = M.b("julia_mode" + k, false);
				M.v3("julia_c" + k, juliaC, 0, 0, 0);
				M.v3("fractal_constant_factor" + k, constFactor, 1, 1, 1);
				initialW = M.d("initial_waxis" + k, 0.0);
				maxiter = (int)M.i("formula_maxiter" + k, 250, 1, 9999);
				if (ver < 2.29) maxiter = (int)M.i("N", 250, 1, 99999); // N is always registered
				nodeEnabled = M.has("fractal_enable" + k) ? M.b("fractal_enable" + k, true)
				                                           : F(objectId).b("fractal_enable", true);
			}
			else
			{
				flags |= enabled.empty() ? MBX_FRACT_DISABLED : MBX_FRACT_HYBRID;
				if (!enabled.empty()) objectId = enabled[0];
			}
		}
		else if (!hybrid)
		{
			objectId = 1;
			transform_of(1);
			juliaMode = M.b("julia_mode", false);
			M.v3("julia_c", juliaC, 0, 0, 0);
			M.v3("fractal_constant_factor", constFactor, 1, 1, 1);
			initialW = M.d("initial_waxis", 0.0);
			maxiter = (int)M.i("N", 250, 1, 99999);
			nodeEnabled = F(1).b("fractal_enable", true);
		}
		else
		{
			// MigrateToObjectsTree, hybrid mode (settings.cpp:2690-2760): a hybrid node with the
			// legacy transform and sequence-wide values, and one child per container whose formula
			// is not none; a child switched off by fractal_enable(_N) leaves the render list
			// (objects_tree.cpp:234-270) and so the sequence
			hybridNode = true;
			for (int n : enabled)
				if (F(n).b("fractal_enable", true)) hybridIds.push_back(n);
			// the members' own transforms are not used: the distance code evaluates the sequence at
			// the hybrid node and skips its fractal children (calculate_distance.cpp:414-436)
			if (hybridIds.empty()) flags |= MBX_FRACT_DISABLED;
			objectId = hybridIds.empty() ? (enabled.empty() ? 1 : enabled[0]) : hybridIds[0];
			M.v3("fractal_rotation", rotation, 0, 0, 0);
			M.v3("fractal_position", position, 0, 0, 0);
			M.v3("repeat", repeat, 0, 0, 0);
			juliaMode = M.b("julia_mode", false);
			M.v3("julia_c", juliaC, 0, 0, 0);
			M.v3("fractal_constant_factor", constFactor, 1, 1, 1);
			initialW = M.d("initial_waxis", 0.0);
			maxiter = (int)M.i("N", 250, 1, 99999);
		}
		if (doc.primitives) flags |= MBX_FRACT_PRIMITIVES;
	}
	else
	{
		// objects tree (objects_tree.cpp:36-96): without node lines the tree is InitNodeParams(1)'s
		// one Mandelbulb node on fractal 1 (settings.cpp:715-733, initparameters.cpp:2102-2125)
		std::vector<int> ids(doc.nodeIds.begin(), doc.nodeIds.end());
		if (ids.empty()) ids.push_back(1);
		int fractalNodes = 0, otherNodes = 0, nodeId = -1;
		int hybridId = -1, hybridNodes = 0, childrenOfHybrid = 0, hybridParent = 0;
		std::vector<int> hybridChildren;
		for (int id : ids)
		{
			char pfx[32];
			std::snprintf(pfx, sizeof(pfx), "node_%04d_", id);
			const std::string *def = M.raw(std::string(pfx) + "definition");
			std::vector<std::string> parts = split(def ? *def : std::string("mandelbulb 1,1,1,0,1"), ',');
			if (parts.size() >= 5 && to_int(parts[2]) == 10)
			{
				hybridNodes++;
				hybridId = id;
				hybridParent = (int)to_int(parts[3]);
			}
		}
		for (int id : ids)
		{
			char pfx[32];
			std::snprintf(pfx, sizeof(pfx), "node_%04d_", id);
			const std::string *def = M.raw(std::string(pfx) + "definition");
			std::vector<std::string> parts = split(def ? *def : std::string("mandelbulb 1,1,1,0,1"), ',');
			if (parts.size() >= 5 && to_int(parts[2]) == 1 && hybridNodes == 1 && to_int(parts[3]) == hybridId)
			{
				childrenOfHybrid++;
				if (M.b(std::string(pfx) + "enabled", true)) hybridChildren.push_back((int)to_int(parts[4]));
			}
		}
		for (int id : ids)
		{
			char pfx[32];
			std::snprintf(pfx, sizeof(pfx), "node_%04d_", id);
			const std::string *def = M.raw(std::string(pfx) + "definition");
			const std::string d = def ? *def : std::string("mandelbulb 1,1,1,0,1");
			std::vector<std::string> parts = split(d, ',');
			const int type = parts.size() >= 5 ? (int)to_int(parts[2]) : 0;
			const bool en = M.b(std::string(pfx) + "enabled", true);
			if (type == 1)
			{
				fractalNodes++;
				if (nodeId < 0 || (en && !nodeEnabled))
				{
					nodeId = id;
					objectId = (int)to_int(parts[4]);
					nodeEnabled = en;
				}
			}
			else if (type == 2) flags |= MBX_FRACT_PRIMITIVES;
			else otherNodes++;
		}
		// a hybrid node at the root whose children are all the fractal nodes is a hybrid sequence
		const bool hybridTree = hybridNodes == 1 && hybridParent == 0 && otherNodes == 1 &&
			childrenOfHybrid == fractalNodes && fractalNodes > 0;
		if (hybridTree)
		{
			hybridNode = true;
			hybridIds = hybridChildren;
			nodeId = hybridId;
			if (hybridIds.empty()) flags |= MBX_FRACT_DISABLED;
			else objectId = hybridIds[0];
			char pfx[32];
			std::snprintf(pfx, sizeof(pfx), "node_%04d_", hybridId);
			nodeEnabled = M.b(std::string(pfx) + "enabled", true);
		}
		else if (fractalNodes > 1 || otherNodes > 0)
			flags |= MBX_FRACT_HYBRID;
		if (nodeId < 0) flags |= MBX_FRACT_DISABLED;
		else
		{
			char pfx[32];
			std::snprintf(pfx, sizeof(pfx), "node_%04d_", nodeId);
			const std::string p(pfx);
			M.v3(p + "rotation", rotation, 0, 0, 0);
			M.v3(p + "position", position, 0, 0, 0);
			M.v3(p + "repeat", repeat, 0, 0, 0);
			nodeScale = M.d(p + "scale", 1.0);
			material = (int)M.i(p + "material", 1);
			juliaMode = M.b(p + "julia_mode", false);
			M.v3(p + "julia_c", juliaC, 0, 0, 0);
			M.v3(p + "fractal_constant_factor", constFactor, 1, 1, 1);
			initialW = M.d(p + "initial_waxis", 0.0);
			maxiter = (int)M.i(p + "formula_maxiter", 250, 1, 9999);
		}
	}
	if (!nodeEnabled) flags |= MBX_FRACT_DISABLED;
	if (position[0] != 0.0 || position[1] != 0.0 || position[2] != 0.0 || repeat[0] != 0.0 ||
		repeat[1] != 0.0 || repeat[2] != 0.0 || nodeScale != 1.0 || M.b("limits_enabled", false))
		flags |= MBX_FRACT_TRANSFORM;

	// ---- the parameter block ------------------------------------------------------------------
	mbx_params p;
	mbx_params_default(&p);
	p.imageWidth = (int)M.i("image_width", 800, 5, 65535);
	p.imageHeight = (int)M.i("image_height", 600, 5, 65535);
	M.v3("camera", p.camera, 3.0, -6.0, 2.0);
	M.v3("target", p.target, 0.0, 0.0, 0.0);
	M.v3("camera_top", p.topVector, -0.1277753, 0.2555506, 0.958314);
	p.perspectiveType = (int)M.e("perspective_type", 0, {"three_point", "fish_eye", "equirectangular", "fish_eye_cut"});
	p.fov = M.d("fov", 53.13, 0.0, 720.0);
	if (ver < 2.21)
	{
		// Compatibility2 (settings.cpp:1611-1627): the old fov was the view-plane half-width
		double fov = p.fov;
		if (fov == 53.13) fov = 1.0;
		double deg = 0.0;
		switch (p.perspectiveType)
		{
			case 0: deg = std::atan(fov / 2.0) * 360.0 / M_PI; break;
			case 1:
			case 3: deg = fov * 180.0; break;
			case 2: deg = fov * 360.0; break;
		}
		p.fov = std::min(std::max(deg, 0.0), 720.0);
	}
	// stereo, and the Monte Carlo lens of the fish-eye and equirectangular projections
	// (render_worker.cpp:1845-1866)
	if (M.b("stereo_enabled", false) ||
		(p.perspectiveType != 0 && M.b("DOF_monte_carlo", false) && M.b("DOF_enabled", false)))
		flags |= MBX_FRACT_PERSPECTIVE;
	p.sweetSpotHAngle = M.d("sweet_spot_horizontal_angle", 0.0, -180.0, 180.0);
	p.sweetSpotVAngle = M.d("sweet_spot_vertical_angle", 0.0, -90.0, 90.0);
	p.minN = (int)M.i("minN", 1, 0, 99999);
	p.detailLevel = M.d("detail_level", 1.0, 1e-8, 1e8);
	p.DEThresh = M.d("DE_thresh", 0.01, 1e-15, 1e5);
	p.smoothness = M.d("smoothness", 1.0, 1e-15, 1e15);
	p.iterThreshMode = M.b("iteration_threshold_mode", false);
	p.DEFactor = M.d("DE_factor", 1.0, 1e-15, 1e15);
	p.viewDistanceMax = M.d("view_distance_max", 50.0, 1e-15, 1e15);
	p.interiorMode = M.b("interior_mode", false);
	p.constantDEThreshold = M.b("constant_DE_threshold", false);
	p.maxRaymarchingSteps = (int)M.i("max_raymarching_steps", 10000, 1, 100000000);
	p.advancedQuality = M.b("advanced_quality", false);
	p.absMaxMarchingStep = M.d("abs_max_marching_step", 3.0, 1e-15, 1e15);
	p.absMinMarchingStep = M.d("abs_min_marching_step", 1e-15, 1e-15, 1e15);
	p.relMaxMarchingStep = M.d("rel_max_marching_step", 1e4, 1e-15, 1e15);
	p.relMinMarchingStep = M.d("rel_min_marching_step", 1e-2, 1e-15, 1e15);
	p.detailSizeMax = M.d("detail_size_max", 1.0);
	p.detailSizeMin = M.d("detail_size_min", 1e-12);
	p.deltaDERelativeDelta = M.d("deltade_relative_delta", 0.01, 1e-15, 1e6);
	// sImageAdjustments (fractparams.cpp:155-159): float members read from double parameters
	p.image.brightness = (float)M.d("brightness", 1.0, 0.0, 1e15);
	p.image.contrast = (float)M.d("contrast", 1.0, 0.0, 1e15);
	p.image.gamma = (float)M.d("gamma", 1.0, 0.0, 1e15);
	p.image.saturation = (float)M.d("saturation", 1.0, 0.0, 1e15);
	p.image.hdr = M.b("hdr", false) ? 1 : 0;
	// sCommonParams::foldings (fractparams.cpp:346-351) and slow shading (fractparams.cpp:218)
	p.boxFolding = M.b("box_folding", false) ? 1 : 0;
	p.boxFoldingLimit = M.d("box_folding_limit", 1.0);
	p.boxFoldingValue = M.d("box_folding_value", 2.0);
	p.sphericalFolding = M.b("spherical_folding", false) ? 1 : 0;
	p.sphericalFoldingOuter = M.d("spherical_folding_outer", 1.0);
	p.sphericalFoldingInner = M.d("spherical_folding_inner", 0.5);
	p.slowShading = M.b("slow_shading", false) ? 1 : 0;

	// ambient occlusion
	p.ambientOcclusion = M.d("ambient_occlusion", 1.0, 0.0, 1e15);
	p.ambientOcclusionQuality = (int)M.i("ambient_occlusion_quality", 4, 1, 10);
	p.ambientOcclusionFastTune = M.d("ambient_occlusion_fast_tune", 1.0, 1e-5, 1e5);
	p.ambientOcclusionEnabled = M.b("ambient_occlusion_enabled", false);
	p.ambientOcclusionMode = (int)M.i("ambient_occlusion_mode", MBX_AO_SCREEN_SPACE);
	M.rgb("ambient_occlusion_color", p.ambientOcclusionColor, 65535, 65535, 65535);
	// screen-space AO (mode 2) is rendered by mbx_render_ssao after the render
	p.SSAORandomMode = M.b("SSAO_random_mode", false);
	if (p.ambientOcclusionEnabled && p.ambientOcclusionMode == MBX_AO_MULTIPLE_RAYS) flags |= MBX_FRACT_AO_LIGHTMAP;

	// background, fill light, Monte Carlo
	p.backgroundBrightness = M.d("background_brightness", 1.0, 0.0, 1000.0);
	p.backgroundGamma = M.d("background_gamma", 1.0, 0.0, 1000.0);
	M.rgb("background_color_1", p.backgroundColor1, 0, 38306, 65535);
	M.rgb("background_color_2", p.backgroundColor2, 65535, 65535, 65535);
	M.rgb("background_color_3", p.backgroundColor3, 0, 10000, 500);
	p.background3ColorsEnable = M.b("background_3_colors_enable", true);
	p.reflectionsMax = (int)M.i("reflections_max", 5, 0, 1000);
	M.rgb("fill_light_color", p.fillLightColor, 0, 0, 0);
	p.DOFMonteCarlo = M.b("DOF_monte_carlo", false);
	p.DOFSamples = (int)M.i("DOF_samples", 100);
	p.DOFMonteCarloGlobalIllumination = M.b("DOF_MC_global_illumination", false);
	p.monteCarloGIRadianceLimit = M.d("MC_GI_radiance_limit", 10.0, 0.001, 1e10);
	p.DOFMinSamples = (int32_t)M.i("DOF_min_samples", 10);
	p.DOFMaxNoise = M.d("DOF_max_noise", 1.0, 0.00001, 100.0);
	p.aoLightMapRotation = M.d("ao_light_map_rotation", 0.0, -180.0, 180.0); // initparameters.cpp:220
	// the Monte Carlo lens, chromatic aberration and soft shadows (fractparams.cpp:116-130, 180)
	p.DOFEnabled = M.b("DOF_enabled", false) ? 1 : 0;
	p.DOFFocus = M.d("DOF_focus", 6.0, 0.0, 10000.0);
	p.DOFRadius = M.d("DOF_radius", 10.0, 0.0, 200.0);
	p.DOFMonteCarloChromaticAberration = M.b("DOF_MC_CA_enable", false) ? 1 : 0;
	p.DOFMonteCarloCACameraDispersion = M.d("DOF_MC_CA_camera_dispersion", 1.0, 1e-15, 1000.0);
	p.monteCarloSoftShadows = M.b("MC_soft_shadows_enable", false) ? 1 : 0;
	// the OpenCL nebula engine renders these scenes (render_job.cpp:664-673)
	if (M.b("nebula_mode", false)) flags |= MBX_FRACT_NEBULA;
	p.N = (int32_t)M.i("N", 250, 1, 99999); // params->N: slow shading's pseudo-distance (fractparams.cpp:189)

	// material of the node (InitMaterialParams :1600-1810, material.cpp:330-349); a file that names
	// no material at all gets material 1 with its defaults (CheckIfMaterialsAreDefined)
	{
		const std::string m = "mat" + std::to_string(material) + "_";
		M.rgb(m + "surface_color", p.surfaceColor, 50000, 50000, 50000);
		p.shading = M.d(m + "shading", 1.0, 0.0, 1e15);
		p.luminosity = M.d(m + "luminosity", 0.0, 0.0, 1e15);
		M.rgb(m + "luminosity_color", p.luminosityColor, 65535, 65535, 65535);
		p.luminosityEmissive = M.d(m + "luminosity_emissive", 1.0, 0.0, 1e15);
		// palette colouring
		p.useColorsFromPalette = M.b(m + "use_colors_from_palette", true);
		p.surfaceGradientEnable = M.b(m + "surface_gradient_enable", true);
		p.coloringAlgorithm = (int)M.i(m + "fractal_coloring_algorithm", 0, 0, 4);
		p.coloringPreV215 = M.b(m + "fractal_coloring_color_preV215_false", false);
		p.coloring4D = M.b(m + "fractal_coloring_color_4D_enabled_false", false);
		p.coloringSpeed = M.d(m + "coloring_speed", 1.0, 0.0, 1e15);
		p.coloringPaletteOffset = M.d(m + "coloring_palette_offset", 0.0, 0.0, 256.0);
		p.coloringSphereRadius = M.d(m + "fractal_coloring_sphere_radius", 1.0, 0.0, 1e20);
		M.v3(m + "fractal_coloring_line_direction", p.coloringLineDirection, 1.0, 0.0, 0.0);
		const std::string *gradient = M.raw(m + "surface_color_gradient");
		const std::string gradientText = gradient ? *gradient : std::string(kDefaultSurfaceGradient);
		p.surfaceGradientCount = parse_gradient(gradientText, p.surfaceGradientPosition, p.surfaceGradientColor);
		if (ver < 2.19)
		{
			// Compatibility2 (settings.cpp:1563-1593): palette offset and speed were per colour
			const double paletteSize = (double)(split(gradientText, ' ').size() / 2);
			p.coloringPaletteOffset = std::min(std::max(p.coloringPaletteOffset / paletteSize, 0.0), 256.0);
			p.coloringSpeed = std::min(std::max(p.coloringSpeed * 10.0 / paletteSize, 0.0), 1e15);
		}
		if (p.useColorsFromPalette &&
			(M.b(m + "fractal_coloring_extra_color_enabled_false", false) ||
				M.b(m + "specular_gradient_enable", false) || M.b(m + "diffuse_gradient_enable", false) ||
				M.b(m + "luminosity_gradient_enable", false) || M.b(m + "roughness_gradient_enable", false) ||
				M.b(m + "reflectance_gradient_enable", false) || M.b(m + "transparency_gradient_enable", false)))
			flags |= MBX_FRACT_PALETTE;
		// specular
		p.specularPlasticEnable = M.b(m + "specular_plastic_enable", true);
		p.metallic = M.b(m + "metallic", true);
		p.specular = M.d(m + "specular", 5.0, 0.0, 1e15);
		p.specularWidth = M.d(m + "specular_width", 0.05, 1e-10, 1e15);
		p.specularMetallic = M.d(m + "specular_metallic", 2.0, 0.0, 1e15);
		p.specularMetallicWidth = M.d(m + "specular_metallic_width", 1.0, 1e-10, 1e15);
		p.specularMetallicRoughness = M.d(m + "specular_metallic_roughness", 0.1, 0.0, 1e4);
		M.rgb(m + "specular_color", p.specularColor, 65535, 65535, 65535);
		const bool rt = M.b("raytraced_reflections", false);
		if (rt && (M.d(m + "reflectance", 0.0) > 0.0 || M.d(m + "transparency_of_surface", 0.0) > 0.0))
			flags |= MBX_FRACT_REFLECTIONS;
		if (M.b(m + "iridescence_enabled", false) || M.b(m + "rough_surface", false) ||
			M.b(m + "subsurface_scattering", false) || M.b(m + "perlin_noise_enable", false))
			flags |= MBX_FRACT_MATERIAL;
		for (const char *t : {"use_color_texture", "use_diffusion_texture", "use_displacement_texture",
				 "use_luminosity_texture", "use_normal_map_texture", "use_reflectance_texture",
				 "use_transparency_texture"})
			if (M.b(m + t, false)) flags |= MBX_FRACT_TEXTURES;
	}

	// light #1: defined when the file names it; pre-2.25 files always get it (settings.cpp:1643-1654)
	{
		mbx_light &L = p.light;
		const bool defined = doc.lightsSeen.count(1) || ver < 2.25;
		if (!defined) L.enabled = 0;
		else
		{
			L.enabled = M.b("light1_enabled", true);
			L.castShadows = M.b("light1_cast_shadows", true);
			L.penetrating = M.b("light1_penetrating", true);
			L.relativePosition = M.b("light1_relative_position", true);
			if (ver < 2.25 && doc.lightsSeen.count(1))
				set3(L.rotation, M.d("light1_alpha", -45.0), M.d("light1_beta", 45.0), 0.0);
			else if (ver >= 2.25)
				M.v3("light1_rotation", L.rotation, -45.0, 45.0, 0.0);
			L.intensity = M.d("light1_intensity", 1.0, 0.0, 1e10);
			L.visibility = M.d("light1_visibility", 1.0, 0.0, 1e10);
			L.size = M.d("light1_size", 0.5, 0.0, 1e10);
			L.softShadowCone = M.d("light1_soft_shadow_cone", 1.0, 0.0, 1e10);
			L.contourSharpness = M.d("light1_contour_sharpness", 1.0, 0.0, 1e10);
			M.rgb("light1_color", L.color, 65535, 65535, 65535);
			// settings.cpp:1700-1712: pre-2.28 iteration fog scaled light #1 by its brightness boost
			if (ver < 2.28 && M.b("iteration_fog_enable", false))
			{
				double boost = M.d("iteration_fog_brightness_boost", 1.0, 0.0, 1e6);
				if (ver <= 2.12) boost = 100.0; // settings.cpp:1547-1551
				L.intensity /= boost;
			}
			// cLight (light.cpp:77-87): intensity, visibility and size scale by the all_lights_ factors
			L.intensity *= M.d("all_lights_intensity", 1.0, 0.0, 100000.0);
			L.visibility *= M.d("all_lights_visibility", 1.0, 0.0, 100000.0);
			L.size *= M.d("all_lights_size", 1.0, 0.0, 100000.0);
			const long long type = M.e("light1_type", 0, {"directional", "point", "spot", "projection", "beam"});
			if (L.enabled && (type != 0 || M.b("light1_volumetric", false))) flags |= MBX_FRACT_LIGHTS;
		}
		for (int id : doc.lightsSeen)
			if (id >= 2 && M.b("light" + std::to_string(id) + "_enabled", false)) flags |= MBX_FRACT_LIGHTS;
		if (M.b("random_lights_group", false) || M.b("fake_lights_enabled", false)) flags |= MBX_FRACT_LIGHTS;
	}

	// glow (rendered); the volumetric shader's own step recalculation and glow on GI bounces are not
	p.glowEnabled = M.b("glow_enabled", true);
	p.glowIntensity = M.d("glow_intensity", 0.2, 0.0, 1e15);
	M.rgb("glow_color_1", p.glowColor1, 40984, 44713, 49490);
	M.rgb("glow_color_2", p.glowColor2, 57192, 60888, 62408);
	if (p.glowEnabled && p.glowIntensity > 0.0 &&
		(M.d("volumetric_light_DE_Factor", 1.0, 1e-15, 1e15) != 1.0 ||
			(p.DOFMonteCarlo && p.DOFMonteCarloGlobalIllumination && M.b("MC_global_illumination_volumetric", false))))
		flags |= MBX_FRACT_GLOW;
	// fog (initparameters.cpp:247-272, 315-319): basic and distance fog are rendered, lit by
	// (1, 1, 1); iteration fog, clouds, shadows in fog and the volumetric shader's own march (a DE
	// factor other than 1) are not
	p.basicFogEnabled = M.b("basic_fog_enabled", false);
	p.basicFogVisibility = M.d("basic_fog_visibility", 20.0, 0.0, 1e15);
	M.rgb("basic_fog_color", p.basicFogColor, 59399, 61202, 65535);
	p.basicFogCastShadows = M.b("basic_fog_cast_shadows", false);
	p.volFogEnabled = M.b("volumetric_fog_enabled", false);
	p.volFogDensity = M.d("volumetric_fog_density", 0.5, 0.0, 1e15);
	p.volFogColour1Distance = M.d("volumetric_fog_colour_1_distance", 1.0, 0.0, 1e15);
	p.volFogColour2Distance = M.d("volumetric_fog_colour_2_distance", 2.0, 0.0, 1e15);
	p.volFogDistanceFactor = M.d("volumetric_fog_distance_factor", 1.0, 0.0, 1e15);
	p.volFogDistanceFromSurface = M.d("volumetric_fog_distance_from_surface", 1e-15, 0.0, 1e15);
	M.rgb("fog_color_1", p.volFogColour[0], 30000, 30000, 30000);
	M.rgb("fog_color_2", p.volFogColour[1], 0, 30000, 65535);
	M.rgb("fog_color_3", p.volFogColour[2], 65535, 65535, 65535);
	p.distanceFogShadows = M.b("distance_fog_shadows", false);
	p.iterFogEnabled = M.b("iteration_fog_enable", false);
	p.cloudsEnabled = M.b("clouds_enable", false);
	p.volumetricLightDEFactor = M.d("volumetric_light_DE_Factor", 1.0, 1e-15, 1e15);
	{
		const bool fogOn = p.basicFogEnabled || p.volFogEnabled;
		if (p.iterFogEnabled || p.cloudsEnabled || (p.basicFogEnabled && p.basicFogCastShadows) ||
			(p.volFogEnabled && p.distanceFogShadows) || (fogOn && p.volumetricLightDEFactor != 1.0))
			flags |= MBX_FRACT_FOG;
	}
	// other effects the path does not render
	if (M.b("textured_background", false) || M.b("env_mapping_enable", false)) flags |= MBX_FRACT_TEXTURES;
	// brightness, contrast, gamma, saturation and hdr are read into params.image (mbx_compile_image)
	// DOF without Monte Carlo is the post-processing blur; with it, the lens of the pixel loop
	if (M.b("hdr_blur_enabled", false) || (M.b("DOF_enabled", false) && !p.DOFMonteCarlo) ||
		M.b("post_chromatic_aberration_enabled", false))
		flags |= MBX_FRACT_POSTPROCESS;
	if (doc.animation) flags |= MBX_FRACT_ANIMATION;
	p.antialiasingEnabled = M.b("antialiasing_enabled", false);
	p.antialiasingSize = (int)M.i("antialiasing_size", 2, 1, 10);
	{
		// the projections and antialiasing render without the lens variant's effects
		const bool lensFx = (p.DOFMonteCarlo && (p.DOFEnabled || p.DOFMonteCarloChromaticAberration)) ||
			p.basicFogEnabled || p.volFogEnabled;
		if (lensFx && p.perspectiveType != 0) flags |= MBX_FRACT_PERSPECTIVE;
		if (lensFx && p.antialiasingEnabled) flags |= MBX_FRACT_ANTIALIASING;
	}

	// ---- the fractal containers (sFractal, fractal.cpp:41-120) ---------------------------------
	auto read_container = [&](int n, mbx_fractal &fr) {
		const View fv = F(n);
		fr.formula = formula_of(n);
		fr.formulaIterations = (int)fv.i("formula_iterations", 1, 1, 65536);
		fr.formulaWeight = fv.d("formula_weight", 1.0, 0.0, 1.0);
		fr.formulaStartIteration = (int)fv.i("formula_start_iteration", 0, 0, 65536);
		fr.formulaStopIteration = (int)fv.i("formula_stop_iteration", 250, 0, 65536);
		fr.checkForBailout = fv.b("check_for_bailout", true);
		fr.dontAddCConstant = fv.b("dont_add_c_constant", false);
		// MigrateLegacyParamsToFractal (settings.cpp:2328-2340): a true main-section
		// dont_add_c_constant_N sets fractal N's flag
		if (ver < 2.35 && n <= 8 && M.b("dont_add_c_constant_" + std::to_string(n), false)) fr.dontAddCConstant = 1;
		fr.bulb.power = fv.d("power", 9.0);
		fr.bulb.alphaAngleOffset = fv.d("alpha_angle_offset", 0.0);
		fr.bulb.betaAngleOffset = fv.d("beta_angle_offset", 0.0);
		mbx_fractal_mandelbox &mb = fr.mandelbox;
		mb.scale = fv.d("mandelbox_scale", 2.0);
		mb.foldingLimit = fv.d("mandelbox_folding_limit", 1.0);
		mb.foldingValue = fv.d("mandelbox_folding_value", 2.0);
		mb.foldingSphericalMin = fv.d("mandelbox_folding_min_radius", 0.5);
		mb.foldingSphericalFixed = fv.d("mandelbox_folding_fixed_radius", 1.0);
		fv.v3("mandelbox_offset", mb.offset, 0, 0, 0);
		fv.v3("mandelbox_rotation_main", mb.rotationMain, 0, 0, 0);
		for (int i = 1; i <= 3; i++)
		{
			fv.v3("mandelbox_rotation_neg_" + std::to_string(i), mb.rotation[0][i - 1], 0, 0, 0);
			fv.v3("mandelbox_rotation_pos_" + std::to_string(i), mb.rotation[1][i - 1], 0, 0, 0);
		}
		fv.v3("mandelbox_color", mb.colorFactor, 0.03, 0.05, 0.07);
		mb.colorFactorSp1 = fv.d("mandelbox_color_Sp1", 0.2);
		mb.colorFactorSp2 = fv.d("mandelbox_color_Sp2", 0.2);
		mb.colorFactorR = fv.d("mandelbox_color_R", 0.0);
		mb.rotationsEnabled = fv.b("mandelbox_rotations_enabled", false);
		mb.mainRotationEnabled = fv.b("mandelbox_main_rotation_enabled", false);
		mbx_fractal_ifs &ifs = fr.IFS;
		ifs.scale = fv.d("IFS_scale", 2.0);
		fv.v3("IFS_offset", ifs.offset, 1, 0, 0);
		fv.v3("IFS_rotation", ifs.rotation, 0, 0, 0);
		fv.v3("IFS_edge", ifs.edge, 0, 0, 0);
		for (int i = 0; i < MBX_IFS_VECTOR_COUNT; i++)
		{
			const std::string k = "_" + std::to_string(i);
			fv.v3("IFS_direction" + k, ifs.direction[i], 1, 0, 0);
			fv.v3("IFS_rotations" + k, ifs.rotations[i], 0, 0, 0);
			ifs.distance[i] = fv.d("IFS_distance" + k, 0.0);
			ifs.intensity[i] = fv.d("IFS_intensity" + k, 1.0);
			ifs.enabled[i] = fv.b("IFS_enabled" + k, false);
		}
		ifs.absX = fv.b("IFS_abs_x", false);
		ifs.absY = fv.b("IFS_abs_y", false);
		ifs.absZ = fv.b("IFS_abs_z", false);
		ifs.mengerSpongeMode = fv.b("IFS_menger_sponge_mode", false);
		ifs.rotationEnabled = fv.b("IFS_rotation_enabled", false);
		ifs.edgeEnabled = fv.b("IFS_edge_enabled", false);
		// settings.cpp:1507-1513, 1629-1641
		if (ver <= 2.06 && n <= 4) ifs.rotationEnabled = 1;
		if (ver < 2.22 && (ifs.edge[0] != 0.0 || ifs.edge[1] != 0.0 || ifs.edge[2] != 0.0)) ifs.edgeEnabled = 1;
		mbx_fractal_transform &t = fr.transform;
		t.scale3 = fv.d("transf_scale_3", 3.0);
		t.aexionCadd = fv.d("cadd", -1.3);
		t.foldingIntPowFoldFactor = fv.d("boxfold_bulbpow2_folding_factor", 2.0);
		t.foldingIntPowZFactor = fv.d("boxfold_bulbpow2_z_factor", 5.0);
		t.analyticDEEnabled = fv.b("analyticDE_enabled", true);
		t.analyticDEEnabledFalse = fv.b("analyticDE_enabled_false", false);
		t.analyticDEScale1 = fv.d("analyticDE_scale_1", 1.0);
		t.analyticDEOffset1 = fv.d("analyticDE_offset_1", 1.0);
		t.analyticDEOffset2 = fv.d("analyticDE_offset_2", 1.0);
		// what the device path does not render: formulas outside the table, rotated Mandelbox folds,
		// the Menger sponge's fold colour with palette colouring
		static const int kTable[] = {MBX_FORMULA_MANDELBULB, MBX_FORMULA_MANDELBULB_POWER2, MBX_FORMULA_HYPERCOMPLEX,
			MBX_FORMULA_QUATERNION, MBX_FORMULA_MENGER_SPONGE, MBX_FORMULA_MANDELBOX, MBX_FORMULA_KALEIDOSCOPIC_IFS,
			MBX_FORMULA_XENODREAMBUIE, MBX_FORMULA_BOX_FOLD_BULB_POW2, MBX_FORMULA_AEXION};
		bool known = false;
		for (int id : kTable) known |= fr.formula == id;
		if (fr.formula != MBX_FORMULA_NONE && !known) flags |= MBX_FRACT_FORMULA;
		if (fr.formula == MBX_FORMULA_MENGER_SPONGE && p.useColorsFromPalette &&
			fv.b("fold_color_aux_color_enabled_false", false))
			flags |= MBX_FRACT_PALETTE;
	};
	{
		// the rendered container, or the first of the hybrid's
		mbx_fractal &fr = p.fractal;
		read_container(objectId, fr);
		if (fr.formula == MBX_FORMULA_NONE) flags |= MBX_FRACT_DISABLED;
		if (hybridNode && !hybridIds.empty())
		{
			p.hybridCount = (int)hybridIds.size();
			for (size_t k = 1; k < hybridIds.size(); k++) read_container(hybridIds[k], p.hybridFractals[k - 1]);
			p.repeatFrom = (int)M.i("repeat_from", 1, 1, 9);
			// a hybrid with palette colouring and the extra colour options uses scales this path
			// does not apply (fractal_coloring.cpp:309-315)
			if (p.useColorsFromPalette && M.b(std::string("mat") + std::to_string(material) +
						"_fractal_coloring_extra_color_options_false", false))
				flags |= MBX_FRACT_PALETTE;
		}
		p.linearDEOffset = M.d("linear_DE_offset", 0.0);
		// sequence-wide values (fractparams.cpp, hybrid_fractal_sequences.cpp:181-287)
		fr.formulaMaxiter = maxiter;
		fr.deltaDEMethod = (int)M.i("delta_DE_method", 0, 0, 2);
		fr.deltaDEFunction = (int)M.i("delta_DE_function", 0, 0, 6);
		// settings.cpp:1500-1505: before 2.07 every non-linear DE function became logarithmic;
		// settings.cpp:1596-1606: before 2.20 "preferred" method meant "preferred" function
		if (ver <= 2.06 && fr.deltaDEFunction != MBX_DE_FUNCTION_LINEAR) fr.deltaDEFunction = MBX_DE_FUNCTION_LOGARITHMIC;
		if (ver < 2.20 && fr.deltaDEMethod == 0) fr.deltaDEFunction = 0;
		fr.useDefaultBailout = M.b("use_default_bailout", true);
		fr.bailout = M.d("bailout", 100.0, 1.0, 1e15);
		fr.juliaMode = juliaMode;
		std::memcpy(fr.juliaC, juliaC, sizeof(juliaC));
		std::memcpy(fr.constantFactor, constFactor, sizeof(constFactor));
		fr.initialWAxis = initialW;
		std::memcpy(p.formulaRotation, rotation, sizeof(rotation));
	}

	*params = p;
	if (info)
	{
		info->lines = doc.lines;
		info->errors = doc.errors;
		info->formula = p.fractal.formula;
		info->unsupported = flags;
		std::snprintf(info->message, sizeof(info->message), "%s", doc.firstError.c_str());
	}
	return MBX_OK;
}

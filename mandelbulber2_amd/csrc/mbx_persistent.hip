// mbx_persistent.hip - persistent work-stealing raymarch kernel for gfx950.
//
// A fixed grid of waves drains a queue of 8x8 pixel tiles covering every frame of a batch.
// Inside a wave each lane owns one pixel and runs the pixel's whole program -- primary march,
// binary search, 6-tap normal, AO rays, shadow march, GI bounces, Monte Carlo samples -- as a
// state machine whose every step is one DE query.  The loop body advances each lane's DE by one
// formula iteration; a lane whose DE finished takes the next state transition and, when its
// pixel is done, takes the next pixel of the wave's tile pool.  So the 64 lanes stay busy on
// formula iterations whatever the spread of step counts between rays, which is where the plain
// one-ray-per-lane kernel loses most of its time.
//
// The per-pixel semantics are those of mbx_render.hip (and of the reference path it cites):
// same march, same random stream per pixel, same shading order.
#include <hip/hip_runtime.h>

#include "mbx_device.h"
#include "mbx_launch.h"

namespace mbx
{

// Actions of a pixel's program.  A DE query remembers the action that consumes its result.
enum Act : int
{
	A_SAMPLE = 0,
	A_MARCH_DE,
	A_MARCH_END,
	A_SEARCH_CHECK,
	A_SEARCH_DE,
	A_SEARCH_DONE,
	A_NORMALS,
	A_NORMAL_DE,
	A_SHADE,
	A_AO_RAY,
	A_AO_RAY_DE,
	A_AO_NEXT,
	A_AO_FAST_DE,
	A_LIGHT,
	A_SHADOW_DE,
	A_COMBINE,
	A_GI_BOUNCE,
	A_GI_NEXT,
	A_GI_DE,
	A_GI_DONE,
	A_SAMPLE_DONE,
};

__device__ __forceinline__ float thresh_at(const FrameConst &f, V3 p)
{
	if (f.constantDEThreshold) return f.DEThresh;
	return length(v3(f.camera[0], f.camera[1], f.camera[2]) - p) * f.threshK;
}

// BackgroundShader: 3-colour gradient and the directional light's disk (shader_background.cpp)
__device__ __forceinline__ V3 background(const FrameConst &f, V3 viewVector)
{
	const V3 vn = normalize(viewVector);
	float c[3];
	const float grad = vn.z + 1.0f;
#pragma unroll
	for (int k = 0; k < 3; k++)
	{
		float v;
		if (!f.background3Colors)
			v = f.bg1[k];
		else if (grad < 1.0f)
			v = f.bg3[k] * (1.0f - grad) + f.bg2[k] * grad;
		else
			v = f.bg2[k] * (2.0f - grad) + f.bg1[k] * (grad - 1.0f);
		c[k] = v * f.bgBrightness;
	}
	V3 out;
	if (f.bgInvGamma == 1.0f)
		out = v3(c[0], c[1], c[2]);
	else
		out = v3(powf(c[0], f.bgInvGamma), powf(c[1], f.bgInvGamma), powf(c[2], f.bgInvGamma));
	if (f.lightEnabled)
	{
		const V3 ld = v3(f.lightDir[0], f.lightDir[1], f.lightDir[2]);
		const float a = -(dot(vn, ld) - 1.0f) * 360.0f / f.lightSize;
		const float pw = 6.0f * f.lightContourSharpness;
		const float ap = (pw == 6.0f) ? (a * a) * (a * a) * (a * a) : powf(a, pw);
		const float intensity = 1.0f / (1.0f + ap) * f.lightVisibility * f.lightIntensity;
		out = out + v3(f.lightColor[0], f.lightColor[1], f.lightColor[2]) * intensity;
	}
	return out;
}

__device__ __forceinline__ V3 clamp_rgb(V3 c, float lim)
{
	return v3(fminf(fmaxf(c.x, 0.0f), lim), fminf(fmaxf(c.y, 0.0f), lim), fminf(fmaxf(c.z, 0.0f), lim));
}

// Lane state of one pixel's program.
struct Lane
{
	int pix; // pixel index in the current frame
	int phase; // action consuming the pending DE result; -1 when the pixel is complete
	uint32_t seed;
	// primary ray
	V3 dir;
	float invDirLen;
	float scan, step, dist, thr;
	int counter, lastIters, sub;
	int mfound, mdead;
	int rep;
	V3 acc, sampleRGB;
	// hit being shaded (primary hit or GI bounce hit)
	V3 hp, hn;
	float hthr, hdelta;
	int ctx; // 0 primary hit, 1 GI bounce hit
	V3 primaryRGB;
	// shadow / AO marches
	float si, sstep, shade, maxSoft, shadowVal, aoShadow, aoLast;
	int scount, aoIdx, aoEnd;
	V3 aoAcc;
	// GI
	int bounce;
	V3 giOut, giDir, giOrigin;
	// outputs: sample 0 (found, iters, normal) and last sample (depth)
	float outDepth;
	int outFound, outIters, outSteps;
	V3 outNormal;
};

template <int KIND, int FEAT>
struct Machine
{
	static constexpr bool kShadowOn = (FEAT & kFeatShadow) != 0;
	static constexpr bool kAoOn = (FEAT & kFeatAO) != 0;
	static constexpr bool kMcOn = (FEAT & kFeatMC) != 0;

	const FrameConst &f;
	const V3 cam;
	Lane L;
	DEState<KIND> de;
	float detail;
	Work w;

	__device__ __forceinline__ Machine(const FrameConst &frame)
		: f(frame), cam(v3(frame.camera[0], frame.camera[1], frame.camera[2]))
	{
		w.de_evals = 0u;
		w.iterations = 0u;
		L.pix = -1;
		L.phase = -1;
	}

	__device__ __forceinline__ void query(V3 q, float detailSize, int consumer)
	{
		de_begin<KIND>(de, q);
		detail = detailSize;
		L.phase = consumer;
	}

	__device__ __forceinline__ void begin_pixel(int pix)
	{
		L.pix = pix;
		L.seed = pixel_seed((uint32_t)pix, f.randomSeed);
		L.rep = 0;
		L.acc = v3(0.0f, 0.0f, 0.0f);
		L.outSteps = 0;
		run(A_SAMPLE, 0.0f, 0);
	}

	__device__ __forceinline__ void on_de(float d, int iters) { run(L.phase, d, iters); }

	__device__ __forceinline__ void record_primary(bool found)
	{
		L.outDepth = found ? L.scan : 1e20f;
		if (L.rep == 0)
		{
			L.outFound = found ? 1 : 0;
			L.outIters = L.lastIters;
			L.outNormal = v3(0.0f, 0.0f, 0.0f);
		}
	}

	__device__ __forceinline__ V3 normal_tap(int k) const
	{
		const float d = L.hthr * f.smoothness * ((k & 1) ? -1.0f : 1.0f);
		const int axis = k >> 1;
		return v3(axis == 0 ? d : 0.0f, axis == 1 ? d : 0.0f, axis == 2 ? d : 0.0f);
	}

	__device__ __forceinline__ V3 ao_vector(int i) const
	{
		return v3(f.aoVectors[i][0], f.aoVectors[i][1], f.aoVectors[i][2]);
	}

	// Run the pixel's program from action `act` until it issues the next DE query or finishes.
	__device__ __forceinline__ void run(int act, float d, int iters)
	{
		const V3 lv = v3(f.lightDir[0], f.lightDir[1], f.lightDir[2]);
		const V3 colour = v3(f.surfaceColor[0], f.surfaceColor[1], f.surfaceColor[2]);
		for (;;)
		{
			switch (act)
			{
				// ---- primary ray: doWork sample set-up and RayMarching (render_worker.cpp:708-951)
				case A_SAMPLE:
				{
					const int xs = L.pix % f.width, ys = L.pix / f.width;
					float ix = (f.invWidth * (float)xs - 0.5f) * f.aspectRatio;
					float iy = 0.5f - f.invHeight * (float)ys;
					if (kMcOn && f.monteCarlo)
					{
						ix += ((float)random_mod(L.seed, 1000) * 1e-3f - 0.5f) * f.invWidth * f.aspectRatio;
						iy += ((float)random_mod(L.seed, 1000) * 1e-3f - 0.5f) * f.invHeight;
					}
					const V3 vv = normalize(v3(ix * f.fov, 1.0f, iy * f.fov));
					L.dir = normalize(rot(f.mRot, vv));
					L.invDirLen = 1.0f / length(L.dir);
					L.scan = 0.0f;
					L.step = 0.0f;
					L.dist = 0.0f;
					L.counter = 1;
					L.lastIters = 0;
					L.ctx = 0;
					L.thr = thresh_at(f, cam);
					query(cam, L.thr, A_MARCH_DE);
					return;
				}
				case A_MARCH_DE:
				{
					L.dist = d;
					L.lastIters = iters;
					if (d < L.thr)
					{
						L.mfound = 1;
						L.mdead = 0;
						act = A_MARCH_END;
						break;
					}
					const float jitter = 1.0f - (float)random_mod(L.seed, 1000) * 1e-4f;
					L.step = fminf((d - 0.5f * L.thr) * f.DEFactor * jitter, 3.0f);
					L.scan += L.step * L.invDirLen;
					if (L.scan > f.viewDistanceMax || L.counter >= f.maxRaymarchingSteps)
					{
						L.mfound = 0;
						L.mdead = 0;
						act = A_MARCH_END;
						break;
					}
					L.counter++;
					const V3 q = cam + L.dir * L.scan;
					if ((q.x == de.p.x && q.y == de.p.y && q.z == de.p.z) || !finite3(q))
					{
						L.mfound = 1;
						L.mdead = 1;
						act = A_MARCH_END;
						break;
					}
					L.thr = thresh_at(f, q);
					query(q, L.thr, A_MARCH_DE);
					return;
				}
				case A_MARCH_END:
				{
					if (L.mfound && !L.mdead)
					{
						L.step *= 0.5f;
						L.sub = 0;
						act = A_SEARCH_CHECK;
						break;
					}
					record_primary(L.mfound != 0);
					if (L.mfound)
					{
						L.hp = cam + L.dir * L.scan;
						L.hthr = L.thr;
						act = A_NORMALS;
						break;
					}
					L.sampleRGB = background(f, L.dir);
					act = A_SAMPLE_DONE;
					break;
				}
				case A_SEARCH_CHECK:
				{
					L.counter++;
					const float lim = L.thr * f.searchLimit;
					if (L.dist < L.thr && L.dist > lim)
					{
						act = A_SEARCH_DONE;
						break;
					}
					if (L.dist > L.thr)
						L.scan += L.step;
					else if (L.dist < lim)
						L.scan -= L.step;
					const V3 q = cam + L.dir * L.scan;
					L.thr = thresh_at(f, q);
					query(q, L.thr, A_SEARCH_DE);
					return;
				}
				case A_SEARCH_DE:
				{
					L.dist = d;
					L.lastIters = iters;
					L.step *= 0.5f;
					L.sub++;
					act = (L.sub >= 30) ? A_SEARCH_DONE : A_SEARCH_CHECK;
					break;
				}
				case A_SEARCH_DONE:
				{
					record_primary(true);
					L.hp = cam + L.dir * L.scan;
					L.hthr = L.thr;
					act = A_NORMALS;
					break;
				}
				// ---- CalculateNormals (shader_calculate_normals.cpp:44-123)
				case A_NORMALS:
				{
					L.sub = 0;
					L.hn = v3(0.0f, 0.0f, 0.0f);
					query(L.hp + normal_tap(0), L.hthr, A_NORMAL_DE);
					return;
				}
				case A_NORMAL_DE:
				{
					const float sd = (L.sub & 1) ? -d : d;
					const int axis = L.sub >> 1;
					L.hn.x += axis == 0 ? sd : 0.0f;
					L.hn.y += axis == 1 ? sd : 0.0f;
					L.hn.z += axis == 2 ? sd : 0.0f;
					L.sub++;
					if (L.sub < 6)
					{
						query(L.hp + normal_tap(L.sub), L.hthr, A_NORMAL_DE);
						return;
					}
					V3 n = L.hn;
					if (n.x == 0.0f && n.y == 0.0f && n.z == 0.0f)
						n = v3(1.0f, 0.0f, 0.0f);
					else
						n = normalize(n);
					if (!finite3(n)) n = v3(1.0f, 0.0f, 0.0f);
					L.hn = n;
					if (L.ctx == 0)
					{
						if (L.rep == 0) L.outNormal = n;
						L.hdelta = length(cam - L.hp) * f.deltaK;
					}
					else
						L.hdelta = L.hthr;
					act = A_SHADE;
					break;
				}
				// ---- ObjectShader: AO, then LightShading with AuxShadow (shader_object.cpp:40-229)
				case A_SHADE:
				{
					L.aoAcc = v3(0.0f, 0.0f, 0.0f);
					act = A_LIGHT;
					if (kAoOn && f.aoEnabled)
					{
						if (f.aoMode == MBX_AO_FAST)
						{
							if (f.aoQuality * f.aoQuality > 1)
							{
								L.aoIdx = 1;
								L.aoLast = 1e20f;
								L.maxSoft = 0.0f;
								query(L.hp + L.hn * L.hthr, L.hthr, A_AO_FAST_DE);
								return;
							}
							L.aoAcc = v3(1.0f, 1.0f, 1.0f);
						}
						else
						{
							int start = 0, end = f.aoCount - 1;
							if (kMcOn && f.monteCarlo)
							{
								start = random_mod(L.seed, f.aoCount - 1);
								end = start;
							}
							L.aoIdx = start;
							L.aoEnd = end;
							act = A_AO_RAY;
						}
					}
					break;
				}
				case A_AO_RAY:
				{
					if (!kAoOn)
					{
						act = A_LIGHT;
						break;
					}
					if (L.aoIdx > L.aoEnd)
					{
						if (!(kMcOn && f.monteCarlo)) L.aoAcc = L.aoAcc * (1.0f / (float)f.aoCount);
						act = A_LIGHT;
						break;
					}
					const float end_dist = L.hdelta * (float)f.height;
					if (L.hthr < end_dist)
					{
						L.si = L.hthr;
						query(L.hp + ao_vector(L.aoIdx) * L.si, L.hthr, A_AO_RAY_DE);
						return;
					}
					L.aoShadow = 1.0f;
					act = A_AO_NEXT;
					break;
				}
				case A_AO_RAY_DE:
				{
					if (!kAoOn)
					{
						act = A_LIGHT;
						break;
					}
					const float end_dist = L.hdelta * (float)f.height;
					if (d < L.hthr)
					{
						L.aoShadow = fmaxf(1.0f - (end_dist - L.si) / end_dist, 0.0f);
						act = A_AO_NEXT;
						break;
					}
					L.si += d * 2.0f;
					if (L.si < end_dist)
					{
						query(L.hp + ao_vector(L.aoIdx) * L.si, L.hthr, A_AO_RAY_DE);
						return;
					}
					L.aoShadow = 1.0f;
					act = A_AO_NEXT;
					break;
				}
				case A_AO_NEXT:
				{
					if (!kAoOn)
					{
						act = A_LIGHT;
						break;
					}
					L.aoAcc = L.aoAcc +
						v3(f.aoColors[L.aoIdx][0], f.aoColors[L.aoIdx][1], f.aoColors[L.aoIdx][2]) * L.aoShadow;
					L.aoIdx++;
					act = A_AO_RAY;
					break;
				}
				case A_AO_FAST_DE:
				{
					if (!kAoOn)
					{
						act = A_LIGHT;
						break;
					}
					const float scan = (float)(L.aoIdx * L.aoIdx) * L.hthr;
					float dd = d;
					if (dd > L.aoLast * 2.0f) dd = L.aoLast * 2.0f;
					L.aoLast = dd;
					L.maxSoft += exp2f(-(float)L.aoIdx) * (scan - f.aoFastTune * dd) / L.hthr;
					L.aoIdx++;
					if (L.aoIdx < f.aoQuality * f.aoQuality)
					{
						query(L.hp + L.hn * ((float)(L.aoIdx * L.aoIdx) * L.hthr), L.hthr, A_AO_FAST_DE);
						return;
					}
					const float ao = fmaxf(1.0f - 0.2f * L.maxSoft, 0.0f);
					L.aoAcc = v3(ao, ao, ao);
					act = A_LIGHT;
					break;
				}
				case A_LIGHT:
				{
					L.shade = 0.0f;
					L.shadowVal = 1.0f;
					act = A_COMBINE;
					if (f.lightEnabled)
					{
						float shade = fmaxf(dot(L.hn, lv), 0.0f);
						shade = 1.0f - f.shading + shade * f.shading;
						shade = fminf(shade * f.lightIntensity, 500.0f);
						L.shade = shade;
						if (kShadowOn && f.castShadows)
						{
							if (shade > 0.001f)
							{
								const float distance =
									f.penetrating ? L.hdelta * (float)f.height : f.viewDistanceMaxLight;
								L.si = L.hthr;
								L.scount = 0;
								L.maxSoft = 0.0f;
								if (L.si < distance)
								{
									query(L.hp + lv * L.si, L.hthr, A_SHADOW_DE);
									return;
								}
							}
							else
								L.shadowVal = 0.0f;
						}
					}
					break;
				}
				case A_SHADOW_DE:
				{
					if (!kShadowOn)
					{
						act = A_COMBINE;
						break;
					}
					const float distance = f.penetrating ? L.hdelta * (float)f.height : f.viewDistanceMaxLight;
					if (f.softShadows)
					{
						float angle = (d - L.hthr) / L.si;
						if (angle < 0.0f || d < L.hthr) angle = 0.0f;
						float soft = 1.0f - angle / f.softRange;
						if (f.penetrating) soft *= (distance - L.si) / distance;
						L.maxSoft = fmaxf(L.maxSoft, fmaxf(soft, 0.0f));
					}
					if (d < L.hthr)
					{
						const float sh = f.penetrating ? fmaxf(1.0f - (distance - L.si) / distance, 0.0f) : 0.0f;
						L.shadowVal = f.softShadows ? 1.0f - L.maxSoft : sh;
						act = A_COMBINE;
						break;
					}
					L.scount++;
					L.si += fmaxf(fminf(d, 1e6f) * f.DEFactor, 1e-15f);
					if (L.scount > f.maxRaymarchingSteps || !(L.si < distance))
					{
						L.shadowVal = f.softShadows ? 1.0f - L.maxSoft : 1.0f;
						act = A_COMBINE;
						break;
					}
					query(L.hp + lv * L.si, L.hthr, A_SHADOW_DE);
					return;
				}
				case A_COMBINE:
				{
					const V3 light = v3(f.lightColor[0], f.lightColor[1], f.lightColor[2]) * (L.shade * L.shadowVal);
					V3 out;
					out.x = (f.fillLight[0] + L.aoAcc.x * f.aoColor[0]) * colour.x + light.x * colour.x + f.luminosity[0];
					out.y = (f.fillLight[1] + L.aoAcc.y * f.aoColor[1]) * colour.y + light.y * colour.y + f.luminosity[1];
					out.z = (f.fillLight[2] + L.aoAcc.z * f.aoColor[2]) * colour.z + light.z * colour.z + f.luminosity[2];
					if (L.ctx == 0)
					{
						L.primaryRGB = out;
						if (kMcOn && f.monteCarlo && f.globalIllumination)
						{
							L.giOut = v3(0.0f, 0.0f, 0.0f);
							L.bounce = 0;
							act = A_GI_BOUNCE;
							break;
						}
						L.sampleRGB = v3(fmaxf(out.x, 0.0f), fmaxf(out.y, 0.0f), fmaxf(out.z, 0.0f));
						act = A_SAMPLE_DONE;
						break;
					}
					if (!kMcOn)
					{
						act = A_SAMPLE_DONE;
						break;
					}
					// GI bounce hit: the ObjectShader result counts once plainly and once clamped
					// (shader_global_illumination.cpp:171-190)
					L.giOut = L.giOut + v3(out.x * colour.x, out.y * colour.y, out.z * colour.z);
					const V3 c = clamp_rgb(out, f.giRadianceLimit);
					L.giOut = L.giOut + v3(c.x * colour.x, c.y * colour.y, c.z * colour.z);
					L.bounce++;
					act = A_GI_BOUNCE;
					break;
				}
				// ---- GlobalIlumination bounce loop (shader_global_illumination.cpp:41-209)
				case A_GI_BOUNCE:
				{
					if (!kMcOn)
					{
						act = A_SAMPLE_DONE;
						break;
					}
					if (L.bounce >= f.reflectionsMax)
					{
						act = A_GI_DONE;
						break;
					}
					const float rx = (float)(random_mod(L.seed, 20000) - 10000) * 1e-4f;
					const float ry = (float)(random_mod(L.seed, 20000) - 10000) * 1e-4f;
					const float rz = (float)(random_mod(L.seed, 20000) - 10000) * 1e-4f;
					L.giDir = normalize(L.hn + v3(rx * 1.2f, ry * 1.2f, rz * 1.2f));
					L.giOrigin = L.hp;
					L.si = L.hthr;
					L.scount = 0;
					act = A_GI_NEXT;
					break;
				}
				case A_GI_NEXT:
				{
					if (!kMcOn)
					{
						act = A_SAMPLE_DONE;
						break;
					}
					if (L.si < f.viewDistanceMax && L.scount < f.maxRaymarchingSteps)
					{
						const V3 q = L.giOrigin + L.giDir * L.si;
						L.sstep = thresh_at(f, q);
						query(q, L.sstep, A_GI_DE);
						return;
					}
					const V3 c = clamp_rgb(background(f, L.giDir), f.giRadianceLimit);
					L.giOut = L.giOut + v3(c.x * colour.x, c.y * colour.y, c.z * colour.z);
					act = A_GI_DONE;
					break;
				}
				case A_GI_DE:
				{
					if (!kMcOn)
					{
						act = A_SAMPLE_DONE;
						break;
					}
					const float thr = L.sstep;
					if (d < thr)
					{
						if (L.si < thr * 2.0f)
						{
							act = A_GI_DONE;
							break;
						}
						L.hp = L.giOrigin + L.giDir * L.si;
						L.hthr = thr;
						L.ctx = 1;
						act = A_NORMALS;
						break;
					}
					L.si += d * f.DEFactor;
					L.scount++;
					act = A_GI_NEXT;
					break;
				}
				case A_GI_DONE:
				{
					if (!kMcOn)
					{
						act = A_SAMPLE_DONE;
						break;
					}
					const V3 o = L.primaryRGB + L.giOut;
					L.sampleRGB = v3(fmaxf(o.x, 0.0f), fmaxf(o.y, 0.0f), fmaxf(o.z, 0.0f));
					act = A_SAMPLE_DONE;
					break;
				}
				case A_SAMPLE_DONE:
				default:
				{
					L.acc = L.acc + L.sampleRGB;
					L.outSteps += L.counter;
					L.rep++;
					if (kMcOn && f.monteCarlo && L.rep < f.samples)
					{
						act = A_SAMPLE;
						break;
					}
					L.phase = -1;
					return;
				}
			}
		}
	}
};

// Frame of global tile t (tiles of all frames are numbered consecutively).
__device__ __forceinline__ int frame_of(const FrameJob *__restrict__ jobs, int numFrames, unsigned int t)
{
	int fr = 0;
	while (fr + 1 < numFrames && t >= (unsigned int)jobs[fr].tileEnd) fr++;
	return fr;
}

// Persistent kernel: every wave loops over tiles taken from the global queue (one atomic per
// 64 pixels).  A wave stays on one frame at a time so that the frame constants remain
// wave-uniform; a tile of the next frame waits until the wave's lanes have drained.
template <int KIND, int FEAT>
__global__ __launch_bounds__(64) void persistent_kernel(const FrameConst *__restrict__ frames,
	const FrameJob *__restrict__ jobs, int numFrames, unsigned int *__restrict__ queue,
	unsigned long long *__restrict__ counters)
{
	const int lane = threadIdx.x;
	const unsigned long long lowerMask = (1ull << lane) - 1ull;
	const unsigned int totalTiles = (unsigned int)jobs[numFrames - 1].tileEnd;
	unsigned long long de_evals = 0, iterations = 0;

	unsigned int next = 0;
	if (lane == 0) next = atomicAdd(queue, 1u);
	// wave-uniform from here on: the frame constants are then scalar loads
	next = __builtin_amdgcn_readfirstlane(__shfl(next, 0, 64));
	while (next < totalTiles)
	{
		const int fr = frame_of(jobs, numFrames, next);
		const FrameConst &f = frames[fr];
		const FrameJob &job = jobs[fr];
		const unsigned int tileBegin = fr == 0 ? 0u : (unsigned int)jobs[fr - 1].tileEnd;
		const unsigned int tileEnd = (unsigned int)job.tileEnd;
		const int tilesX = (f.width + 7) >> 3;
		Machine<KIND, FEAT> m(f);
		unsigned int tile = next;
		int slot = 0;
		bool tileValid = true;
		bool active = false;
		while (true)
		{
			const unsigned long long need = __ballot(!active);
			if (need && tileValid)
			{
				if (slot >= 64)
				{
					unsigned int t = 0;
					if (lane == 0) t = atomicAdd(queue, 1u);
					t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
					if (t < tileEnd)
					{
						tile = t;
						slot = 0;
					}
					else
					{
						next = t; // first tile of a later frame, or the end of the queue
						tileValid = false;
					}
				}
				if (tileValid)
				{
					const int rank = __popcll(need & lowerMask);
					const int s = slot + rank;
					if (!active && s < 64)
					{
						const int lt = (int)(tile - tileBegin);
						const int xs = (lt % tilesX) * 8 + (s & 7);
						const int ys = (lt / tilesX) * 8 + (s >> 3);
						if (xs < f.width && ys < f.height)
						{
							m.begin_pixel(ys * f.width + xs);
							active = true;
						}
					}
					slot = min(slot + __popcll(need), 64);
				}
			}
			if (__ballot(active) == 0ull)
			{
				if (!tileValid) break;
				continue;
			}
			if (active)
			{
				de_step<KIND>(m.de, f);
				if (m.de.status != 0)
				{
					int iters;
					const float d = de_finish<KIND>(m.de, f, m.detail, iters);
					m.w.de_evals++;
					m.w.iterations += (KIND == kBoxDelta ? 4u : 1u) * (uint32_t)iters;
					m.on_de(d, iters);
					if (m.L.phase < 0)
					{
						const size_t i = (size_t)m.L.pix;
						const int reps = ((FEAT & kFeatMC) && f.monteCarlo) ? f.samples : 1;
						const V3 rgb = reps > 1 ? m.L.acc * (1.0f / (float)reps) : m.L.acc;
						if (job.rgb)
						{
							job.rgb[3 * i + 0] = rgb.x;
							job.rgb[3 * i + 1] = rgb.y;
							job.rgb[3 * i + 2] = rgb.z;
						}
						if (job.depth) job.depth[i] = m.L.outDepth;
						if (job.iters) job.iters[i] = m.L.outIters;
						if (job.mask) job.mask[i] = (uint8_t)m.L.outFound;
						if (job.normal)
						{
							job.normal[3 * i + 0] = m.L.outNormal.x;
							job.normal[3 * i + 1] = m.L.outNormal.y;
							job.normal[3 * i + 2] = m.L.outNormal.z;
						}
						if (job.steps) job.steps[i] = m.L.outSteps;
						active = false;
					}
				}
			}
		}
		de_evals += m.w.de_evals;
		iterations += m.w.iterations;
	}
	if (counters)
	{
#pragma unroll
		for (int off = 32; off > 0; off >>= 1)
		{
			de_evals += __shfl_xor(de_evals, off, 64);
			iterations += __shfl_xor(iterations, off, 64);
		}
		if (lane == 0)
		{
			atomicAdd(&counters[0], de_evals);
			atomicAdd(&counters[1], iterations);
		}
	}
}

template <int KIND, int FEAT>
static hipError_t launch_persistent_t(const FrameConst *frames, const FrameJob *jobs, int numFrames,
	unsigned int *queue, unsigned long long *counters, int numWaves, hipStream_t stream)
{
	hipLaunchKernelGGL((persistent_kernel<KIND, FEAT>), dim3(numWaves), dim3(64), 0, stream, frames,
		jobs, numFrames, queue, counters);
	return hipGetLastError();
}

template <int KIND>
static hipError_t launch_persistent_k(int feat, const FrameConst *frames, const FrameJob *jobs,
	int numFrames, unsigned int *queue, unsigned long long *counters, int numWaves, hipStream_t stream)
{
	if (feat & kFeatMC)
		return launch_persistent_t<KIND, kFeatShadow | kFeatAO | kFeatMC>(frames, jobs, numFrames, queue,
			counters, numWaves, stream);
	if (feat & (kFeatShadow | kFeatAO))
		return launch_persistent_t<KIND, kFeatShadow | kFeatAO>(frames, jobs, numFrames, queue, counters,
			numWaves, stream);
	return launch_persistent_t<KIND, 0>(frames, jobs, numFrames, queue, counters, numWaves, stream);
}

hipError_t launch_persistent(int deKind, int feat, const FrameConst *frames, const FrameJob *jobs,
	int numFrames, unsigned int *queue, unsigned long long *counters, int numWaves, hipStream_t stream)
{
	switch (deKind)
	{
		case kBulb8:
			return launch_persistent_k<kBulb8>(feat, frames, jobs, numFrames, queue, counters, numWaves, stream);
		case kBulbTrig:
			return launch_persistent_k<kBulbTrig>(feat, frames, jobs, numFrames, queue, counters, numWaves, stream);
		case kBoxAnalytic:
			return launch_persistent_k<kBoxAnalytic>(feat, frames, jobs, numFrames, queue, counters, numWaves,
				stream);
		case kBoxDelta:
			return launch_persistent_k<kBoxDelta>(feat, frames, jobs, numFrames, queue, counters, numWaves, stream);
		case kIfsAnalytic:
			return launch_persistent_k<kIfsAnalytic>(feat, frames, jobs, numFrames, queue, counters, numWaves,
				stream);
		default: return hipErrorInvalidValue;
	}
}

// Small device-side clear of the queue head and counters (graph-capturable, no memset node).
__global__ void clear_kernel(unsigned int *queue, unsigned long long *counters)
{
	if (threadIdx.x == 0) queue[0] = 0u;
	if (threadIdx.x < 4) counters[threadIdx.x] = 0ull;
}

hipError_t launch_clear(unsigned int *queue, unsigned long long *counters, hipStream_t stream)
{
	hipLaunchKernelGGL(clear_kernel, dim3(1), dim3(64), 0, stream, queue, counters);
	return hipGetLastError();
}

} // namespace mbx

"""Scene parameter blocks: reference defaults and the benchmark / parity workloads.

Defaults are the values cParameterContainer registers in mandelbulber2/src/initparameters.cpp
(line numbers cited per field); colours given there as 16-bit sRGB become float(c / 65536) as
toRGBFloat makes them (color_structures.hpp:106-109).  Only parameters on the raymarch path are
represented.  The library holds the same table (mbx_params_default); a test keeps them equal.
"""
import ctypes
import math
import struct

import numpy as np

from . import abi


def _u16(rgb16):
    return [float(np.float32(c / 65536.0)) for c in rgb16]


def default_fractal(f, formula):
    """One fractal container's defaults (InitFractalParams, initparameters.cpp:695-1010) with the
    sequence-wide values (:128-156) and node defaults (:2102-2125)."""
    f.formula = formula
    f.formulaMaxiter = 250
    f.deltaDEMethod = 0
    f.deltaDEFunction = 0
    f.useDefaultBailout = 1
    f.juliaMode = 0
    f.dontAddCConstant = 0
    f.checkForBailout = 1                      # :714
    f.bailout = 100.0
    abi.set3(f.juliaC, (0.0, 0.0, 0.0))
    abi.set3(f.constantFactor, (1.0, 1.0, 1.0))
    f.initialWAxis = 0.0
    f.formulaIterations = 1                    # :701-704
    f.formulaWeight = 1.0
    f.formulaStartIteration = 0
    f.formulaStopIteration = 250
    f.bulb.power = 9.0
    f.bulb.alphaAngleOffset = 0.0
    f.bulb.betaAngleOffset = 0.0
    mb = f.mandelbox
    mb.scale = 2.0
    mb.foldingLimit = 1.0
    mb.foldingValue = 2.0
    mb.foldingSphericalMin = 0.5
    mb.foldingSphericalFixed = 1.0
    abi.set3(mb.colorFactor, (0.03, 0.05, 0.07))
    mb.colorFactorSp1 = 0.2
    mb.colorFactorSp2 = 0.2
    mb.colorFactorR = 0.0                      # :764
    ifs = f.IFS
    ifs.scale = 2.0
    abi.set3(ifs.offset, (1.0, 0.0, 0.0))
    for i in range(abi.MBX_IFS_VECTOR_COUNT):
        abi.set3(ifs.direction[i], (1.0, 0.0, 0.0))
        ifs.distance[i] = 0.0
        ifs.intensity[i] = 1.0
        ifs.enabled[i] = 0
    t = f.transform
    t.scale3 = 3.0                             # :1010
    t.aexionCadd = -1.3                        # :722
    t.foldingIntPowFoldFactor = 2.0            # :784
    t.foldingIntPowZFactor = 5.0               # :785
    t.analyticDEEnabled = 1                    # :904-910
    t.analyticDEEnabledFalse = 0
    t.analyticDEScale1 = 1.0
    t.analyticDEOffset1 = 1.0
    t.analyticDEOffset2 = 1.0
    return f


def default_params(width=640, height=480):
    """Reference defaults (initparameters.cpp) for the fields of mbx_params."""
    p = abi.MbxParams()
    p.imageWidth = width
    p.imageHeight = height
    p.perspectiveType = 0                      # three-point (initparameters.cpp:121)
    p.maxRaymarchingSteps = 10000              # :167
    p.minN = 1                                 # :131
    p.iterThreshMode = 0                       # :137
    p.interiorMode = 0                         # :147
    p.constantDEThreshold = 0                  # :148
    abi.set3(p.camera, (3.0, -6.0, 2.0))       # :113
    abi.set3(p.target, (0.0, 0.0, 0.0))        # :114
    abi.set3(p.topVector, (-0.1277753, 0.2555506, 0.958314))  # :116
    p.fov = 53.13                              # :120
    p.sweetSpotHAngle = 0.0                    # :124
    p.sweetSpotVAngle = 0.0                    # :125
    p.detailLevel = 1.0                        # :133
    p.DEThresh = 0.01                          # :135
    p.DEFactor = 1.0                           # :139
    p.smoothness = 1.0                         # :136
    p.viewDistanceMax = 50.0                   # :141
    # material #1 (initparameters.cpp:1640-1781)
    abi.set3(p.surfaceColor, _u16((50000, 50000, 50000)))
    p.shading = 1.0
    p.luminosity = 0.0
    abi.set3(p.luminosityColor, _u16((65535, 65535, 65535)))
    abi.set3(p.fillLightColor, (0.0, 0.0, 0.0))  # :309
    # light #1 (initparameters.cpp:1841-1960)
    L = p.light
    L.enabled = 1
    L.castShadows = 1
    L.penetrating = 1
    L.relativePosition = 1
    abi.set3(L.rotation, (-45.0, 45.0, 0.0))
    L.intensity = 1.0
    L.visibility = 1.0
    L.size = 0.5
    L.softShadowCone = 1.0
    L.contourSharpness = 1.0
    abi.set3(L.color, _u16((65535, 65535, 65535)))
    # ambient occlusion (initparameters.cpp:213-219)
    p.ambientOcclusionEnabled = 0
    p.ambientOcclusionMode = abi.MBX_AO_SCREEN_SPACE
    p.ambientOcclusionQuality = 4
    p.aoVectorsCount = 0
    p.ambientOcclusion = 1.0
    abi.set3(p.ambientOcclusionColor, _u16((65535, 65535, 65535)))
    p.ambientOcclusionFastTune = 1.0
    # background (initparameters.cpp:227-245)
    p.background3ColorsEnable = 1
    abi.set3(p.backgroundColor1, _u16((0, 38306, 65535)))
    abi.set3(p.backgroundColor2, _u16((65535, 65535, 65535)))
    abi.set3(p.backgroundColor3, _u16((0, 10000, 500)))
    p.backgroundBrightness = 1.0
    p.backgroundGamma = 1.0
    # Monte Carlo (initparameters.cpp:220-231, 128)
    p.DOFMonteCarlo = 0
    p.DOFSamples = 100
    p.DOFMonteCarloGlobalIllumination = 0
    p.reflectionsMax = 5
    p.monteCarloGIRadianceLimit = 10.0
    p.rngMode = abi.MBX_RNG_PER_PIXEL
    p.randomSeed = 0
    # fractal containers: #1 Mandelbulb, the others none; a single formula (initparameters.cpp:149-157)
    default_fractal(p.fractal, abi.MBX_FORMULA_MANDELBULB)
    for k in range(abi.MBX_MAX_FRACTALS - 1):
        default_fractal(p.hybridFractals[k], abi.MBX_FORMULA_NONE)
    p.hybridCount = 0
    p.repeatFrom = 1
    p.linearDEOffset = 0.0
    # advanced quality march limits (initparameters.cpp:159-166)
    p.advancedQuality = 0
    p.absMaxMarchingStep = 3.0
    p.absMinMarchingStep = 1e-15
    p.relMaxMarchingStep = 1e4
    p.relMinMarchingStep = 1e-2
    p.detailSizeMax = 1.0
    p.detailSizeMin = 1e-12
    p.deltaDERelativeDelta = 0.01
    p.image.brightness = p.image.contrast = p.image.gamma = p.image.saturation = 1.0
    p.image.hdr = 0
    # global foldings (initparameters.cpp:199-204) and slow shading (:140)
    p.boxFolding = 0
    p.boxFoldingLimit = 1.0
    p.boxFoldingValue = 2.0
    p.sphericalFolding = 0
    p.sphericalFoldingOuter = 1.0
    p.sphericalFoldingInner = 0.5
    p.slowShading = 0
    p.N = 250
    p.DOFMinSamples = 10
    p.DOFMaxNoise = 1.0
    p.aoLightMapRotation = 0.0
    p.DOFEnabled = 0
    p.DOFMonteCarloChromaticAberration = 0
    p.DOFFocus = 6.0
    p.DOFRadius = 10.0
    p.DOFMonteCarloCACameraDispersion = 1.0
    p.monteCarloSoftShadows = 0
    # palette colouring of material #1 (InitMaterialParams, initparameters.cpp:1600-1810)
    p.useColorsFromPalette = 1
    p.surfaceGradientEnable = 1
    p.coloringAlgorithm = 0
    p.coloringSpeed = 1.0
    p.coloringPaletteOffset = 0.0
    p.coloringSphereRadius = 1.0
    abi.set3(p.coloringLineDirection, (1.0, 0.0, 0.0))
    set_gradient(p, DEFAULT_GRADIENT)
    # specular of material #1
    p.specularPlasticEnable = 1
    p.metallic = 1
    p.specular = 5.0
    p.specularWidth = 0.05
    p.specularMetallic = 2.0
    p.specularMetallicWidth = 1.0
    p.specularMetallicRoughness = 0.1
    abi.set3(p.specularColor, _u16((65535, 65535, 65535)))
    p.luminosityEmissive = 1.0
    # glow (initparameters.cpp:222-223, 240-241)
    p.glowEnabled = 1
    p.glowIntensity = 0.2
    abi.set3(p.glowColor1, _u16((40984, 44713, 49490)))
    abi.set3(p.glowColor2, _u16((57192, 60888, 62408)))
    # screen-space AO (initparameters.cpp:221), fog (:247-272, 315-319), volumetric light DE
    # factor (:350)
    p.SSAORandomMode = 0
    p.basicFogEnabled = 0
    p.basicFogCastShadows = 0
    p.basicFogVisibility = 20.0
    abi.set3(p.basicFogColor, _u16((59399, 61202, 65535)))
    p.volFogEnabled = 0
    p.distanceFogShadows = 0
    p.volFogDensity = 0.5
    p.volFogColour1Distance = 1.0
    p.volFogColour2Distance = 2.0
    p.volFogDistanceFactor = 1.0
    p.volFogDistanceFromSurface = 1e-15
    for k, c in enumerate(((30000, 30000, 30000), (0, 30000, 65535), (65535, 65535, 65535))):
        abi.set3(p.volFogColour[k], _u16(c))
    p.iterFogEnabled = 0
    p.cloudsEnabled = 0
    p.volumetricLightDEFactor = 1.0
    p.antialiasingEnabled = 0
    p.antialiasingSize = 2
    return p


DEFAULT_GRADIENT = [(0, 0xfd6029), (1000, 0x698403), (2000, 0xfff59b), (3000, 0xf5bd22),
                    (4000, 0x0b5e87), (5000, 0xc68876), (6000, 0xa51c64), (7000, 0x3b9fee),
                    (8000, 0xd4ffd4), (9000, 0xaba53c)]


def set_gradient(p, stops):
    """Surface gradient from (position 0..10000, 0xRRGGBB) pairs as cColorGradient holds it
    (color_gradient.cpp:305-354): positions / 10000 as float, the first colour again at 1.0,
    sorted by position."""
    pts = [(float(np.float32(pos) / np.float32(10000.0)), rgb) for pos, rgb in stops]
    pts.insert(1, (1.0, pts[0][1]))
    pts.sort(key=lambda s: s[0])
    p.surfaceGradientCount = len(pts)
    for i in range(abi.MBX_MAX_GRADIENT_STOPS):
        pos, rgb = pts[i] if i < len(pts) else (0.0, 0)
        p.surfaceGradientPosition[i] = pos
        abi.set3(p.surfaceGradientColor[i], ((rgb >> 16) & 255, (rgb >> 8) & 255, rgb & 255))


def plain_material(p):
    """The benchmark scenes' material: the material colour (palette off), Lambert light without
    specular highlights, no glow."""
    p.useColorsFromPalette = 0
    p.specular = 0.0
    p.specularMetallic = 0.0
    p.specularPlasticEnable = 0
    p.metallic = 0
    p.glowEnabled = 0
    return p


def reference_material(p):
    """Material #1 and glow as the reference initialises them (palette colouring from the default
    surface gradient, plastic and metallic specular highlights, glow on), restored onto a scene
    made with plain_material: the fields from useColorsFromPalette to glowColor2."""
    d = default_params(p.imageWidth, p.imageHeight)
    off = abi.MbxParams.useColorsFromPalette.offset
    end = abi.MbxParams.glowColor2.offset + abi.MbxParams.glowColor2.size
    ctypes.memmove(ctypes.addressof(p) + off, ctypes.addressof(d) + off, end - off)
    return p


# ------------------------------------------------------------------------------------------
# workloads
# ------------------------------------------------------------------------------------------

def mandelbulb8(width=640, height=480, seed=0):
    """Plain CPU-runnable case: Mandelbulb power 8, analytic log DE, bailout 10, 250 iterations,
    DE factor 1, detail 1, camera at distance 3.5 looking at the origin, FOV 53 degrees, one
    directional light without shadows, Lambert shading, no AO/GI/glow, 1 sample per pixel.
    The 1920x1080 variant is the headline throughput workload."""
    p = plain_material(default_params(width, height))
    abi.set3(p.camera, (0.0, -3.5, 0.0))
    abi.set3(p.target, (0.0, 0.0, 0.0))
    abi.set3(p.topVector, (0.0, 0.0, 1.0))
    p.fov = 53.0
    p.light.castShadows = 0
    p.randomSeed = seed
    f = p.fractal
    f.formula = abi.MBX_FORMULA_MANDELBULB
    f.bulb.power = 8.0
    f.formulaMaxiter = 250
    return p


def mandelbulb8_orbit(width=1920, height=1080, frames=8, step_deg=3.0):
    """Frames of the Mandelbulb-8 scene along a camera orbit: frame k has the camera at distance
    3.5 in the z = 0 plane, k * step_deg degrees round the z axis from (0, -3.5, 0), looking at
    the origin (an animation's consecutive frames)."""
    out = []
    for k in range(frames):
        p = mandelbulb8(width, height, seed=k)
        a = math.radians(step_deg * k)
        abi.set3(p.camera, (3.5 * math.sin(a), -3.5 * math.cos(a), 0.0))
        out.append(p)
    return out


def splitmix64(state):
    """SplitMix64 step: returns (value, new_state)."""
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31), state


class SeededStream:
    """Deterministic uniforms / normals from SplitMix64 (Box-Muller for normals)."""

    def __init__(self, seed):
        self.state = seed & 0xFFFFFFFFFFFFFFFF

    def uniform(self, a=0.0, b=1.0):
        v, self.state = splitmix64(self.state)
        return a + (b - a) * ((v >> 11) * (1.0 / 9007199254740992.0))

    def normal(self):
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def fibonacci_hemisphere(n):
    """n unit directions on the upper (z >= 0) hemisphere by the Fibonacci spiral."""
    golden = math.pi * (3.0 - math.sqrt(5.0))
    out = []
    for i in range(n):
        z = 1.0 - (i + 0.5) / n
        r = math.sqrt(max(0.0, 1.0 - z * z))
        phi = golden * i
        out.append((r * math.cos(phi), r * math.sin(phi), z))
    return out


def mandelbox_shadow_ao(width=1920, height=1080, seed=1, scale=2.0):
    """Mandelbox (folding limit 1, min radius 0.5, fixed radius 1), delta DE with the linear DE
    function, 50 iterations; hard shadows from the directional light and 16-ray AO with a fixed
    Fibonacci-hemisphere direction set; camera on a seeded sphere of radius 6-10 looking at the
    origin.  scale 2.0 for the first run, -1.5 for the second."""
    p = plain_material(default_params(width, height))
    rs = SeededStream(seed)
    d = [rs.normal() for _ in range(3)]
    n = math.sqrt(sum(c * c for c in d))
    radius = rs.uniform(6.0, 10.0)
    abi.set3(p.camera, [c / n * radius for c in d])
    abi.set3(p.target, (0.0, 0.0, 0.0))
    abi.set3(p.topVector, (0.0, 0.0, 1.0))
    p.light.castShadows = 1
    p.light.softShadowCone = 0.0
    p.light.penetrating = 1
    p.ambientOcclusionEnabled = 1
    p.ambientOcclusionMode = abi.MBX_AO_MULTIPLE_RAYS
    dirs = fibonacci_hemisphere(16)
    p.aoVectorsCount = len(dirs)
    for i, v in enumerate(dirs):
        abi.set3(p.aoVectors[i], v)
        abi.set3(p.aoColors[i], (1.0, 1.0, 1.0))
    p.randomSeed = seed
    f = p.fractal
    f.formula = abi.MBX_FORMULA_MANDELBOX
    f.formulaMaxiter = 50
    f.deltaDEMethod = 1
    f.deltaDEFunction = 1
    f.mandelbox.scale = scale
    f.mandelbox.foldingLimit = 1.0
    f.mandelbox.foldingValue = 2.0
    f.mandelbox.foldingSphericalMin = 0.5
    f.mandelbox.foldingSphericalFixed = 1.0
    return p


def menger_ifs(width=320, height=240, seed=2):
    """Kaleidoscopic IFS in Menger-sponge mode (abs folds, scale 3, offset (1,1,1)) as in
    mandelbulber2/deploy/share/mandelbulber2/examples/IFS_001.fract:38-49."""
    p = plain_material(default_params(width, height))
    abi.set3(p.camera, (2.2, -2.6, 1.7))
    abi.set3(p.target, (0.0, 0.0, 0.0))
    abi.set3(p.topVector, (0.0, 0.0, 1.0))
    p.light.castShadows = 0
    p.randomSeed = seed
    f = p.fractal
    f.formula = abi.MBX_FORMULA_KALEIDOSCOPIC_IFS
    f.formulaMaxiter = 30
    ifs = f.IFS
    ifs.scale = 3.0
    abi.set3(ifs.offset, (1.0, 1.0, 1.0))
    ifs.absX = ifs.absY = ifs.absZ = 1
    ifs.mengerSpongeMode = 1
    return p


ANIMATION_RESOLUTIONS = ((1920, 1080), (1280, 720), (640, 480))


def animation64(seed=3, frames=64, downscale=1):
    """Batched animation: `frames` Mandelbulb frames with power_k = 2 + 10 k / (frames - 1)
    (general-power trigonometric form), the camera on a seeded random walk over the shell of
    radius 2.5-4 around the origin, each frame's resolution drawn from 1920x1080, 1280x720 and
    640x480 by the seed (divided by `downscale` for small parity runs), 4 Monte Carlo samples
    per pixel with global illumination of depth 2.  Returns the list of parameter blocks."""
    rs = SeededStream(seed)
    d = [rs.normal() for _ in range(3)]
    radius = rs.uniform(2.5, 4.0)
    out = []
    for k in range(frames):
        n = math.sqrt(sum(c * c for c in d))
        d = [c / n for c in d]
        w, h = ANIMATION_RESOLUTIONS[int(rs.uniform(0.0, 3.0)) % 3]
        p = mandelbulb8(max(8, w // downscale), max(8, h // downscale), seed=seed * 1000 + k)
        abi.set3(p.camera, [c * radius for c in d])
        p.fractal.bulb.power = 2.0 + 10.0 * k / max(1, frames - 1)
        p.DOFMonteCarlo = 1
        p.DOFSamples = 4
        p.DOFMonteCarloGlobalIllumination = 1
        p.reflectionsMax = 2
        out.append(p)
        # random walk: perturb the direction, then the radius, re-projected onto the shell
        d = [c + 0.06 * rs.normal() for c in d]
        radius = min(4.0, max(2.5, radius + 0.05 * rs.normal()))
    return out


def _mandelbox_de(p, scale=2.0, L=1.0, V=2.0, rmin=0.5, rfix=1.0, bailout=100.0, maxiter=50):
    """Analytic linear DE r / DE of the Mandelbox orbit of p (c = p), FP64: the same formula and
    orbit as cFractalMandelbox::FormulaCode (fractal_mandelbox.cpp:31-107) with Compute's
    bailout (compute_fractal.cpp:299-321); used here only to place the deep-zoom camera."""
    mR2, fR2 = rmin * rmin, rfix * rfix
    factor1 = fR2 / mR2
    x, y, z = p
    DE = 1.0
    r = math.sqrt(x * x + y * y + z * z)
    for _ in range(maxiter):
        if abs(x) > L:
            x = math.copysign(V, x) - x
        if abs(y) > L:
            y = math.copysign(V, y) - y
        if abs(z) > L:
            z = math.copysign(V, z) - z
        r2 = x * x + y * y + z * z
        k = factor1 if r2 < mR2 else (fR2 / r2 if r2 < fR2 else 1.0)
        x, y, z = x * k * scale + p[0], y * k * scale + p[1], z * k * scale + p[2]
        DE = DE * k * abs(scale) + 1.0
        r = math.sqrt(x * x + y * y + z * z)
        if r > bailout:
            break
    return r / abs(DE)


def deep_zoom_mandelbox(width=1024, height=1024, seed=5, surface_eps=1e-11, standoff=1e-9):
    """Deep zoom: Mandelbox scale 2 (analytic linear DE, 50 iterations).  From a seeded exterior
    point at radius 8 the FP64 DE is sphere-traced toward the origin until the distance is
    below surface_eps; the camera sits `standoff` back along that ray and looks at the surface
    point.  Constant DE threshold 1e-11, 1024x1024."""
    rs = SeededStream(seed)
    d = [rs.normal() for _ in range(3)]
    n = math.sqrt(sum(c * c for c in d))
    start = [8.0 * c / n for c in d]
    aim = [0.25 * rs.normal() for _ in range(3)]
    ray = [a - s0 for a, s0 in zip(aim, start)]
    n = math.sqrt(sum(c * c for c in ray))
    ray = [c / n for c in ray]
    t = 0.0
    for _ in range(200000):
        q = [s0 + c * t for s0, c in zip(start, ray)]
        dist = _mandelbox_de(q)
        if dist < surface_eps:
            break
        t += dist
    else:
        raise RuntimeError("deep-zoom surface search did not converge")
    surface = [s0 + c * t for s0, c in zip(start, ray)]
    p = plain_material(default_params(width, height))
    abi.set3(p.camera, [sf - c * standoff for sf, c in zip(surface, ray)])
    abi.set3(p.target, surface)
    abi.set3(p.topVector, (0.0, 0.0, 1.0))
    p.fov = 53.0
    p.constantDEThreshold = 1
    p.DEThresh = surface_eps
    p.light.castShadows = 0
    p.randomSeed = seed
    f = p.fractal
    f.formula = abi.MBX_FORMULA_MANDELBOX
    f.formulaMaxiter = 50
    f.deltaDEMethod = 0  # the formula's default: analytic
    f.deltaDEFunction = 0
    f.mandelbox.scale = 2.0
    f.mandelbox.foldingLimit = 1.0
    f.mandelbox.foldingValue = 2.0
    f.mandelbox.foldingSphericalMin = 0.5
    f.mandelbox.foldingSphericalFixed = 1.0
    return p


def inverse_rendering(width=512, height=512, seed=0):
    """Differentiable-render scene: Mandelbulb (trigonometric form, analytic log DE) with a
    seeded power near 8 and node rotation, camera at distance ~3 looking at the origin, one
    light at an absolute angle, Lambert, no shadows/AO.  The ten differentiable parameters are
    power, rotation x/y/z (degrees), camera x/y/z and target x/y/z."""
    rs = SeededStream(1000 + seed)
    p = mandelbulb8(width, height, seed=seed)
    p.fractal.bulb.power = 8.0 + rs.uniform(-1.0, 1.0)
    abi.set3(p.formulaRotation, [rs.uniform(-30.0, 30.0) for _ in range(3)])
    d = [rs.normal() for _ in range(3)]
    n = math.sqrt(sum(c * c for c in d))
    radius = rs.uniform(2.6, 3.4)
    abi.set3(p.camera, [c / n * radius for c in d])
    abi.set3(p.target, [rs.uniform(-0.1, 0.1) for _ in range(3)])
    abi.set3(p.topVector, (0.0, 0.0, 1.0))
    p.light.relativePosition = 0
    p.light.castShadows = 0
    return p


def params_bytes(p):
    return bytes(memoryview(p))


def copy_params(p):
    q = abi.MbxParams()
    struct_bytes = params_bytes(p)
    abi.C.memmove(abi.C.addressof(q), struct_bytes, len(struct_bytes))
    return q

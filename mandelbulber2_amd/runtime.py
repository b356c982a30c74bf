"""Python binding of libmbx.so (the C ABI in include/mbx.h).

The library is the product: HIP kernels for gfx950 plus the C++ host.  This module only
loads it with ctypes and moves torch device buffers across the boundary; it has no compute
path of its own and raises if the library is missing.
"""
import ctypes as C
import os

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmbx.so")

# every entry point include/mbx.h declares
EXPORTS = ["mbx_version", "mbx_sizeof_params", "mbx_ctx_create", "mbx_ctx_destroy",
           "mbx_last_error", "mbx_scene_upload", "mbx_scene_info", "mbx_render", "mbx_distance",
           "mbx_get_stats", "mbx_render_batch", "mbx_tune", "mbx_get_tune", "mbx_autotune",
           "mbx_render_diff", "mbx_adjoint", "mbx_params_default", "mbx_load_fract",
           "mbx_compile_image", "mbx_write_png", "mbx_compile_zbuffer", "mbx_render_batch_noalloc",
           "mbx_formula_count", "mbx_formula_get", "mbx_formula_iterate", "mbx_prepare_ao_vectors",
           "mbx_render_ssao", "mbx_write_tiff", "mbx_write_exr", "mbx_float_to_half"]

_lib = None


class MbxError(RuntimeError):
    pass


def load():
    """Load the in-tree libmbx.so (built by mandelbulber2_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MbxError("libmbx.so is not built: run `python -m mandelbulber2_amd.build` "
                       "(there is no fallback path)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.mbx_version.restype = C.c_int
    L.mbx_sizeof_params.restype = C.c_size_t
    L.mbx_ctx_create.argtypes = [C.c_int, C.c_void_p, P(C.c_void_p)]
    L.mbx_ctx_destroy.argtypes = [C.c_void_p]
    L.mbx_ctx_destroy.restype = None
    L.mbx_last_error.argtypes = [C.c_void_p]
    L.mbx_last_error.restype = C.c_char_p
    L.mbx_scene_upload.argtypes = [C.c_void_p, P(abi.MbxParams)]
    L.mbx_scene_info.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.mbx_render.argtypes = [C.c_void_p, P(abi.MbxOutputs)]
    L.mbx_distance.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p]
    L.mbx_get_stats.argtypes = [C.c_void_p, P(abi.MbxStats)]
    L.mbx_render_batch.argtypes = [C.c_void_p, P(abi.MbxParams), P(abi.MbxOutputs), C.c_int]
    L.mbx_render_batch_noalloc.argtypes = [C.c_void_p, P(abi.MbxParams), P(abi.MbxOutputs), C.c_int]
    L.mbx_tune.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.mbx_get_tune.argtypes = [C.c_void_p, C.c_int, P(C.c_int)]
    L.mbx_render_diff.argtypes = [C.c_void_p, P(abi.MbxOutputs), C.c_void_p]
    L.mbx_adjoint.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.mbx_autotune.argtypes = [C.c_void_p, P(abi.MbxOutputs), C.c_int, P(C.c_int), P(C.c_double)]
    L.mbx_params_default.argtypes = [P(abi.MbxParams)]
    L.mbx_params_default.restype = None
    L.mbx_load_fract.argtypes = [C.c_char_p, C.c_size_t, P(abi.MbxParams), P(abi.MbxFractInfo)]
    L.mbx_compile_image.argtypes = [C.c_void_p, P(abi.MbxImageAdjustments), C.c_void_p, C.c_int32,
                                    C.c_int32, C.c_void_p, C.c_void_p]
    L.mbx_compile_zbuffer.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_void_p]
    L.mbx_write_png.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
    L.mbx_write_tiff.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
    L.mbx_write_exr.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, P(C.c_char_p), P(C.c_int32),
                                P(C.c_void_p), P(C.c_int32), C.c_int32]
    L.mbx_float_to_half.argtypes = [C.c_float]
    L.mbx_float_to_half.restype = C.c_uint16
    L.mbx_render_ssao.argtypes = [C.c_void_p, P(abi.MbxParams), C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p]
    L.mbx_prepare_ao_vectors.argtypes = [P(abi.MbxParams), C.c_void_p, C.c_int, C.c_int]
    L.mbx_formula_count.argtypes = []
    L.mbx_formula_count.restype = C.c_int
    L.mbx_formula_get.argtypes = [C.c_int, P(abi.MbxFormulaInfo)]
    L.mbx_formula_iterate.argtypes = [C.c_void_p, P(abi.MbxFractal), C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
    _lib = L
    return L


def params_default():
    """The reference's parameter defaults as the library holds them (mbx_params_default)."""
    p = abi.MbxParams()
    load().mbx_params_default(C.byref(p))
    return p


def load_fract(text):
    """Read a .fract settings text (str or bytes) with mbx_load_fract.  Returns (params, info);
    raises MbxError with the reader's message when the file is rejected."""
    if isinstance(text, str):
        text = text.encode("utf-8")
    p = abi.MbxParams()
    info = abi.MbxFractInfo()
    rc = load().mbx_load_fract(text, len(text), C.byref(p), C.byref(info))
    if rc != 0:
        raise MbxError("mbx_load_fract: %s (%d)" % (info.message.decode(errors="replace"), rc))
    return p, info


def prepare_ao_vectors(params, lightmap=None):
    """Multi-ray AO directions and colours from a decoded lightmap ((H, W, 3) floats, the 8-bit
    value / 256 per channel), or from the white substitute with None (mbx_prepare_ao_vectors);
    fills params in place and returns the direction count."""
    L = load()
    if lightmap is None:
        rc = L.mbx_prepare_ao_vectors(C.byref(params), None, 0, 0)
    else:
        import numpy as np
        t = np.ascontiguousarray(lightmap, np.float32)
        if t.ndim != 3 or t.shape[2] != 3:
            raise MbxError("prepare_ao_vectors: (H, W, 3) lightmap")
        rc = L.mbx_prepare_ao_vectors(C.byref(params), t.ctypes.data_as(C.c_void_p), t.shape[1], t.shape[0])
    if rc != 0:
        raise MbxError("mbx_prepare_ao_vectors failed (%d)" % rc)
    return params.aoVectorsCount


def formulas():
    """The device's formula table with its metadata (mbx_formula_count / mbx_formula_get): a list
    of abi.MbxFormulaInfo."""
    L = load()
    out = []
    for k in range(L.mbx_formula_count()):
        info = abi.MbxFormulaInfo()
        if L.mbx_formula_get(k, C.byref(info)) != 0:
            raise MbxError("mbx_formula_get(%d) failed" % k)
        out.append(info)
    return out


def write_png(path, pixels):
    """ImageFileSavePNG::SavePNG through mbx_write_png: `pixels` is a C-contiguous numpy array of
    shape (H, W, 3) or (H, W), dtype uint8 or uint16."""
    import numpy as np
    a = np.ascontiguousarray(pixels)
    if a.dtype not in (np.uint8, np.uint16) or a.ndim not in (2, 3) or (a.ndim == 3 and a.shape[2] != 3):
        raise MbxError("write_png: (H, W, 3) or (H, W) uint8 / uint16 pixels")
    bits = 8 if a.dtype == np.uint8 else 16
    channels = 3 if a.ndim == 3 else 1
    rc = load().mbx_write_png(str(path).encode(), a.shape[1], a.shape[0], bits, channels,
                              a.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise MbxError("mbx_write_png failed (%d)" % rc)


def write_tiff(path, pixels):
    """ImageFileSaveTIFF::SaveTIFF through mbx_write_tiff: `pixels` is a C-contiguous numpy array
    of shape (H, W), (H, W, 3) or (H, W, 4), dtype uint8, uint16 or float32."""
    import numpy as np
    a = np.ascontiguousarray(pixels)
    bits = {np.dtype(np.uint8): 8, np.dtype(np.uint16): 16, np.dtype(np.float32): 32}.get(a.dtype)
    if bits is None or a.ndim not in (2, 3) or (a.ndim == 3 and a.shape[2] not in (3, 4)):
        raise MbxError("write_tiff: (H, W), (H, W, 3) or (H, W, 4) uint8 / uint16 / float32 pixels")
    channels = 1 if a.ndim == 2 else a.shape[2]
    rc = load().mbx_write_tiff(str(path).encode(), a.shape[1], a.shape[0], bits, channels,
                               a.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise MbxError("mbx_write_tiff failed (%d)" % rc)


def write_exr(path, channels, half=True, linear=False):
    """ImageFileSaveEXR::SaveEXR through mbx_write_exr: `channels` maps channel names to (H, W)
    float arrays, or a name tuple such as ("R", "G", "B") to an (H, W, 3) array; HALF channels
    unless half is False."""
    import numpy as np
    names, planes, strides, keep = [], [], [], []
    H = W = None
    for key, arr in channels.items():
        a = np.ascontiguousarray(arr, np.float32)
        keep.append(a)
        keys = (key,) if isinstance(key, str) else tuple(key)
        if a.ndim != (2 if len(keys) == 1 else 3) or (a.ndim == 3 and a.shape[2] != len(keys)):
            raise MbxError("write_exr: channel %s has shape %s" % (key, a.shape))
        if H is None:
            H, W = a.shape[0], a.shape[1]
        if (a.shape[0], a.shape[1]) != (H, W):
            raise MbxError("write_exr: channels of different sizes")
        for k, name in enumerate(keys):
            names.append(name.encode())
            planes.append(a.ctypes.data + 4 * k)
            strides.append(len(keys))
    n = len(names)
    c_names = (C.c_char_p * n)(*names)
    c_types = (C.c_int32 * n)(*([1 if half else 2] * n))
    c_planes = (C.c_void_p * n)(*planes)
    c_strides = (C.c_int32 * n)(*strides)
    rc = load().mbx_write_exr(str(path).encode(), W, H, n, c_names, c_types, c_planes, c_strides, int(linear))
    if rc != 0:
        raise MbxError("mbx_write_exr failed (%d)" % rc)


MBX_TUNE_KERNEL = 0
MBX_TUNE_WAVES = 1
MBX_TUNE_GRAPH = 2
MBX_TUNE_DELTA_DE = 3
MBX_TUNE_TILE_WIDTH = 4
MBX_TUNE_FAR_FIELD = 5
MBX_TUNE_PRECISION = 6
MBX_TUNE_TIMING = 7
MBX_TUNE_SPLIT = 8
KERNEL_GRID = 0
KERNEL_PLAIN = 1
KERNEL_PERSISTENT = 2


class Context:
    """One mbx context on a HIP device and stream (torch's current stream by default)."""

    def __init__(self, device=0, stream=None):
        import torch
        self.torch = torch
        self.L = load()
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device)
        self.stream = stream
        handle = C.c_void_p()
        rc = self.L.mbx_ctx_create(device, C.c_void_p(stream.cuda_stream), C.byref(handle))
        if rc != 0:
            raise MbxError("mbx_ctx_create failed: %d (%s)" % (rc, self.L.mbx_last_error(None).decode()))
        self.h = handle
        self.params = None

    def close(self):
        if self.h:
            self.L.mbx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise MbxError("%s failed (%d): %s" % (what, rc,
                                                  self.L.mbx_last_error(self.h).decode()))

    def upload(self, params):
        self._check(self.L.mbx_scene_upload(self.h, C.byref(params)), "mbx_scene_upload")
        self.params = params

    def scene_info(self):
        import numpy as np
        m = np.zeros(9, np.float64)
        fov = np.zeros(1, np.float64)
        light = np.zeros(3, np.float64)
        kind = C.c_int()
        self._check(self.L.mbx_scene_info(self.h, m.ctypes.data_as(C.c_void_p),
                                          fov.ctypes.data_as(C.c_void_p),
                                          light.ctypes.data_as(C.c_void_p), C.byref(kind)),
                    "mbx_scene_info")
        return m.reshape(3, 3), float(fov[0]), light, kind.value

    def alloc_outputs(self, params=None, optional=False):
        """Device planes for mbx_render; with optional, also the world / specular / shadow / GI
        channels (3 floats per pixel each)."""
        torch = self.torch
        p = params or self.params
        n = p.imageWidth * p.imageHeight
        dev = torch.device("cuda", self.device)
        extra = {}
        if optional:
            for name in ("world", "specular", "shadow", "gi"):
                extra[name] = torch.empty((p.imageHeight, p.imageWidth, 3), dtype=torch.float32, device=dev)
        return {**extra, "rgb": torch.empty((p.imageHeight, p.imageWidth, 3), dtype=torch.float32, device=dev),
                "depth": torch.empty((p.imageHeight, p.imageWidth), dtype=torch.float32, device=dev),
                "iters": torch.empty((p.imageHeight, p.imageWidth), dtype=torch.int32, device=dev),
                "mask": torch.empty((p.imageHeight, p.imageWidth), dtype=torch.uint8, device=dev),
                "normal": torch.empty((p.imageHeight, p.imageWidth, 3), dtype=torch.float32, device=dev),
                "steps": torch.empty((p.imageHeight, p.imageWidth), dtype=torch.int32, device=dev),
                "colour": torch.empty((p.imageHeight, p.imageWidth, 3), dtype=torch.uint8, device=dev),
                "opacity": torch.empty((p.imageHeight, p.imageWidth), dtype=torch.uint16, device=dev),
                "_n": n}

    @staticmethod
    def outputs_struct(bufs):
        o = abi.MbxOutputs()
        for name in ("rgb", "depth", "iters", "mask", "normal", "steps", "colour", "opacity", "world", "specular",
                     "shadow", "gi"):
            t = bufs.get(name)
            setattr(o, name, t.data_ptr() if t is not None else None)
        return o

    def render(self, bufs):
        o = self.outputs_struct(bufs)
        self._check(self.L.mbx_render(self.h, C.byref(o)), "mbx_render")

    def render_batch(self, params_list, bufs_list):
        n = len(params_list)
        arr = (abi.MbxParams * n)(*params_list)
        outs = (abi.MbxOutputs * n)(*[self.outputs_struct(b) for b in bufs_list])
        self._check(self.L.mbx_render_batch(self.h, arr, outs, n), "mbx_render_batch")

    def render_batch_noalloc(self, params_list, bufs_list):
        """mbx_render_batch_noalloc: the batch render with the library's static device scratch
        (at most MBX_NOALLOC_MAX_FRAMES frames, no device allocation)."""
        n = len(params_list)
        arr = (abi.MbxParams * max(n, 1))(*params_list)
        outs = (abi.MbxOutputs * max(n, 1))(*[self.outputs_struct(b) for b in bufs_list])
        self._check(self.L.mbx_render_batch_noalloc(self.h, arr, outs, n), "mbx_render_batch_noalloc")

    def compile_image(self, rgb, adj=None, rgb16=None, rgb8=None):
        """cImage::CompileImage + ConvertTo8bitChar on the device: `rgb` is an (H, W, 3) float32
        CUDA tensor (the render's rgb plane); returns the (H, W, 3) uint16 and uint8 tensors.
        `adj` defaults to the uploaded scene's image adjustments."""
        torch = self.torch
        if adj is None:
            if self.params is None:
                raise MbxError("compile_image: no image adjustments given and no scene uploaded")
            adj = self.params.image
        H, W = rgb.shape[0], rgb.shape[1]
        if rgb.dtype != torch.float32 or not rgb.is_contiguous() or tuple(rgb.shape) != (H, W, 3):
            raise MbxError("compile_image: contiguous (H, W, 3) float32 image")
        if rgb16 is None:
            rgb16 = torch.empty((H, W, 3), dtype=torch.uint16, device=rgb.device)
        if rgb8 is None:
            rgb8 = torch.empty((H, W, 3), dtype=torch.uint8, device=rgb.device)
        self._check(self.L.mbx_compile_image(self.h, C.byref(adj), C.c_void_p(rgb.data_ptr()), W, H,
                                             C.c_void_p(rgb16.data_ptr()), C.c_void_p(rgb8.data_ptr())),
                    "mbx_compile_image")
        return rgb16, rgb8

    def compile_zbuffer(self, depth, bits=16, logarithmic=True, invert=False, constant_range=False,
                        min_depth=1e-4, max_depth=10.0):
        """The z-buffer grey image of an (H, W) float32 depth plane: (H, W) uint16 or uint8."""
        torch = self.torch
        if depth.dtype != torch.float32 or not depth.is_contiguous() or depth.dim() != 2:
            raise MbxError("compile_zbuffer: contiguous (H, W) float32 depth")
        out = torch.empty(depth.shape, dtype=torch.uint16 if bits == 16 else torch.uint8, device=depth.device)
        self._check(self.L.mbx_compile_zbuffer(self.h, C.c_void_p(depth.data_ptr()), depth.shape[1], depth.shape[0],
                                               bits, int(logarithmic), int(invert), int(constant_range),
                                               min_depth, max_depth, C.c_void_p(out.data_ptr())),
                    "mbx_compile_zbuffer")
        return out

    def render_ssao(self, bufs, params=None, rgb_out=None):
        """Screen-space AO (mbx_render_ssao) over the planes of a render: bufs holds the (H, W)
        float32 depth, (H, W, 3) uint8 colour, (H, W) uint16 opacity (or None) and (H, W, 3)
        float32 rgb tensors; returns the post-effect image (a new tensor unless rgb_out is given,
        which may be bufs["rgb"] itself)."""
        torch = self.torch
        p = params or self.params
        H, W = p.imageHeight, p.imageWidth
        checks = (("depth", torch.float32, (H, W)), ("colour", torch.uint8, (H, W, 3)),
                  ("rgb", torch.float32, (H, W, 3)))
        for name, dt, shape in checks:
            t = bufs.get(name)
            if t is None or t.dtype != dt or not t.is_contiguous() or tuple(t.shape) != shape:
                raise MbxError("render_ssao: contiguous %s %s plane %s" % (str(dt), name, shape))
        op = bufs.get("opacity")
        if op is not None and (op.dtype != torch.uint16 or not op.is_contiguous() or tuple(op.shape) != (H, W)):
            raise MbxError("render_ssao: contiguous uint16 opacity plane (H, W)")
        if rgb_out is None:
            rgb_out = torch.empty_like(bufs["rgb"])
        self._check(self.L.mbx_render_ssao(self.h, C.byref(p), C.c_void_p(bufs["depth"].data_ptr()),
                                           C.c_void_p(bufs["colour"].data_ptr()),
                                           C.c_void_p(op.data_ptr()) if op is not None else None,
                                           C.c_void_p(bufs["rgb"].data_ptr()),
                                           C.c_void_p(rgb_out.data_ptr())), "mbx_render_ssao")
        return rgb_out

    def tune(self, key, value):
        self._check(self.L.mbx_tune(self.h, key, value), "mbx_tune")

    def alloc_checkpoints(self, params=None):
        """Device buffer for one mbx_ray_checkpoint (24 bytes) per pixel."""
        p = params or self.params
        return self.torch.empty((p.imageHeight * p.imageWidth * 3,), dtype=self.torch.float64,
                                device=self.torch.device("cuda", self.device))

    def render_diff(self, bufs, checkpoints):
        o = self.outputs_struct(bufs)
        self._check(self.L.mbx_render_diff(self.h, C.byref(o), C.c_void_p(checkpoints.data_ptr())),
                    "mbx_render_diff")

    def adjoint(self, checkpoints, dL_drgb, pixel_grad=None):
        """dL/dtheta (numpy, MBX_DIFF_NPARAMS) for dL_drgb (float32 device tensor, 3 per pixel)."""
        import numpy as np
        g = np.zeros(abi.MBX_DIFF_NPARAMS, np.float64)
        assert dL_drgb.dtype == self.torch.float32 and dL_drgb.is_contiguous()
        pg = C.c_void_p(pixel_grad.data_ptr()) if pixel_grad is not None else None
        self._check(self.L.mbx_adjoint(self.h, C.c_void_p(checkpoints.data_ptr()),
                                       C.c_void_p(dL_drgb.data_ptr()),
                                       g.ctypes.data_as(C.c_void_p), pg), "mbx_adjoint")
        return g

    def get_tune(self, key):
        v = C.c_int()
        self._check(self.L.mbx_get_tune(self.h, key, C.byref(v)), "mbx_get_tune")
        return v.value

    def autotune(self, bufs, reps=3):
        o = self.outputs_struct(bufs)
        w = C.c_int()
        ms = C.c_double()
        self._check(self.L.mbx_autotune(self.h, C.byref(o), reps, C.byref(w), C.byref(ms)),
                    "mbx_autotune")
        return w.value, ms.value

    def stats(self):
        s = abi.MbxStats()
        self._check(self.L.mbx_get_stats(self.h, C.byref(s)), "mbx_get_stats")
        return {"rays": s.rays, "de_evals": s.de_evals, "iterations": s.iterations,
                "kernel_ms": s.kernel_ms, "far_steps": s.far_steps}

    def formula_iterate(self, fractal, z, aux):
        """One FormulaCode call of fractal.formula per point on the device (mbx_formula_iterate):
        z is (n, 4) float64, aux an array of n abi.MbxExtendedAux; returns the updated copies."""
        import numpy as np
        z = np.ascontiguousarray(z, np.float64).reshape(-1, 4).copy()
        n = len(z)
        a = (abi.MbxExtendedAux * n)(*aux)
        self._check(self.L.mbx_formula_iterate(self.h, C.byref(fractal), n, z.ctypes.data_as(C.c_void_p),
                                               C.cast(a, C.c_void_p), z.ctypes.data_as(C.c_void_p),
                                               C.cast(a, C.c_void_p)), "mbx_formula_iterate")
        return z, list(a)

    def distance(self, pts, detail_size):
        import numpy as np
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        n = len(pts)
        d = np.zeros(n, np.float32)
        it = np.zeros(n, np.int32)
        self._check(self.L.mbx_distance(self.h, pts.ctypes.data_as(C.c_void_p), n,
                                        float(detail_size), d.ctypes.data_as(C.c_void_p),
                                        it.ctypes.data_as(C.c_void_p)), "mbx_distance")
        return d, it


def render_to_host(params, device=0, kernel=KERNEL_GRID, tune=None, optional=False):
    """Upload, render and copy all planes back as numpy arrays (plus work statistics and the
    kernel kind the scene was specialised for).  `tune` maps MBX_TUNE_* keys to values; with
    optional, the world / specular / shadow / GI channels too."""
    import torch
    ctx = Context(device)
    try:
        ctx.tune(MBX_TUNE_KERNEL, kernel)
        for k, v in (tune or {}).items():
            ctx.tune(k, v)
        ctx.upload(params)
        bufs = ctx.alloc_outputs(optional=optional)
        ctx.render(bufs)
        st = ctx.stats()
        # the post effects cRenderer runs after the render (render_image.cpp:498-508): with
        # screen-space AO the post image is the render's image plus the SSAO term
        if params.ambientOcclusionEnabled and params.ambientOcclusionMode == abi.MBX_AO_SCREEN_SPACE:
            bufs["post_rgb"] = ctx.render_ssao(bufs)
        torch.cuda.synchronize(device)
        out = {k: v.cpu().numpy() for k, v in bufs.items() if not k.startswith("_")}
        out.update(st)
        out["de_kind"] = ctx.scene_info()[3]
        return out
    finally:
        ctx.close()

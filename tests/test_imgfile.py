"""TIFF and OpenEXR writers (mbx_write_tiff, mbx_write_exr) on the host.

* TIFF: read back by Pillow (its libtiff decoder) and by a field reader written here from the
  TIFF 6.0 layout (little-endian header, one IFD, deflated strips of 64 rows): pixels equal, the
  fields the reference sets (compression 32946, contiguous samples, fill order 1, sample format).
* OpenEXR: decoded by a reader written here from the OpenEXR file layout (magic 20000630,
  attributes, scanline offset table, ZIPS chunks: zlib, then the inverse of the format's delta
  predictor and byte interleave); the header bytes of a one-pixel FLOAT image against a
  hand-assembled byte string; HALF values against IEEE binary16 closed forms (numpy's float16,
  round to nearest even).
"""
import struct
import zlib

import numpy as np
import pytest

from mandelbulber2_amd import runtime


@pytest.fixture(scope="module")
def lib():
    return runtime.load()


# ------------------------------------------------------------------------------------------
# TIFF
# ------------------------------------------------------------------------------------------
def _tiff_fields(data):
    assert data[:4] == b"II*\x00"
    ifd = struct.unpack_from("<I", data, 4)[0]
    n = struct.unpack_from("<H", data, ifd)[0]
    fields = {}
    for k in range(n):
        tag, typ, count, val = struct.unpack_from("<HHII", data, ifd + 2 + 12 * k)
        size = {3: 2, 4: 4}[typ] * count
        if size <= 4:
            raw = data[ifd + 2 + 12 * k + 8: ifd + 2 + 12 * k + 8 + size]
        else:
            raw = data[val: val + size]
        fmt = "<%d%s" % (count, "H" if typ == 3 else "I")
        fields[tag] = struct.unpack(fmt, raw)
    assert struct.unpack_from("<I", data, ifd + 2 + 12 * n)[0] == 0
    return fields


def _tiff_pixels(data, dtype, channels):
    f = _tiff_fields(data)
    W, H = f[256][0], f[257][0]
    raw = b"".join(zlib.decompress(data[o: o + c]) for o, c in zip(f[273], f[279]))
    return f, np.frombuffer(raw, dtype=dtype).reshape((H, W) if channels == 1 else (H, W, channels))


@pytest.mark.parametrize("shape,dtype", [((70, 33, 3), np.uint8), ((130, 17, 3), np.uint16), ((40, 21), np.uint16),
                                         ((65, 9, 4), np.uint8), ((64, 11), np.float32), ((10, 12, 3), np.float32)])
def test_tiff_round_trip(tmp_path, lib, shape, dtype):
    rng = np.random.default_rng(len(shape) * 7 + shape[0])
    if dtype == np.float32:
        a = rng.random(shape, dtype=np.float32)
    else:
        a = rng.integers(0, np.iinfo(dtype).max + 1, shape, dtype=dtype)
    path = tmp_path / "x.tiff"
    runtime.write_tiff(path, a)
    data = path.read_bytes()
    channels = 1 if len(shape) == 2 else shape[2]
    f, px = _tiff_pixels(data, a.dtype.newbyteorder("<"), channels)
    assert np.array_equal(px, a)
    bits = a.dtype.itemsize * 8
    assert f[258] == (bits,) * channels and f[277] == (channels,)
    assert f[259] == (32946,) and f[284] == (1,) and f[266] == (1,) and f[278] == (64,)
    assert f[262] == ((2,) if channels >= 3 else (1,))
    assert f[339] == ((3,) if dtype == np.float32 else (1,)) * channels
    assert len(f[273]) == (shape[0] + 63) // 64
    if channels == 4:
        assert f[338] == (2,)
    if dtype != np.float32 and channels != 4:
        from PIL import Image
        with Image.open(path) as im:
            b = np.asarray(im)
        if dtype == np.uint16 and channels == 1:
            b = b.astype(np.uint16)
        if dtype == np.uint16 and channels == 3:
            a = a >> 8          # Pillow opens 48-bit RGB as 8-bit RGB (the high bytes)
        assert np.array_equal(b, a)


def test_tiff_rejects_bad_arguments(tmp_path, lib):
    import ctypes as C
    buf = np.zeros((4, 4, 3), np.uint8)
    for bits, ch in ((12, 3), (8, 2), (8, 5)):
        assert lib.mbx_write_tiff(str(tmp_path / "b.tiff").encode(), 4, 4, bits, ch,
                                  buf.ctypes.data_as(C.c_void_p)) == -1


# ------------------------------------------------------------------------------------------
# OpenEXR
# ------------------------------------------------------------------------------------------
def _exr_read(data):
    assert struct.unpack_from("<I", data, 0)[0] == 20000630
    assert struct.unpack_from("<I", data, 4)[0] == 2
    pos, attrs = 8, {}
    while data[pos] != 0:
        e = data.index(b"\x00", pos)
        name = data[pos:e].decode()
        e2 = data.index(b"\x00", e + 1)
        typ = data[e + 1:e2].decode()
        size = struct.unpack_from("<I", data, e2 + 1)[0]
        attrs[name] = (typ, data[e2 + 5: e2 + 5 + size])
        pos = e2 + 5 + size
    pos += 1
    chl = attrs["channels"][1]
    chans, p = [], 0
    while chl[p] != 0:
        e = chl.index(b"\x00", p)
        name = chl[p:e].decode()
        ptype, linear, xs, ys = struct.unpack_from("<iB3xii", chl, e + 1)
        chans.append((name, ptype, linear))
        p = e + 1 + 16
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from("<%dQ" % H, data, pos)
    sizes = [W * (2 if t == 1 else 4) for _, t, _ in chans]
    line = sum(sizes)
    out = {n: np.zeros((H, W), np.float32) for n, _, _ in chans}
    for y in range(H):
        yy, n = struct.unpack_from("<ii", data, offsets[y])
        assert yy == y
        blob = data[offsets[y] + 8: offsets[y] + 8 + n]
        if n < line:
            t = bytearray(zlib.decompress(blob))
            for k in range(1, len(t)):
                t[k] = (t[k - 1] + t[k] - 128) & 255
            half = (line + 1) // 2
            raw = bytearray(line)
            raw[0::2] = t[:half]
            raw[1::2] = t[half:]
        else:
            raw = blob
        at = 0
        for (name, ptype, _), sz in zip(chans, sizes):
            out[name][y] = np.frombuffer(bytes(raw[at:at + sz]), "<f2" if ptype == 1 else "<f4").astype(np.float32)
            at += sz
    return attrs, chans, out


def test_exr_round_trip_half_and_float(tmp_path, lib):
    H, W = 37, 53
    rng = np.random.default_rng(5)
    rgb = (rng.random((H, W, 3)) * 4.0 - 1.0).astype(np.float32)
    z = (rng.random((H, W)) * 100.0).astype(np.float32)
    z[3, 4] = 1e20                                   # a miss: HALF infinity
    path = tmp_path / "a.exr"
    runtime.write_exr(path, {("R", "G", "B"): rgb, "Z": z}, half=True)
    attrs, chans, out = _exr_read(path.read_bytes())
    assert [c[0] for c in chans] == ["B", "G", "R", "Z"]      # sorted by name
    assert attrs["compression"] == ("compression", b"\x02")   # ZIPS
    for k, name in enumerate("RGB"):
        assert np.array_equal(out[name], rgb[..., k].astype(np.float16).astype(np.float32))
    assert np.array_equal(out["Z"], z.astype(np.float16).astype(np.float32))
    assert np.isinf(out["Z"][3, 4])
    path2 = tmp_path / "b.exr"
    runtime.write_exr(path2, {("n.X", "n.Y", "n.Z"): rgb, "A": z}, half=False, linear=True)
    _, chans2, out2 = _exr_read(path2.read_bytes())
    assert [c[0] for c in chans2] == ["A", "n.X", "n.Y", "n.Z"] and all(c[2] == 1 for c in chans2)
    assert np.array_equal(out2["n.Y"], rgb[..., 1]) and np.array_equal(out2["A"], z)


def test_exr_header_bytes_of_one_pixel(tmp_path, lib):
    """A 1x1 FLOAT channel "Y": magic, version, the eight required attributes in name order,
    the end-of-header byte, one offset, then chunk (y = 0, 4 bytes, the float raw: the zlib
    stream would be longer than 4 bytes)."""
    path = tmp_path / "one.exr"
    runtime.write_exr(path, {"Y": np.full((1, 1), 0.75, np.float32)}, half=False)
    data = path.read_bytes()

    def attr(name, typ, value):
        return name.encode() + b"\x00" + typ.encode() + b"\x00" + struct.pack("<I", len(value)) + value

    hdr = struct.pack("<II", 20000630, 2)
    hdr += attr("channels", "chlist", b"Y\x00" + struct.pack("<iB3xii", 2, 0, 1, 1) + b"\x00")
    hdr += attr("compression", "compression", b"\x02")
    box = struct.pack("<iiii", 0, 0, 0, 0)
    hdr += attr("dataWindow", "box2i", box) + attr("displayWindow", "box2i", box)
    hdr += attr("lineOrder", "lineOrder", b"\x00")
    hdr += attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    hdr += attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    hdr += attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    hdr += b"\x00"
    want = hdr + struct.pack("<Q", len(hdr) + 8) + struct.pack("<ii", 0, 4) + struct.pack("<f", 0.75)
    assert data == want


@pytest.mark.parametrize("v", [0.0, -0.0, 1.0, -2.5, 0.1, 65504.0, 65519.0, 65520.0, 1e5, 6.1e-5, 5.96e-8, 2.9e-8,
                               3.0e-8, 1e-9, 1.0009765625, 1.00048828125, 1.00146484375, float("inf")])
def test_half_rounding(lib, v):
    with np.errstate(over="ignore"):
        want = np.array([v], np.float32).astype(np.float16).view(np.uint16)[0]
    assert lib.mbx_float_to_half(v) == want, (v, hex(lib.mbx_float_to_half(v)), hex(want))


def test_half_nan_stays_nan(lib):
    h = lib.mbx_float_to_half(float("nan"))
    assert (h & 0x7c00) == 0x7c00 and (h & 0x3ff) != 0

"""mbx_write_png (ImageFileSavePNG::SavePNG, mandelbulber2/src/file_image.cpp:481-600): the file
is a PNG that an independent reader decodes to the same samples — a chunk walker with CRC checks
and zlib here, and Pillow for the 8-bit files.  Host only."""
import struct
import zlib

import numpy as np
import pytest

from mandelbulber2_amd import runtime


def _decode(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xffffffff
        chunks.append((kind, body))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, bits, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (comp, filt, inter) == (0, 0, 0)
    ch = {2: 3, 0: 1}[ctype]
    raw = zlib.decompress(chunks[1][1])
    row = w * ch * bits // 8
    assert len(raw) == h * (row + 1)
    rows = [raw[y * (row + 1):(y + 1) * (row + 1)] for y in range(h)]
    assert all(r[0] == 0 for r in rows)
    px = np.frombuffer(b"".join(r[1:] for r in rows), ">u2" if bits == 16 else np.uint8)
    shape = (h, w, 3) if ch == 3 else (h, w)
    return px.reshape(shape).astype(np.uint16 if bits == 16 else np.uint8), bits


@pytest.mark.parametrize("dtype,shape", [(np.uint8, (37, 53, 3)), (np.uint16, (31, 45, 3)),
                                         (np.uint16, (20, 17)), (np.uint8, (9, 64))])
def test_png_round_trip(tmp_path, dtype, shape):
    rs = np.random.default_rng(3)
    px = rs.integers(0, np.iinfo(dtype).max + 1, size=shape, dtype=np.int64).astype(dtype)
    path = tmp_path / "img.png"
    runtime.write_png(path, px)
    got, bits = _decode(path)
    assert bits == (8 if dtype == np.uint8 else 16)
    np.testing.assert_array_equal(got, px)
    if dtype == np.uint8:
        PIL = pytest.importorskip("PIL.Image")
        np.testing.assert_array_equal(np.asarray(PIL.open(path)), px)


def test_png_rejects_bad_arguments(tmp_path):
    with pytest.raises(runtime.MbxError):
        runtime.write_png(tmp_path / "x.png", np.zeros((4, 4, 2), np.uint8))
    with pytest.raises(runtime.MbxError):
        runtime.write_png(tmp_path / "x.png", np.zeros((4, 4, 3), np.float32))
    with pytest.raises(runtime.MbxError):
        runtime.write_png(tmp_path / "no_such_dir" / "x.png", np.zeros((4, 4, 3), np.uint8))

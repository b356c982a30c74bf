"""mbx_compile_image (cImage::CalculatePixel + ConvertTo8bitChar on the device) against the oracle.

Bar: bit-exact 16-bit and 8-bit output without the tone map; with hdr the device's tanhf and the
host's may differ in the last bit, which moves a 16-bit value by at most a few units where the
gamma curve is steep: >= 99.9% of samples exact, 8-bit within 1.
"""
import numpy as np
import pytest

from mandelbulber2_amd import abi, scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from mandelbulber2_amd import runtime
    runtime.load()
    return runtime


def _adj(brightness=1.0, contrast=1.0, gamma=1.0, saturation=1.0, hdr=0):
    a = abi.MbxImageAdjustments()
    a.brightness, a.contrast, a.gamma, a.saturation, a.hdr = brightness, contrast, gamma, saturation, hdr
    return a


CASES = {
    "defaults": _adj(),
    "adjusted": _adj(brightness=1.3, contrast=0.8, gamma=2.2, saturation=1.4),
    "grey_gamma_half": _adj(gamma=0.5, saturation=0.0),
    "hdr": _adj(brightness=1.7, contrast=1.2, gamma=1.6, saturation=0.7, hdr=1),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_compile_image_parity(rt, oracle_mod, name):
    import torch
    adj = CASES[name]
    rs = np.random.default_rng(11)
    img = rs.uniform(-0.3, 1.6, size=(123, 257, 3)).astype(np.float32)   # odd size: a partial block
    print("[case] compile_image %s %s" % (name, img.shape), flush=True)
    ctx = rt.Context(0)
    ctx.params = scene.default_params(8, 8)
    d16, d8 = ctx.compile_image(torch.from_numpy(img).cuda(), adj)
    torch.cuda.synchronize()
    g16, g8 = d16.cpu().numpy(), d8.cpu().numpy()
    o16, o8 = oracle_mod.compile_image(adj, img)
    ctx.close()
    if not adj.hdr:
        np.testing.assert_array_equal(g16, o16)
        np.testing.assert_array_equal(g8, o8)
    else:
        assert (g16 == o16).mean() >= 0.999
        assert np.abs(g8.astype(int) - o8.astype(int)).max() <= 1


def test_unaligned_buffers_take_the_one_pixel_path(rt, oracle_mod):
    """Buffers offset by one pixel (12, 6 and 3 bytes) are not 16/8/4-byte aligned: the kernel's
    one-pixel path must give the same samples as the four-pixel path."""
    import torch
    adj = CASES["adjusted"]
    rs = np.random.default_rng(12)
    H, W = 37, 61
    img = rs.uniform(-0.3, 1.6, size=(H * W + 1, 3)).astype(np.float32)
    flat = torch.from_numpy(img).cuda()
    o16 = torch.zeros((H * W + 1, 3), dtype=torch.uint16, device="cuda")
    o8 = torch.zeros((H * W + 1, 3), dtype=torch.uint8, device="cuda")
    ctx = rt.Context(0)
    ctx.params = scene.default_params(8, 8)
    print("[case] compile_image unaligned %dx%d" % (W, H), flush=True)
    ctx.compile_image(flat[1:].view(H, W, 3), adj, o16[1:].view(H, W, 3), o8[1:].view(H, W, 3))
    torch.cuda.synchronize()
    w16, w8 = oracle_mod.compile_image(adj, img[1:])
    np.testing.assert_array_equal(o16[1:].cpu().numpy(), w16)
    np.testing.assert_array_equal(o8[1:].cpu().numpy(), w8)
    assert (o16[0].cpu().numpy() == 0).all() and (o8[0].cpu().numpy() == 0).all()
    ctx.close()


def test_rendered_frame_with_the_scene_adjustments(rt, oracle_mod):
    """A rendered Mandelbulb frame compiled with the image adjustments of its parameter block, and
    a second compile with another gamma (the gamma table is rebuilt) on the same context."""
    import torch
    p = scene.reference_material(scene.mandelbulb8(96, 72))
    p.image.brightness, p.image.contrast, p.image.gamma, p.image.saturation = 1.2, 1.1, 2.2, 1.3
    ctx = rt.Context(0)
    ctx.upload(p)
    bufs = ctx.alloc_outputs()
    ctx.render(bufs)
    rgb = bufs["rgb"].reshape(p.imageHeight, p.imageWidth, 3).contiguous()
    d16, d8 = ctx.compile_image(rgb)
    torch.cuda.synchronize()
    host = rgb.cpu().numpy()
    o16, o8 = oracle_mod.compile_image(p.image, host)
    np.testing.assert_array_equal(d16.cpu().numpy(), o16)
    np.testing.assert_array_equal(d8.cpu().numpy(), o8)
    adj = _adj(gamma=0.8)
    d16, _ = ctx.compile_image(rgb, adj)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d16.cpu().numpy(), oracle_mod.compile_image(adj, host)[0])
    ctx.close()


def test_compile_image_validation(rt):
    import torch
    ctx = rt.Context(0)
    ctx.params = scene.default_params(8, 8)
    img = torch.zeros((4, 4, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(rt.MbxError, match="gamma"):
        ctx.compile_image(img, _adj(gamma=-1.0))
    with pytest.raises(rt.MbxError):
        ctx.compile_image(img.to(torch.float64))
    ctx.close()


@pytest.mark.parametrize("bits", [16, 8])
@pytest.mark.parametrize("mode", ["log", "linear", "log_inverted", "constant_range"])
def test_zbuffer_image_parity(rt, oracle_mod, mode, bits):
    """mbx_compile_zbuffer on a rendered depth plane (hits and misses) against the oracle: the
    frame's range by the device reduction, then the grey values.  Linear scale bit-exact; the
    logarithmic scale calls the device's logf, whose last bit may differ from the host's: >= 99.9%
    exact and within 1."""
    import torch
    p = scene.mandelbulb8(160, 120)
    ctx = rt.Context(0)
    ctx.upload(p)
    bufs = ctx.alloc_outputs()
    ctx.render(bufs)
    depth = bufs["depth"].reshape(p.imageHeight, p.imageWidth).contiguous()
    kw = {"bits": bits, "logarithmic": mode != "linear", "invert": mode == "log_inverted",
          "constant_range": mode == "constant_range", "min_depth": 1.5, "max_depth": 4.0}
    print("[case] zbuffer %s %d-bit" % (mode, bits), flush=True)
    g = ctx.compile_zbuffer(depth, **kw)
    torch.cuda.synchronize()
    host = depth.cpu().numpy()
    want = oracle_mod.compile_zbuffer(host, **kw)
    got = g.cpu().numpy()
    ctx.close()
    assert (host > 1e19).any() and (host < 1e19).any()
    if mode == "linear":
        np.testing.assert_array_equal(got, want)
    else:
        assert (got == want).mean() >= 0.999
        assert np.abs(got.astype(int) - want.astype(int)).max() <= 1

"""Each formula of the device's table run alone through the C ABI (mbx_formula_iterate: one
<Name>Iteration call per point on the GPU) against the reference's own formula sources
(tests/golden/formula_golden.json, from oracle/_ref/formula_ref).

Bars: formulas built from +, -, *, /, sqrt and fabs are bit-exact (the device's FP64 division
and square root are correctly rounded, and the functions are compiled without contraction);
the Mandelbulb and Xenodreambuie call pow / asin / acos / atan2 / sin / cos, whose device
implementations differ from glibc's by an ulp or so: z and DE within 1e-12 relative (of the
vector's length for z)."""
import numpy as np
import pytest

from mandelbulber2_amd import abi
from tests import golden_formulas as gf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from mandelbulber2_amd import runtime
    runtime.load()
    c = runtime.Context(0)
    yield c
    c.close()


def _aux(z_in, aux_in, it, cin):
    out = []
    for k in range(len(z_in)):
        a = abi.MbxExtendedAux()
        for j in range(4):
            a.c[j] = a.old_z[j] = z_in[k, j]
            a.const_c[j] = cin[k, j] if cin is not None else z_in[k, j]
        a.pos_neg, a.DE0, a.dist, a.pseudoKleinianDE = 1.0, 0.0, 1000.0, 1.0
        a.r, a.DE, a.color = aux_in[k]
        a.actualScale, a.temp1000, a.i = 2.0, 1000.0, it
        out.append(a)
    return out


def test_every_table_formula_against_the_reference(ctx):
    seen = set()
    for tag, p, it, z_in, aux_in, cin, z_out, aux_out, cout in gf.cases(gf.load()):
        formula, trig = gf.FORMULAS[tag]
        seen.add(formula)
        print("[formula %s (%d): %d points, i = %d]" % (tag, formula, len(z_in), it), flush=True)
        z, aux = ctx.formula_iterate(p.fractal, z_in, _aux(z_in, aux_in, it, cin))
        de = np.array([a.DE for a in aux])
        if trig:
            scale = np.linalg.norm(z_out, axis=1)[:, None]
            assert np.all(np.abs(z - z_out) <= 1e-12 * scale), (tag, np.max(np.abs(z - z_out) / scale))
            np.testing.assert_allclose(de, aux_out[:, 1], rtol=1e-12, err_msg=tag)
        else:
            np.testing.assert_array_equal(z, z_out, err_msg=tag)
            np.testing.assert_array_equal(de, aux_out[:, 1], err_msg=tag)
        if tag == "mandelbox":
            np.testing.assert_array_equal(np.array([a.color for a in aux]), aux_out[:, 2], err_msg=tag)
        if cout is not None:
            np.testing.assert_array_equal(np.array([list(a.const_c) for a in aux]), cout, err_msg=tag)
        # the formula leaves r and the fields it does not own as they came in
        np.testing.assert_array_equal(np.array([a.r for a in aux]), aux_in[:, 0], err_msg=tag)
        assert all(a.i == it and a.temp1000 == 1000.0 for a in aux), tag
    from mandelbulber2_amd import runtime
    assert seen == {i.internalID for i in runtime.formulas()}


def test_formula_outside_the_table_is_rejected(ctx):
    from mandelbulber2_amd import runtime, scene
    p = scene.default_params()
    p.fractal.formula = 9           # fast_mandelbulb_power2, not in the device's table
    with pytest.raises(runtime.MbxError):
        ctx.formula_iterate(p.fractal, np.zeros((1, 4)), [abi.MbxExtendedAux()])

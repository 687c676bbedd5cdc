"""GPU parity of the five proximity ops through the C ABI against the CPU oracle.

Bars:
* F32 / U32 / S32 / F64 sources, SqrDistance / AbsDistance / CrossCorr: bit-exact against
  the reference-order restatement with fused multiply-add (cl-files/ImageProximity.cl:55-70
  under OpenCL's default FP contraction);
* 8/16-bit integer sources, SqrDistance / AbsDistance / CrossCorr: bit-exact against the
  exact int64 sums rounded to float;
* the Norm ops (denominator from the FP64 sum-of-squares integral): within 1e-5 relative
  of the reference-order map (plus a 1e-6 absolute floor) wherever the window holds at
  least 1e-6 of the image's energy, and everywhere within the float32 accumulation bound
  plus the integral's conditioning term of the FP64 oracle (see norm_close).
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import OP_CODE, band_oracle, integral_kappa, run_op, say

pytestmark = pytest.mark.gpu

RAW_OPS = ["SqrDistance", "AbsDistance", "CrossCorr"]
NORM_OPS = ["SqrDistance_Norm", "CrossCorr_Norm"]
SIZES = [(37, 29), (64, 64), (130, 70)]
TEMPLATES = [(1, 1), (2, 3), (5, 5), (8, 8), (16, 16), (17, 15), (31, 31), (33, 7), (12, 40)]


@pytest.fixture(scope="module")
def ocip():
    from openclipp_amd import ocip as mod
    ctx = mod.initialize()
    yield mod
    mod.lib.ocipUninitialize(ctx)


def rand(rng, shape, dtype):
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        return rng.integers(info.min, info.max, size=shape, endpoint=True).astype(dtype)
    return rng.random(shape).astype(dtype)


def norm_close(got, ref, f64, n, src=None, tpl=None):
    """Norm bar.  The denominator's window sum of s^2 is a 4-corner difference of the FP64
    integral, whose absolute error is ~eps64 * (total image energy); its relative error is
    therefore eps64 * kappa with kappa = total energy / window energy.  Where the window
    holds at least 1e-6 of the image energy (kappa <= 1e6) the bar is 1e-5 relative to the
    reference-order map; elsewhere the map is ill-conditioned and is held to the FP64 oracle
    with the conditioning term included."""
    got = got.astype(np.float64)
    ref = ref.astype(np.float64)
    kappa = np.ones_like(f64) if src is None else integral_kappa(src, tpl)
    well = kappa <= 1e6
    d = np.abs(got - ref)
    assert np.all(d[well] <= 1e-5 * np.abs(ref[well]) + 1e-6), np.max(d[well])
    bound = (3 * (n + 4) * 2.0 ** -24 + 16 * 2.0 ** -53 * kappa) * np.abs(f64) + 1e-6
    assert np.all(np.abs(got - f64) <= bound)


@pytest.mark.parametrize("W,H", SIZES)
@pytest.mark.parametrize("tw,th", TEMPLATES)
def test_f32_raw_bit_exact(ocip, W, H, tw, th):
    if tw > W or th > H:
        pytest.skip("template larger than source")
    rng = np.random.default_rng(W * 1000 + H * 10 + tw * 7 + th)
    src = rand(rng, (H, W), np.float32)
    tpl = rand(rng, (th, tw), np.float32)
    for op in RAW_OPS:
        say("f32", op, W, H, tw, th)
        got = run_op(ocip, op, src, tpl)
        want = orc.prox_ref(OP_CODE[op], src, tpl, contract=True)
        np.testing.assert_array_equal(got, want, err_msg=op)


@pytest.mark.parametrize("W,H", SIZES)
@pytest.mark.parametrize("tw,th", [(1, 1), (4, 4), (16, 16), (17, 15), (31, 31)])
def test_f32_norm(ocip, W, H, tw, th):
    if tw > W or th > H:
        pytest.skip("template larger than source")
    rng = np.random.default_rng(5 + W + tw)
    src = rand(rng, (H, W), np.float32)
    tpl = rand(rng, (th, tw), np.float32)
    for op in NORM_OPS:
        say("f32", op, W, H, tw, th)
        got = run_op(ocip, op, src, tpl)
        norm_close(got, orc.prox_ref(OP_CODE[op], src, tpl), orc.prox_f64(OP_CODE[op], src, tpl),
                   tw * th, src, tpl)


@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16, np.int16])
@pytest.mark.parametrize("tw,th", [(3, 3), (16, 16), (32, 32), (9, 20), (40, 7), (20, 5), (35, 4)])
def test_integer_exact(ocip, dtype, tw, th):
    rng = np.random.default_rng(tw * th)
    src = rand(rng, (70, 90), dtype)
    tpl = rand(rng, (th, tw), dtype)
    for op in RAW_OPS:
        say(np.dtype(dtype).name, op, tw, th)
        got = run_op(ocip, op, src, tpl)
        want = orc.f64_from_i64(OP_CODE[op], src, tpl).astype(np.float32)
        np.testing.assert_array_equal(got, want, err_msg=op)
    for op in NORM_OPS:
        got = run_op(ocip, op, src, tpl).astype(np.float64)
        want = orc.f64_from_i64(OP_CODE[op], src, tpl)
        np.testing.assert_allclose(got, want, rtol=2.0 ** -23, atol=0)


@pytest.mark.parametrize("dtype", [np.uint32, np.int32])
def test_32bit_integer_reference_order(ocip, dtype):
    """REAL is float for 32-bit integer sources (cl-files/Buffers.h:58-63): bit-exact
    against the reference-order float restatement."""
    rng = np.random.default_rng(77)
    src = rng.integers(0, 5000, size=(50, 61)).astype(dtype)
    tpl = rng.integers(0, 5000, size=(9, 7)).astype(dtype)
    for op in RAW_OPS:
        got = run_op(ocip, op, src, tpl)
        np.testing.assert_array_equal(got, orc.prox_ref(OP_CODE[op], src, tpl), err_msg=op)


def test_f64_source(ocip):
    rng = np.random.default_rng(64)
    src = rng.random((41, 43))
    tpl = rng.random((7, 6))
    for op in RAW_OPS:
        got = run_op(ocip, op, src, tpl)
        assert got.dtype == np.float64
        np.testing.assert_array_equal(got, orc.prox_ref(OP_CODE[op], src, tpl), err_msg=op)
    for op in NORM_OPS:
        got = run_op(ocip, op, src, tpl)
        np.testing.assert_allclose(got, orc.prox_f64(OP_CODE[op], src, tpl), rtol=1e-12)


@pytest.mark.parametrize("nch", [2, 4])
@pytest.mark.parametrize("dtype", [np.float32, np.uint8])
def test_multichannel(ocip, nch, dtype):
    rng = np.random.default_rng(nch)
    src = rand(rng, (33, 40, nch), dtype)
    tpl = rand(rng, (6, 5, nch), dtype)
    for op in RAW_OPS:
        got = run_op(ocip, op, src, tpl)
        if dtype == np.uint8:
            want = orc.f64_from_i64(OP_CODE[op], src, tpl).astype(np.float32)
        else:
            want = orc.prox_ref(OP_CODE[op], src, tpl)
        np.testing.assert_array_equal(got, want, err_msg=op)
    for op in NORM_OPS:
        got = run_op(ocip, op, src, tpl).astype(np.float64)
        np.testing.assert_allclose(got, orc.prox_f64(OP_CODE[op], src, tpl), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("dtype", [np.float32, np.uint8])
def test_colour_images(ocip, dtype):
    """3-channel images are colour images: 4 channels on the device with the fourth set to 255
    (reference ColorImage), converted on send and read.  Each channel of a map equals the
    1-channel map of that channel; the host sees 3 channels; a read restores the host data."""
    rng = np.random.default_rng(5)
    src = rand(rng, (29, 37, 3), dtype)
    tpl = rand(rng, (5, 6, 3), dtype)
    for op in RAW_OPS + NORM_OPS:
        got = run_op(ocip, op, src, tpl)
        assert got.shape == (29, 37, 3)
        for c in range(3):
            s1, t1 = np.ascontiguousarray(src[:, :, c]), np.ascontiguousarray(tpl[:, :, c])
            if op in NORM_OPS:
                np.testing.assert_allclose(got[:, :, c], orc.prox_f64(OP_CODE[op], s1, t1),
                                           rtol=1e-5, atol=1e-6, err_msg=op)
            elif dtype == np.uint8:
                want = orc.f64_from_i64(OP_CODE[op], s1, t1).astype(np.float32)
                np.testing.assert_array_equal(got[:, :, c], want, err_msg=op)
            else:
                np.testing.assert_array_equal(got[:, :, c], orc.prox_ref(OP_CODE[op], s1, t1),
                                              err_msg=op)
    # send, clear the host copy, read back
    img = ocip.Image(src.copy())
    img.send()
    img.arr[...] = 0
    np.testing.assert_array_equal(img.read(), src)
    # the device image has 4 channels: Sum reports the fourth (255 per pixel) as well
    res = (ctypes.c_double * 4)()
    ocip.check(ocip.lib.ocipSum(None, img.handle, res))
    assert list(res)[:3] == [float(src[:, :, c].astype(np.float64).sum()) for c in range(3)]
    assert res[3] == 255.0 * 29 * 37


def test_pitched_images(ocip):
    """Source with padding at the end of each row, template that is an ROI of the source
    (the Bench's template, Bench/src/benchImageProximity.hpp:83-86)."""
    rng = np.random.default_rng(3)
    big = rng.random((60, 80), dtype=np.float32)
    src = big[:, :70]                    # pitch 80 floats, width 70
    tpl = big[10:26, 10:26]              # 16x16 ROI with the source's pitch
    assert src.strides[0] == 320 and tpl.strides[0] == 320
    for op in RAW_OPS:
        got = run_op(ocip, op, src, tpl)
        np.testing.assert_array_equal(got, orc.prox_ref(OP_CODE[op], src, tpl), err_msg=op)


def test_template_too_wide_for_lds_tile(ocip):
    """Templates wider than the LDS tile budget take the direct kernel (same order)."""
    rng = np.random.default_rng(300)
    src = rng.random((24, 320), dtype=np.float32)
    tpl = rng.random((3, 300), dtype=np.float32)
    for op in RAW_OPS:
        got = run_op(ocip, op, src, tpl)
        np.testing.assert_array_equal(got, orc.prox_ref(OP_CODE[op], src, tpl), err_msg=op)


def test_template_equal_to_source(ocip):
    rng = np.random.default_rng(4)
    src = rng.random((21, 18), dtype=np.float32)
    got = run_op(ocip, "CrossCorr", src, src.copy())
    np.testing.assert_array_equal(got, orc.prox_ref(orc.CC, src, src.copy()))


def test_nan_on_empty_window(ocip):
    src = np.zeros((20, 20), np.float32)
    src[10, 10] = 1
    tpl = np.ones((3, 3), np.float32)
    got = run_op(ocip, "CrossCorr_Norm", src, tpl)
    assert np.isnan(got[0, 0]) and got[10, 10] == np.float32(1 / 3)


# ------------------------------------------------------------------------ full-size workloads
@pytest.mark.parametrize("dtype", [np.float32, np.uint8])
def test_mid_size_bands(ocip, dtype):
    """1536 x 1408 with a 16 x 16 template: the tile of 8 outputs per thread (images this
    size keep >= 1024 workgroups with it; smaller ones take the 4-output tile), checked on
    bands of rows against the oracle, bit-exact."""
    rng = np.random.default_rng(1536)
    src = rand(rng, (1408, 1536), dtype)
    tpl = rand(rng, (16, 16), dtype)
    for op in RAW_OPS:
        say("mid-size", np.dtype(dtype).name, op)
        got = run_op(ocip, op, src, tpl)
        if dtype == np.float32:
            fn = lambda c: orc.prox_ref(OP_CODE[op], c, tpl, contract=True)
        else:
            fn = lambda c: orc.f64_from_i64(OP_CODE[op], c, tpl).astype(np.float32)
        for y0, y1 in [(0, 9), (700, 706), (1400, 1408)]:
            np.testing.assert_array_equal(got[y0:y1], band_oracle(fn, src, tpl, y0, y1),
                                          err_msg=op)


def test_plain_1024_invariants(ocip):
    from openclipp_amd import workloads
    src, tpl, (ox, oy) = workloads.plain_1024()
    say("plain_1024")
    ncc = run_op(ocip, "CrossCorr_Norm", src, tpl)
    ssd = run_op(ocip, "SqrDistance", src, tpl)
    y, x = oy + 8, ox + 8
    assert np.unravel_index(np.nanargmax(ncc), ncc.shape) == (y, x)
    assert np.unravel_index(np.argmin(ssd), ssd.shape) == (y, x)
    assert np.nanmax(ncc) <= 1.0 + 1e-6
    for y0, y1 in [(0, 8), (500, 508), (1016, 1024), (y - 2, y + 3)]:
        want = band_oracle(lambda c: orc.prox_ref(orc.SSD, c, tpl), src, tpl, y0, y1)
        np.testing.assert_array_equal(ssd[y0:y1], want)
        want = band_oracle(lambda c: orc.prox_ref(orc.CC_NORM, c, tpl), src, tpl, y0, y1)
        assert np.all(np.abs(ncc[y0:y1] - want) <= 1e-5 * np.abs(want) + 1e-6)


def test_integer_4096_bands(ocip):
    from openclipp_amd import workloads
    src, tpl, (ox, oy) = workloads.integer_4096()
    y, x = oy + 16, ox + 16
    for op in RAW_OPS + NORM_OPS:
        say("integer_4096", op)
        got = run_op(ocip, op, src, tpl)
        for y0, y1 in [(0, 4), (2046, 2050), (4092, 4096), (y - 1, y + 2)]:
            want = band_oracle(lambda c: orc.f64_from_i64(OP_CODE[op], c, tpl), src, tpl, y0, y1)
            if op in RAW_OPS:
                np.testing.assert_array_equal(got[y0:y1], want.astype(np.float32), err_msg=op)
            else:
                np.testing.assert_allclose(got[y0:y1], want, rtol=2.0 ** -23, atol=0, err_msg=op)
        if op == "SqrDistance":
            # exact value at the crop: the sum of squared noise
            noise = tpl.astype(np.int64) - src[oy:oy + 32, ox:ox + 32].astype(np.int64)
            assert got[y, x] == np.float32((noise * noise).sum())


def test_integer_4096_float_maps_reference_order(ocip):
    """The U8 workload's float maps (CrossCorr, CrossCorr_Norm, SqrDistance_Norm) against the
    reference's own float-order sums (REAL = float, ImageProximity.cl:290-412) on 256 bands of
    2 rows spread over the image: within 1e-5 relative.  The library's sums are exact, so the
    difference is the reference's float rounding (at most 7e-6 on these rows)."""
    from openclipp_amd import workloads
    src, tpl, _ = workloads.integer_4096()
    y0s = np.linspace(0, 4094, 256).astype(int)
    rows = np.unique(np.concatenate([y0s, y0s + 1]))
    for op in ["CrossCorr", "CrossCorr_Norm", "SqrDistance_Norm"]:
        say("integer_4096 reference order", op)
        got = run_op(ocip, op, src, tpl)[rows].astype(np.float64)
        want = orc.prox_ref_rows(OP_CODE[op], src, tpl, rows).astype(np.float64)
        rel = np.abs(got - want) / np.abs(want)
        say("  max relative difference", float(rel.max()))
        assert rel.max() <= 1e-5, op


# (the window sums of s^2 take one 128-column chunk per row up to tw = 64 and two beyond, a
# ring of th rows, 64-row bands of 64 columns: shapes across each of these boundaries, images
# whose last block holds one column, and a template nearly as large as the image)
# (image rows, columns), (template rows, columns)
INT_SHAPES = [((37, 45), (1, 1)), ((70, 90), (3, 5)), ((200, 260), (32, 32)),
              ((150, 300), (97, 40)), ((300, 130), (16, 97)), ((129, 250), (64, 64)),
              ((260, 140), (31, 8)), ((52, 70), (50, 65)), ((70, 129), (1, 66)),
              ((90, 257), (5, 97)), ((130, 385), (33, 2))]


@pytest.mark.parametrize("dtype", [np.uint8, np.int8])
@pytest.mark.parametrize("shape,tshape", INT_SHAPES)
def test_int_mfma_equals_valu_path(ocip, monkeypatch, dtype, shape, tshape):
    """8-bit images: the exact-integer MFMA path (BF16 products, FP32 accumulators flushed
    into int32) against the VALU kernel (prox_tile_int, OCIP_INT_MFMA=0), bit for bit, for
    the four cross-term ops; the raw ops also against the exact int64 oracle."""
    rng = np.random.default_rng(shape[0] + tshape[0] * 7 + np.dtype(dtype).itemsize)
    info = np.iinfo(dtype)
    src = rng.integers(info.min, info.max, size=shape, endpoint=True).astype(dtype)
    tpl = rng.integers(info.min, info.max, size=tshape, endpoint=True).astype(dtype)
    for op in ["CrossCorr", "SqrDistance", "CrossCorr_Norm", "SqrDistance_Norm"]:
        say("int mfma", np.dtype(dtype).name, shape, tshape, op)
        monkeypatch.delenv("OCIP_INT_MFMA", raising=False)
        got = run_op(ocip, op, src, tpl)
        monkeypatch.setenv("OCIP_INT_MFMA", "0")
        want = run_op(ocip, op, src, tpl)
        np.testing.assert_array_equal(got, want, err_msg=op)
        if op in RAW_OPS:
            exact = orc.f64_from_i64(OP_CODE[op], src, tpl).astype(np.float32)
            np.testing.assert_array_equal(got, exact, err_msg=op)
    monkeypatch.delenv("OCIP_INT_MFMA", raising=False)


@pytest.mark.parametrize("dtype", [np.uint8, np.int8])
def test_int_mfma_pitched_source(ocip, monkeypatch, dtype):
    """A pitched 8-bit source whose width is not a multiple of 4 (the bytes past the width in
    each row's pitch must read as zero in the window sums and the staged tile): the MFMA path
    equals the VALU path bit for bit."""
    rng = np.random.default_rng(131)
    info = np.iinfo(dtype)
    big = rng.integers(info.min, info.max, size=(90, 136), endpoint=True).astype(dtype)
    src = big[:, :130]
    assert src.strides[0] == 136
    tpl = rng.integers(info.min, info.max, size=(17, 23), endpoint=True).astype(dtype)
    for op in ["SqrDistance", "CrossCorr_Norm"]:
        say("int mfma pitched", np.dtype(dtype).name, op)
        monkeypatch.delenv("OCIP_INT_MFMA", raising=False)
        got = run_op(ocip, op, src, tpl)
        monkeypatch.setenv("OCIP_INT_MFMA", "0")
        want = run_op(ocip, op, src, tpl)
        np.testing.assert_array_equal(got, want, err_msg=op)
    monkeypatch.delenv("OCIP_INT_MFMA", raising=False)


@pytest.mark.parametrize("dtype,value", [(np.uint8, 255), (np.int8, -128), (np.int8, 127)])
def test_int_mfma_largest_products(ocip, dtype, value):
    """Constant images and templates of the largest magnitude: every product is max|s t|, so
    the FP32 accumulators reach the flush bound (a 97-column U8 template flushes every 2 steps)
    and the int32 sums their largest values; against exact int64 sums."""
    src = np.full((300, 260), value, dtype)
    tpl = np.full((96, 97), value, dtype)
    for op in ["CrossCorr", "SqrDistance"]:
        say("int mfma largest", np.dtype(dtype).name, value, op)
        got = run_op(ocip, op, src, tpl)
        exact = orc.f64_from_i64(OP_CODE[op], src, tpl).astype(np.float32)
        np.testing.assert_array_equal(got, exact, err_msg=op)


def test_precision_4096_plain_bands(ocip):
    from openclipp_amd import workloads
    src, tpl, (ox, oy) = workloads.precision_4096()
    say("precision_4096 plain")
    cc = run_op(ocip, "CrossCorr", src, tpl)
    ncc = run_op(ocip, "CrossCorr_Norm", src, tpl)
    y, x = oy + 32, ox + 32
    assert abs(ncc[y, x] - 1.0) < 1e-5
    for y0, y1 in [(0, 2), (2047, 2049), (4094, 4096)]:
        want = band_oracle(lambda c: orc.prox_ref(orc.CC, c, tpl, nthreads=8), src, tpl, y0, y1)
        np.testing.assert_array_equal(cc[y0:y1], want)
    # CrossCorr_Norm on 64 rows spread over the image (the first and last two included)
    # against the reference-order map: within 1e-5 relative (1e-6 absolute floor)
    rows = np.unique(np.concatenate([[0, 1, 4094, 4095], np.linspace(2, 4093, 60).astype(int)]))
    want = orc.prox_ref_rows(orc.CC_NORM, src, tpl, rows).astype(np.float64)
    got = ncc[rows].astype(np.float64)
    assert np.all(np.abs(got - want) <= 1e-5 * np.abs(want) + 1e-6)


def test_golden_fixtures_through_the_abi(ocip):
    """The committed numpy brute-force fixtures (tests/golden) through the C ABI."""
    import glob
    import os
    for path in sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))):
        g = np.load(path)
        src, tpl = g["src"], g["tpl"]
        n = tpl.size
        for op in RAW_OPS + NORM_OPS:
            got = run_op(ocip, op, src, tpl).astype(np.float64)
            if np.issubdtype(src.dtype, np.integer) and op in RAW_OPS:
                np.testing.assert_array_equal(got, g[op + "_i64"].astype(np.float32), err_msg=path)
            else:
                # Norm maps: plus the FP64 integral's conditioning term
                kappa = integral_kappa(src, tpl) if op in NORM_OPS else 0.0
                unit = 1e-12 if src.dtype == np.float64 else 3 * (n + 4) * 2.0 ** -24
                bound = (unit + 16 * 2.0 ** -53 * kappa) * np.abs(g[op]) + \
                    (1e-12 if src.dtype == np.float64 else 1e-6)
                assert np.all(np.abs(got - g[op]) <= bound), (path, op)

"""GPU parity of ocipIntegral / ocipSqrIntegral, ocipMinIndx / ocipMaxIndx and the ABI's
error paths, through the C ABI.

Integral bar: 8/16-bit integer sources are bit-exact against the exact sums (F64 dest) and
against the exact sums rounded once to float (F32 dest); F32 sources are within 1e-13
relative of the exact (long double) sums for F64 dest and within one float ulp for F32;
32-bit integer sources are bit-exact while their sums stay below 2^53, and they (squared) and
F64 sources are otherwise within 257 * 2^-53 * sum|terms| (one more float rounding for F32).

MinIndx / MaxIndx: the value and pixel of the reference-order oracle (its tie order, its float
partials), on every element type.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ocip():
    from openclipp_amd import ocip as mod
    ctx = mod.initialize()
    yield mod
    mod.lib.ocipUninitialize(ctx)


def run_integral(ocip, src, sqr, dtype):
    S = ocip.Image(src)
    D = ocip.Image(shape=src.shape, dtype=dtype)
    prog = ctypes.c_void_p()
    ocip.check(ocip.lib.ocipPrepareIntegral(ctypes.byref(prog), S.handle), "prepare")
    fn = ocip.lib.ocipSqrIntegral if sqr else ocip.lib.ocipIntegral
    ocip.check(fn(prog, S.handle, D.handle), "integral")
    out = D.read().copy()
    ocip.check(ocip.lib.ocipReleaseProgram(prog), "release")
    return out


SHAPES = [(1, 1), (5, 7), (8, 64), (37, 1023), (130, 1025), (64, 4096), (300, 2049)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16, np.int16])
@pytest.mark.parametrize("sqr", [False, True])
def test_integral_integer_exact(ocip, shape, dtype, sqr):
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    info = np.iinfo(dtype)
    src = rng.integers(info.min, info.max, size=shape, endpoint=True).astype(dtype)
    exact = orc.integral_exact(src, sqr)
    say("integral", shape, np.dtype(dtype).name, sqr)
    np.testing.assert_array_equal(run_integral(ocip, src, sqr, np.float64), exact)
    np.testing.assert_array_equal(run_integral(ocip, src, sqr, np.float32), exact.astype(np.float32))


@pytest.mark.parametrize("shape", [(33, 70), (257, 1100)])
@pytest.mark.parametrize("sqr", [False, True])
def test_integral_float(ocip, shape, sqr):
    rng = np.random.default_rng(1)
    src = rng.random(shape).astype(np.float32)
    exact = orc.integral_exact(src, sqr)
    np.testing.assert_allclose(run_integral(ocip, src, sqr, np.float64), exact, rtol=1e-13)
    got32 = run_integral(ocip, src, sqr, np.float32).astype(np.float64)
    assert np.all(np.abs(got32 - exact) <= np.spacing(exact.astype(np.float32)).astype(np.float64))


@pytest.mark.parametrize("shape", [(1, 1), (37, 1023), (300, 2049)])
@pytest.mark.parametrize("dtype", [np.uint32, np.int32, np.float64])
@pytest.mark.parametrize("sqr", [False, True])
def test_integral_wide_sources(ocip, shape, dtype, sqr):
    """32-bit integer and F64 sources over their full range.  The FP64 sums are exact while
    every partial stays below 2^53 (32-bit sums here); otherwise (squares of 32-bit values,
    F64 sources) the error of any summation tree of depth d over the terms is at most
    d 2^-53 sum|terms| (+ 2^-53 per rounded square): d < 256 covers the in-tile column run
    (32 rows), the wavefront scan and the carries over tile rows and columns.  F32 dest adds
    one rounding to float."""
    rng = np.random.default_rng(shape[0] * 5 + shape[1] + np.dtype(dtype).itemsize)
    if dtype == np.float64:
        src = rng.standard_normal(shape) * 1e3
    else:
        info = np.iinfo(dtype)
        src = rng.integers(info.min, info.max, size=shape, endpoint=True, dtype=np.int64)
        src = src.astype(dtype)
    say("integral", shape, np.dtype(dtype).name, sqr)
    exact = orc.integral_exact(src, sqr)
    a = np.abs(src.astype(np.float64))
    absint = orc.integral_exact(a, sqr)                            # sum |terms|
    exact_fp64 = dtype != np.float64 and not sqr and absint.max() < 2.0 ** 53
    got = run_integral(ocip, src, sqr, np.float64)
    if exact_fp64:
        np.testing.assert_array_equal(got, exact)
    else:
        assert np.all(np.abs(got - exact) <= 257 * 2.0 ** -53 * absint)
    got32 = run_integral(ocip, src, sqr, np.float32).astype(np.float64)
    if exact_fp64:
        np.testing.assert_array_equal(got32, exact.astype(np.float32).astype(np.float64))
    else:
        ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(got32 - exact) <= ulp + 257 * 2.0 ** -53 * absint)


def test_integral_4096_u8_sqr_exact(ocip):
    from openclipp_amd import workloads
    src, _, _ = workloads.integer_4096()
    say("integral 4096 u8")
    got = run_integral(ocip, src, True, np.float64)
    a = src.astype(np.int64)
    want = np.cumsum(np.cumsum(a * a, axis=0), axis=1).astype(np.float64)
    np.testing.assert_array_equal(got, want)
    # invariant: the last element is the total
    assert got[-1, -1] == float((a * a).sum())


def test_integral_argument_errors(ocip):
    src = np.zeros((8, 8), np.uint8)
    S = ocip.Image(src)
    prog = ctypes.c_void_p()
    ocip.check(ocip.lib.ocipPrepareIntegral(ctypes.byref(prog), S.handle))
    bad_type = ocip.Image(shape=(8, 8), dtype=np.int32)
    bad_size = ocip.Image(shape=(8, 9), dtype=np.float32)
    assert ocip.lib.ocipIntegral(prog, S.handle, bad_type.handle) == ocip.CL_INVALID_VALUE
    assert ocip.lib.ocipIntegral(prog, S.handle, bad_size.handle) == ocip.CL_INVALID_IMAGE_SIZE
    assert ocip.lib.ocipIntegral(None, S.handle, bad_size.handle) == ocip.CL_INVALID_PROGRAM
    ocip.check(ocip.lib.ocipReleaseProgram(prog))


# ----------------------------------------------------------------------------- statistics
def _min_max(ocip, src, want_max):
    S = ocip.Image(src)
    v, x, y = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
    fn = ocip.lib.ocipMaxIndx if want_max else ocip.lib.ocipMinIndx
    ocip.check(fn(None, S.handle, ctypes.byref(v), ctypes.byref(x), ctypes.byref(y)))
    S.release()
    return v.value, x.value, y.value


def _tie_heavy(rng, dtype, shape):
    """Few distinct values, so most extremes are tied; 32-bit and F64 values that differ
    below float precision also tie between the groups' float partials."""
    if dtype in (np.uint32, np.int32):
        return (2 ** 30 + rng.integers(0, 4, shape) * 16).astype(dtype)
    if dtype == np.float64:
        return 1.0 + rng.integers(0, 3, shape) * 2.0 ** -40
    if dtype == np.float32:
        return rng.integers(0, 3, shape).astype(np.float32) * np.float32(0.25)
    info = np.iinfo(dtype)
    return rng.integers(max(info.min, -2), min(info.max, 3), shape).astype(dtype)


@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16, np.int16, np.uint32,
                                   np.int32, np.float32, np.float64])
def test_min_max_indx(ocip, dtype):
    """Value and pixel equal the reference-order oracle: random values, tie-heavy values on
    ragged sizes (partial work-groups on both edges), and a single-pixel image."""
    rng = np.random.default_rng(9)
    cases = [rng.integers(0, 100, size=(300, 517)).astype(dtype)
             if np.issubdtype(dtype, np.integer) else rng.random((300, 517)).astype(dtype)]
    for shape in [(1, 1), (16, 256), (17, 257), (45, 530), (300, 517)]:
        cases.append(_tie_heavy(rng, dtype, shape))
    for src in cases:
        say("min/max", np.dtype(dtype).name, src.shape)
        for want_max in (True, False):
            assert _min_max(ocip, src, want_max) == orc.min_max_indx(src, want_max), \
                (src.shape, want_max)


def test_min_max_indx_reference_ties(ocip):
    """Hand-derived known answers of the reference's order: equal maxima at (1, 0) and
    (16, 0) give (16, 0) (work-item 0 holds column 16); (5, 1) and (6, 0) give (6, 0)."""
    img = np.zeros((4, 40), np.uint8)
    img[0, 1] = img[0, 16] = 9
    assert _min_max(ocip, img, True) == (9.0, 16, 0)
    img = np.full((4, 40), 9, np.uint8)
    img[1, 5] = img[0, 6] = 0
    assert _min_max(ocip, img, False) == (0.0, 6, 0)


def test_min_max_indx_nan(ocip):
    a = np.ones((20, 300), np.float32)
    a[3, 7] = 5.0
    a[1, 1] = np.nan
    assert _min_max(ocip, a, True) == (5.0, 7, 3)
    a[0, 0] = np.nan
    v, x, y = _min_max(ocip, a, True)
    assert np.isnan(v) and (x, y) == (0, 0)


def test_max_indx_of_periodic_ncc_map(ocip):
    """A saturated match map: the U8 image repeats the template, so CrossCorr_Norm (exact
    integer sums, exact FP64 window sums) holds bit-identical maxima at every interior
    period.  The located pixel is the reference's tie choice on that map, on the period
    grid."""
    rng = np.random.default_rng(21)
    tpl = rng.integers(1, 256, (12, 10)).astype(np.uint8)
    src = np.tile(tpl, (30, 60))                                 # 360 x 600
    S, T = ocip.Image(src), ocip.Image(tpl)
    D = ocip.Image(shape=src.shape, dtype=np.float32)
    ocip.check(ocip.proximity("CrossCorr_Norm", S, T, D), "CrossCorr_Norm")
    ncc = D.read().copy()
    for i in (S, T, D):
        i.release()
    want = orc.min_max_indx(ncc, True)
    assert (ncc == np.float32(want[0])).sum() > 100              # really tie-heavy
    assert _min_max(ocip, ncc, True) == want
    assert want[1] % 10 == 5 and want[2] % 12 == 6


def test_min_max_indx_4096_u8(ocip):
    """The 4096x4096 U8 workload image: 4096 work-groups, ties everywhere at 0 and 255."""
    from openclipp_amd import workloads
    src = workloads.integer_4096()[0]
    for want_max in (True, False):
        assert _min_max(ocip, src, want_max) == orc.min_max_indx(src, want_max)


SUM_CASES = [((1, 1), 1), ((300, 517), 1), ((64, 4096), 1), ((37, 129), 2), ((20, 300), 4)]


@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16, np.int16, np.int32,
                                   np.float32, np.float64])
@pytest.mark.parametrize("shape,channels", SUM_CASES)
def test_sum_and_sum_sqr(ocip, dtype, shape, channels):
    """ocipSum / ocipSumSqr per channel: 8/16-bit sources bit-exact (exact int64 sums rounded
    once to double); 32-bit and float sources within 4e-16 * n_terms relative of the exact
    sum of |terms|, and within the float-partial bound of the reference-order sum."""
    rng = np.random.default_rng(41)
    full = shape + ((channels,) if channels > 1 else ())
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        src = rng.integers(max(info.min, -30000), min(info.max, 60000) + 1, full).astype(dtype)
    else:
        src = (rng.random(full) - 0.3).astype(dtype)
    S = ocip.Image(src, channels=channels)
    say("sum", dtype.__name__, shape, channels)
    res = (ctypes.c_double * channels)()
    for sqr in (False, True):
        fn = ocip.lib.ocipSumSqr if sqr else ocip.lib.ocipSum
        ocip.check(fn(None, S.handle, res))
        got = np.array(res[:channels])
        want = orc.sum_exact(src, sqr)
        if np.dtype(dtype).itemsize <= 2 and np.issubdtype(dtype, np.integer):
            np.testing.assert_array_equal(got, want)
            continue
        x = src.astype(np.float64).reshape(-1, channels)
        mag = np.abs(x * x if sqr else x).sum(axis=0)
        assert np.all(np.abs(got - want) <= 4e-16 * 64 * mag + 1e-300), (got, want)
        if dtype == np.float32:
            ref = orc.sum_ref(src, sqr)
            assert np.all(np.abs(got - ref) <= 26 * 2.0 ** -24 * mag + 1e-300)


def test_sum_program_and_errors(ocip):
    src = np.arange(12, dtype=np.float32).reshape(3, 4)
    S = ocip.Image(src)
    prog = ctypes.c_void_p()
    ocip.check(ocip.lib.ocipPrepareStatistics(ctypes.byref(prog), S.handle))
    r = ctypes.c_double()
    ocip.check(ocip.lib.ocipSum(prog, S.handle, ctypes.byref(r)))
    assert r.value == 66.0
    ocip.check(ocip.lib.ocipSumSqr(prog, S.handle, ctypes.byref(r)))
    assert r.value == float((np.arange(12) ** 2).sum())
    assert ocip.lib.ocipSum(prog, S.handle, None) == ocip.CL_INVALID_VALUE
    ocip.check(ocip.lib.ocipReleaseProgram(prog))


# ----------------------------------------------------------------------------- ABI behaviour
def test_context_and_images(ocip):
    name = ctypes.create_string_buffer(256)
    ocip.check(ocip.lib.ocipGetDeviceName(name, 256))
    assert b"gfx950" in name.value
    assert ocip.lib.ocipGetDeviceName(ctypes.create_string_buffer(2), 2) == ocip.CL_OUT_OF_RESOURCES
    ocip.check(ocip.lib.ocipFinish())
    # device-only images: Send and Read are no-ops
    D = ocip.Image(shape=(4, 4), dtype=np.float32, device_only=True)
    ocip.check(ocip.lib.ocipSendImage(D.handle))
    ocip.check(ocip.lib.ocipReadImage(D.handle))
    h = ctypes.c_void_p()
    assert ocip.lib.ocipCreateImage(ctypes.byref(h), ocip.SImage(4, 4, 16, 5, ocip.F32), None,
                                    ocip.CL_MEM_READ_WRITE) == ocip.CL_IMAGE_FORMAT_NOT_SUPPORTED


def test_proximity_argument_errors(ocip):
    S = ocip.Image(np.zeros((32, 32), np.float32))
    T = ocip.Image(np.zeros((4, 4), np.float32))
    big = ocip.Image(np.zeros((40, 4), np.float32))
    T8 = ocip.Image(np.zeros((4, 4), np.uint8))
    D = ocip.Image(np.zeros((32, 32), np.float32))
    D64 = ocip.Image(np.zeros((32, 32), np.float64))
    Dsmall = ocip.Image(np.zeros((31, 32), np.float32))
    L = ocip.lib
    assert L.ocipCrossCorr(S.handle, big.handle, D.handle) == ocip.CL_INVALID_VALUE
    assert L.ocipCrossCorr(S.handle, T8.handle, D.handle) == ocip.CL_INVALID_VALUE
    assert L.ocipCrossCorr(S.handle, T.handle, Dsmall.handle) == ocip.CL_INVALID_IMAGE_SIZE
    assert L.ocipCrossCorr(S.handle, T.handle, D64.handle) == ocip.CL_INVALID_VALUE
    assert L.ocipCrossCorr(S.handle, T.handle, None) == ocip.CL_INVALID_MEM_OBJECT
    ocip.check(L.ocipPrepareProximity(S.handle))


def test_send_is_lazy_and_explicit(ocip):
    """A source never sent is sent on first use; later host changes need ocipSendImage."""
    from openclipp_amd import ocip as mod
    src = np.ones((16, 16), np.float32)
    tpl = np.ones((1, 1), np.float32)
    S, T = mod.Image(src), mod.Image(tpl)
    D = mod.Image(shape=(16, 16), dtype=np.float32)
    mod.check(mod.proximity("CrossCorr", S, T, D))
    assert D.read()[3, 3] == 1.0
    src[:] = 2.0
    mod.check(mod.proximity("CrossCorr", S, T, D))
    assert D.read()[3, 3] == 1.0          # device copy unchanged
    S.send()
    mod.check(mod.proximity("CrossCorr", S, T, D))
    assert D.read()[3, 3] == 2.0

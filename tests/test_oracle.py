"""Pins the CPU oracle (oracle/oracle.c) against checks that do not reuse its code.

* numpy sliding-window brute force (tests/brute.py) and scipy.signal.correlate2d
  for the window sums, including the centring anchor of even template sizes;
* exact integer agreement where float32 sums are exact;
* error bounds of float32 accumulation against the FP64 oracle;
* numpy.cumsum for integrals, 4-corner window sums against brute force;
* invariants: NCC in [-1, 1], NCC = 1 and SSD = 0 at an exact crop.
"""
import numpy as np
import pytest
import scipy.signal

from oracle import oracle as orc
from tests import brute

OPS = [orc.SSD, orc.SSD_NORM, orc.SAD, orc.CC, orc.CC_NORM]
INT_TYPES = [np.uint8, np.int8, np.uint16, np.int16, np.uint32, np.int32]


def rand_img(rng, h, w, dtype, c=None, lo=None, hi=None):
    shape = (h, w) if c is None else (h, w, c)
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        lo = info.min if lo is None else lo
        hi = info.max if hi is None else hi
        return rng.integers(lo, hi, size=shape, endpoint=True).astype(dtype)
    return rng.random(shape).astype(dtype)


@pytest.mark.parametrize("tw,th", [(1, 1), (2, 2), (3, 2), (4, 5), (7, 7), (16, 16), (17, 9)])
def test_f64_matches_bruteforce(tw, th):
    rng = np.random.default_rng(100 + tw * 31 + th)
    src = rand_img(rng, 23, 29, np.float32)
    tpl = rand_img(rng, th, tw, np.float32)
    want = brute.maps_f64(src, tpl)
    for op in OPS:
        got = orc.prox_f64(op, src, tpl)
        np.testing.assert_allclose(got, want[orc.OP_NAMES[op]], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("tw,th", [(4, 4), (5, 3), (6, 8)])
def test_cc_matches_scipy_correlate2d(tw, th):
    """correlate2d(mode='full') has the template's (0,0) on source (i-th+1, j-tw+1);
    the reference's centred output D(x,y) is full[y - th/2 + th - 1, x - tw/2 + tw - 1]."""
    rng = np.random.default_rng(7 + tw + th)
    src = rng.random((19, 21))
    tpl = rng.random((th, tw))
    full = scipy.signal.correlate2d(src, tpl, mode="full", boundary="fill")
    H, W = src.shape
    oy, ox = th - 1 - th // 2, tw - 1 - tw // 2
    want = full[oy:oy + H, ox:ox + W]
    got = orc.prox_f64(orc.CC, src, tpl)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    # the wrong anchor (rounding -tw/2 the other way) is caught for even sizes
    if tw % 2 == 0:
        wrong = full[oy:oy + H, ox + 1:ox + 1 + W] if ox + 1 + W <= full.shape[1] else None
        if wrong is not None:
            assert not np.allclose(got, wrong)


@pytest.mark.parametrize("dtype", INT_TYPES)
def test_i64_matches_bruteforce(dtype):
    rng = np.random.default_rng(5)
    src = rand_img(rng, 17, 13, dtype)
    tpl = rand_img(rng, 5, 6, dtype)
    want = brute.maps_i64(src, tpl)
    np.testing.assert_array_equal(orc.prox_i64(orc.I64_SSD, src, tpl), want["SqrDistance"])
    np.testing.assert_array_equal(orc.prox_i64(orc.I64_SAD, src, tpl), want["AbsDistance"])
    np.testing.assert_array_equal(orc.prox_i64(orc.I64_CC, src, tpl), want["CrossCorr"])
    np.testing.assert_array_equal(orc.prox_i64(orc.I64_SS, src, tpl), want["SS"])


@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.int16, np.float32, np.float64])
@pytest.mark.parametrize("contract", [True, False])
def test_ref_exact_when_sums_representable(dtype, contract):
    """Small integer-valued inputs: every float32 partial sum is exact, so the
    reference-order map equals the exact integer map whatever the order."""
    rng = np.random.default_rng(11)
    lo, hi = (0, 15) if dtype != np.int8 and dtype != np.int16 else (-15, 15)
    src = rng.integers(lo, hi, size=(21, 18), endpoint=True).astype(dtype)
    tpl = rng.integers(lo, hi, size=(6, 5), endpoint=True).astype(dtype)
    want = brute.maps_i64(src.astype(np.int64), tpl.astype(np.int64))
    for op, name in [(orc.SSD, "SqrDistance"), (orc.SAD, "AbsDistance"), (orc.CC, "CrossCorr")]:
        got = orc.prox_ref(op, src, tpl, contract=contract)
        np.testing.assert_array_equal(got.astype(np.int64), want[name])


@pytest.mark.parametrize("contract", [True, False])
@pytest.mark.parametrize("tw,th", [(16, 16), (31, 31), (8, 24)])
def test_ref_f32_within_accumulation_bound(contract, tw, th):
    """float32 serial accumulation of n non-negative terms: |err| <= n * 2**-24 * sum|terms|."""
    rng = np.random.default_rng(tw * th)
    src = rand_img(rng, 40, 37, np.float32)
    tpl = rand_img(rng, th, tw, np.float32)
    n = tw * th
    f64 = brute.maps_f64(src, tpl)
    for op, name in [(orc.SSD, "SqrDistance"), (orc.SAD, "AbsDistance"), (orc.CC, "CrossCorr")]:
        got = orc.prox_ref(op, src, tpl, contract=contract).astype(np.float64)
        bound = (n + 2) * 2.0 ** -24 * f64[name] + 1e-30
        assert np.all(np.abs(got - f64[name]) <= bound), name
    for op, name in [(orc.SSD_NORM, "SqrDistance_Norm"), (orc.CC_NORM, "CrossCorr_Norm")]:
        got = orc.prox_ref(op, src, tpl, contract=contract).astype(np.float64)
        np.testing.assert_allclose(got, f64[name], rtol=3 * (n + 4) * 2.0 ** -24)


def test_contract_flag_changes_rounding():
    rng = np.random.default_rng(3)
    src = rand_img(rng, 32, 32, np.float32)
    tpl = rand_img(rng, 16, 16, np.float32)
    a = orc.prox_ref(orc.CC, src, tpl, contract=True)
    b = orc.prox_ref(orc.CC, src, tpl, contract=False)
    assert not np.array_equal(a, b)


@pytest.mark.parametrize("nch", [2, 4])
def test_multichannel_is_per_channel(nch):
    """NBCHAN vector maths is channel-independent (cl-files/Buffers.h:76-84)."""
    rng = np.random.default_rng(nch)
    src = rand_img(rng, 15, 14, np.float32, c=nch)
    tpl = rand_img(rng, 5, 4, np.float32, c=nch)
    for op in OPS:
        got = orc.prox_ref(op, src, tpl)
        for c in range(nch):
            one = orc.prox_ref(op, np.ascontiguousarray(src[:, :, c]),
                               np.ascontiguousarray(tpl[:, :, c]))
            np.testing.assert_array_equal(got[:, :, c], one)


def test_pitched_template_is_read_by_step():
    """Bench templates are ROIs of the source with the source's pitch
    (Bench/src/benchImageProximity.hpp:83-86)."""
    rng = np.random.default_rng(9)
    src = rand_img(rng, 30, 30, np.float32)
    roi = src[10:26, 10:26]                       # view: pitch = 30 floats
    assert roi.strides[0] == 30 * 4
    got = orc.prox_ref(orc.CC_NORM, src, roi)
    ref = orc.prox_ref(orc.CC_NORM, src, np.ascontiguousarray(roi))
    np.testing.assert_array_equal(got, ref)


def test_match_invariants():
    rng = np.random.default_rng(21)
    src = rand_img(rng, 48, 52, np.float32)
    for tw, th in [(16, 16), (9, 12)]:
        ox, oy = 13, 20
        tpl = np.ascontiguousarray(src[oy:oy + th, ox:ox + tw])
        x, y = ox + tw // 2, oy + th // 2
        ncc = orc.prox_f64(orc.CC_NORM, src, tpl)
        assert np.nanmax(ncc) <= 1 + 1e-12 and np.nanmin(ncc) >= -1 - 1e-12
        assert abs(ncc[y, x] - 1) < 1e-12
        assert np.unravel_index(np.nanargmax(ncc), ncc.shape) == (y, x)
        assert orc.prox_f64(orc.SSD, src, tpl)[y, x] == 0
        assert orc.prox_f64(orc.SSD_NORM, src, tpl)[y, x] == 0
        ref = orc.prox_ref(orc.CC_NORM, src, tpl)
        assert abs(ref[y, x] - 1) < 1e-6


def test_norm_nan_on_zero_window():
    """0 / sqrt(0) is NaN, x / 0 is inf: no guard in the reference."""
    src = np.zeros((10, 10), np.float32)
    src[5, 5] = 1
    tpl = np.ones((3, 3), np.float32)
    ncc = orc.prox_ref(orc.CC_NORM, src, tpl)
    assert np.isnan(ncc[0, 0]) and ncc[5, 5] == np.float32(1 / 3)


# ----------------------------------------------------------------------------- integral

@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint32, np.float32, np.float64])
@pytest.mark.parametrize("sqr", [False, True])
def test_integral_exact_matches_cumsum(dtype, sqr):
    rng = np.random.default_rng(2)
    src = rand_img(rng, 37, 70, dtype)
    got = orc.integral_exact(src, sqr)
    if np.issubdtype(dtype, np.integer):
        # Python integers: U32 squares overflow int64 sums
        a = src.astype(object)
        a = a * a if sqr else a
        want = np.cumsum(np.cumsum(a, axis=0), axis=1)
        want = np.array([[float(v) for v in r] for r in want], dtype=np.float64)
        np.testing.assert_array_equal(got, want)
    else:
        np.testing.assert_allclose(got, brute.integral_f64(src, sqr), rtol=1e-13)


@pytest.mark.parametrize("shape", [(8, 64), (16, 128), (19, 77), (1, 1), (9, 65), (64, 3)])
@pytest.mark.parametrize("sqr", [False, True])
def test_integral_ref_exact_for_small_integers(shape, sqr):
    """U8 source: reference F64 scan is exact; the F32 scan is exact while sums < 2**24."""
    rng = np.random.default_rng(shape[0] * 100 + shape[1])
    src = rng.integers(0, 255, size=shape, endpoint=True).astype(np.uint8)
    want = brute.integral_i64(src, sqr).astype(np.float64)
    np.testing.assert_array_equal(orc.integral_ref(src, sqr, np.float64), want)
    if want.max() < 2 ** 24:
        np.testing.assert_array_equal(orc.integral_ref(src, sqr, np.float32), want)


def test_integral_ref_f32_rounding_bound():
    rng = np.random.default_rng(4)
    src = rng.random((200, 300)).astype(np.float32)
    for sqr in (False, True):
        want = brute.integral_f64(src, sqr)
        got = orc.integral_ref(src, sqr, np.float32).astype(np.float64)
        # depth of the reference's summation tree is < 8 + 8 + 8 + 5 + 25 adds
        assert np.all(np.abs(got - want) <= 60 * 2.0 ** -24 * want + 1e-30)


def test_area_sum_matches_window_brute_force():
    rng = np.random.default_rng(8)
    src = rng.random((25, 31))
    I = brute.integral_f64(src, sqr=True)
    for w, h in [(1, 1), (4, 6), (7, 3), (31, 25), (16, 16)]:
        got = orc.area_sum(I, w, h)
        want = (brute.windows(src, w, h) ** 2).sum(axis=(-2, -1))
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-10)


def _bitrev8(v):
    r = np.zeros_like(v)
    for b in range(8):
        r |= ((v >> b) & 1) << (7 - b)
    return r


def _min_max_indx_closed_form(src, want_max):
    """The reference's MinIndx/MaxIndx tie order in closed form (NaN-free images), derived
    from cl-files/Statistics.cl:344-425 without simulating its tree: a halving tree in which
    position p takes p + k only when strictly better settles two tied work-items at the level
    of the lowest bit in which their lids differ, and the one with a 0 there wins; so among
    the work-items holding the group's extreme the winner has the smallest bit-reversed lid
    (lid = 16 ly + lx), and inside a work-item the smallest i (column 256 gxg + lx + 16 i).
    Groups compare their extremes as float32 and the lowest group id wins ties
    (StatisticsHelpers.cpp:105-145)."""
    a = np.asarray(src)
    H, W = a.shape
    ngx = -(-W // 256)
    ys, xs = np.mgrid[0:H, 0:W]
    gid = (ys // 16) * ngx + xs // 256
    lid = 16 * (ys % 16) + xs % 16
    i = (xs % 256) // 16
    val = a.astype(np.float64)                       # exact for every element type
    best_group = {}
    for g in np.unique(gid):
        m = gid == g
        ext = val[m].max() if want_max else val[m].min()
        cand = m & (val == ext)
        key = _bitrev8(lid[cand]) * 16 + i[cand]
        k = np.argmin(key)
        best_group[int(g)] = (np.float32(ext), int(xs[cand][k]), int(ys[cand][k]))
    parts = [best_group[g] for g in sorted(best_group)]
    fl = np.array([p[0] for p in parts], np.float32)
    ext = fl.max() if want_max else fl.min()
    g = int(np.flatnonzero(fl == ext)[0])
    return float(parts[g][0]), parts[g][1], parts[g][2]


def test_min_max_indx_reference_tie_order_known_answers():
    """Hand-derived from the reference's kernel: pixel (16, 0) is work-item 0's i = 1 and
    (1, 0) is work-item 1's i = 0; the tree's last step keeps position 0 -> (16, 0).  (0, 0)
    starts at position 0 and wins every tie.  (5, 1) is lid 21 and (6, 0) lid 6: 21 = 10101b
    and 6 = 00110b first differ in bit 0, where 6 has the 0 -> (6, 0).  Different groups: the
    lower group id wins, (300, 0) (group 1) over (0, 16) (group ngx = 2)."""
    def case(shape, pts, want_max=True, dtype=np.uint8):
        img = np.zeros(shape, dtype) if want_max else np.full(shape, 9, dtype)
        for x, y in pts:
            img[y, x] = 9 if want_max else 0
        return orc.min_max_indx(img, want_max)[1:]
    assert case((4, 40), [(1, 0), (16, 0)]) == (16, 0)
    assert case((4, 40), [(1, 0), (16, 0)], want_max=False) == (16, 0)
    assert case((20, 300), [(0, 0), (17, 3), (299, 19)]) == (0, 0)
    assert case((4, 40), [(5, 1), (6, 0)]) == (6, 0)
    assert case((32, 512), [(0, 16), (300, 0)]) == (300, 0)
    # no ties: the unique extreme
    rng = np.random.default_rng(12)
    src = rng.random((33, 47)).astype(np.float32)
    v, x, y = orc.min_max_indx(src, True)
    k = np.argmax(src)
    assert (y, x) == np.unravel_index(k, src.shape) and v == src.flat[k]


@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16, np.int16, np.uint32,
                                   np.int32, np.float32, np.float64])
def test_min_max_indx_matches_closed_form(dtype):
    """Tie-heavy images (few distinct values) of ragged sizes: the oracle's tree simulation
    against the closed form above.  Wide 32-bit and F64 values also exercise the float
    partials: values that differ below float precision tie between groups."""
    rng = np.random.default_rng(40 + np.dtype(dtype).itemsize)
    for shape in [(1, 1), (3, 17), (16, 256), (17, 257), (45, 530), (70, 300)]:
        if dtype in (np.uint32, np.int32):
            a = (2 ** 30 + rng.integers(0, 4, shape) * 16).astype(dtype)   # ties after rounding
        elif dtype == np.float64:
            a = 1.0 + rng.integers(0, 3, shape) * 2.0 ** -40
        elif dtype == np.float32:
            a = rng.integers(0, 3, shape).astype(np.float32) * np.float32(0.25)
        else:
            info = np.iinfo(dtype)
            a = rng.integers(max(info.min, -2), min(info.max, 3), shape).astype(dtype)
        a = a.astype(dtype)
        for want_max in (True, False):
            assert orc.min_max_indx(a, want_max) == _min_max_indx_closed_form(a, want_max), \
                (shape, want_max)


def test_min_max_indx_nan_follows_the_comparisons():
    """A NaN is never taken; a NaN held by work-item 0's first pixel is never replaced, so the
    result is NaN at (0, 0)."""
    a = np.ones((20, 40), np.float32)
    a[3, 7] = 5.0
    a[1, 1] = np.nan
    assert orc.min_max_indx(a, True) == (5.0, 7, 3)
    a[0, 0] = np.nan
    v, x, y = orc.min_max_indx(a, True)
    assert np.isnan(v) and (x, y) == (0, 0)


def test_sum_ref_exact_where_float_partials_are_exact():
    """Every group partial of these images is an integer below 2^24, so the reference-order
    sum must equal the exact integer sum: a dropped, repeated or misplaced pixel (a wrong
    stride, group size or tree step) changes it."""
    rng = np.random.default_rng(31)
    for shape in [(1, 1), (16, 256), (17, 257), (70, 300), (33, 517)]:
        a = rng.integers(0, 256, shape).astype(np.uint8)
        assert orc.sum_ref(a)[0] == int(a.astype(np.int64).sum())
        b = rng.integers(0, 16, shape).astype(np.uint8)
        assert orc.sum_ref(b, True)[0] == int((b.astype(np.int64) ** 2).sum())
    # channels are independent
    c = rng.integers(0, 256, (20, 300, 4)).astype(np.uint8)
    want = [int(c[:, :, k].astype(np.int64).sum()) for k in range(4)]
    assert list(orc.sum_ref(c)) == want


def test_sum_ref_float_within_accumulation_bound():
    """F32 sums: within the float accumulation bound of the exact sum (16 serial + 8 tree
    additions per group in float), and the rounding really is float (not exact)."""
    rng = np.random.default_rng(32)
    a = (rng.random((100, 700)) - 0.3).astype(np.float32)
    for sqr in (False, True):
        exact = orc.sum_exact(a, sqr)[0]
        x = a.astype(np.float64)
        mag = float(np.abs(x * x if sqr else x).sum())
        got = orc.sum_ref(a, sqr)[0]
        assert abs(got - exact) <= 26 * 2.0 ** -24 * mag
        assert got != exact


def test_sum_exact_against_fsum():
    rng = np.random.default_rng(33)
    a = rng.random((50, 60))
    import math
    assert orc.sum_exact(a)[0] == math.fsum(a.ravel())
    assert orc.sum_exact(a.astype(np.float32), True)[0] == math.fsum(
        float(t) ** 2 for t in a.astype(np.float32).ravel())
    assert orc.sum_exact(np.array([[3, -4]], np.int16), True)[0] == 25.0


def test_cpu_baseline_is_the_reference_order():
    """The benchmark's CPU baseline (oracle/baseline.c, vectorised across adjacent outputs)
    computes the reference's CrossCorr_Norm bit for bit: the same map as the reference-order
    restatement, on ragged sizes, a 1x1 template and windows over every border."""
    rng = np.random.default_rng(71)
    for (H, W), (th, tw) in [((50, 70), (7, 9)), ((130, 200), (16, 16)), ((97, 301), (64, 64)),
                             ((33, 65), (1, 1)), ((20, 129), (20, 12))]:
        src = rng.random((H, W)).astype(np.float32)
        tpl = rng.random((th, tw)).astype(np.float32)
        got, _ = orc.baseline_cc_norm(src, tpl)
        np.testing.assert_array_equal(got, orc.prox_ref(orc.CC_NORM, src, tpl, contract=True))
        band, _ = orc.baseline_cc_norm(src, tpl, rows=(3, H - 2))
        np.testing.assert_array_equal(band[3:H - 2], got[3:H - 2])

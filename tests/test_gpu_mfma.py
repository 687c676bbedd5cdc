"""GPU checks of the MFMA cross-correlation (ocipCrossCorrEx / ocipCrossCorr_NormEx).

Against the FP64 oracle with bounds that follow from the operand precision:
  BF16      |err| <= (2^-7  + acc) * sum|s*t|   (8-bit significands: u = 2^-8 per operand)
  FP16      |err| <= (2^-10 + acc) * sum|s*t|   (11-bit significands: u = 2^-11 per operand)
  BF16x3    |err| <= (2^-14 + acc) * sum|s*t|   (hi + lo leaves 2^-16 per operand, and the
                                                 dropped lo*lo term is below 2^-16)
with acc = 2 (n + 16) 2^-24 for the FP32 accumulation of n products.  Norm maps use the same
bound divided by the exact denominator.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import integral_kappa, run_op, say

pytestmark = pytest.mark.gpu

PREC = {"bf16": 1, "f16": 2, "bf16x3": 3}
UNIT = {"bf16": 2.0 ** -7 + 2.0 ** -16, "f16": 2.0 ** -10 + 2.0 ** -22, "bf16x3": 2.0 ** -14}


@pytest.fixture(scope="module")
def ocip():
    from openclipp_amd import ocip as mod
    ctx = mod.initialize()
    yield mod
    mod.lib.ocipUninitialize(ctx)


def bound(prec, src, tpl):
    n = tpl.size
    sabs = orc.prox_f64(orc.CC, np.abs(src).astype(np.float64), np.abs(tpl).astype(np.float64))
    return (UNIT[prec] + 2 * (n + 16) * 2.0 ** -24) * sabs + 1e-30


TEMPLATES = [(1, 1), (3, 5), (16, 16), (31, 31), (64, 64), (96, 40), (17, 64), (100, 9), (130, 3),
             (132, 5), (260, 3)]


@pytest.mark.parametrize("prec", list(PREC))
@pytest.mark.parametrize("tw,th", TEMPLATES)
def test_crosscorr_bounds(ocip, prec, tw, th):
    rng = np.random.default_rng(tw * 131 + th)
    src = rng.random((150, 260), dtype=np.float32) - np.float32(0.3)
    tpl = rng.random((th, tw), dtype=np.float32) - np.float32(0.3)
    say("mfma", prec, tw, th)
    f64 = orc.prox_f64(orc.CC, src, tpl)
    b = bound(prec, src, tpl)
    got = run_op(ocip, "CrossCorr", src, tpl, precision=PREC[prec]).astype(np.float64)
    err = np.abs(got - f64)
    assert np.all(err <= b), (float(np.max(err / b)), prec)
    ncc64 = orc.prox_f64(orc.CC_NORM, src, tpl)
    s64, t64 = src.astype(np.float64), tpl.astype(np.float64)
    den = np.sqrt(orc.prox_f64(orc.CC, s64 * s64, np.ones_like(t64)) * (t64 * t64).sum())
    got = run_op(ocip, "CrossCorr_Norm", src, tpl, precision=PREC[prec]).astype(np.float64)
    ok = np.isfinite(ncc64)
    cond = 16 * 2.0 ** -53 * integral_kappa(src, tpl) * np.abs(ncc64)
    assert np.all(np.abs(got - ncc64)[ok] <= (b / den + cond)[ok] * 1.01 + 1e-7)


@pytest.mark.parametrize("prec", list(PREC))
def test_image_edges_and_odd_sizes(ocip, prec):
    """Image sizes that are not multiples of the 128x128 tile."""
    rng = np.random.default_rng(2)
    for H, W in [(1, 1), (7, 300), (129, 129), (257, 65)]:
        src = rng.random((H, W), dtype=np.float32)
        tpl = rng.random((min(H, 9), min(W, 12)), dtype=np.float32)
        f64 = orc.prox_f64(orc.CC, src, tpl)
        got = run_op(ocip, "CrossCorr", src, tpl, precision=PREC[prec]).astype(np.float64)
        assert np.all(np.abs(got - f64) <= bound(prec, src, tpl)), (H, W)
        ncc64 = orc.prox_f64(orc.CC_NORM, src, tpl)
        got = run_op(ocip, "CrossCorr_Norm", src, tpl, precision=PREC[prec]).astype(np.float64)
        ok = np.isfinite(ncc64)
        tol = 2 * UNIT[prec] + 2e-6 + 16 * 2.0 ** -53 * integral_kappa(src, tpl)
        assert np.all((np.abs(got - ncc64) <= tol * (1 + np.abs(ncc64)))[ok]), (H, W)


@pytest.mark.parametrize("prec", ["bf16", "bf16x3"])
def test_tall_template(ocip, prec):
    """A template taller than one band of the denominator pass (128 output rows)."""
    rng = np.random.default_rng(11)
    src = rng.random((330, 70), dtype=np.float32)
    tpl = rng.random((310, 5), dtype=np.float32)
    f64 = orc.prox_f64(orc.CC, src, tpl)
    got = run_op(ocip, "CrossCorr", src, tpl, precision=PREC[prec]).astype(np.float64)
    assert np.all(np.abs(got - f64) <= bound(prec, src, tpl))
    ncc64 = orc.prox_f64(orc.CC_NORM, src, tpl)
    got = run_op(ocip, "CrossCorr_Norm", src, tpl, precision=PREC[prec]).astype(np.float64)
    assert np.max(np.abs(got - ncc64)) <= 2 * UNIT[prec] + 1e-6


@pytest.mark.parametrize("prec", ["bf16", "bf16x3"])
def test_norm_zero_windows(ocip, prec):
    """All-zero windows give 0/0 = NaN exactly where the FP64 evaluation does (the
    denominators' column and row sums are differences of two prefixes of one sequence, so an
    all-zero window sums to exactly 0), and windows overlapping the zero block stay within
    the bound."""
    rng = np.random.default_rng(21)
    src = rng.random((160, 200), dtype=np.float32)
    src[30:110, 40:150] = 0
    tpl = rng.random((17, 24), dtype=np.float32)
    ncc64 = orc.prox_f64(orc.CC_NORM, src, tpl)
    got = run_op(ocip, "CrossCorr_Norm", src, tpl, precision=PREC[prec]).astype(np.float64)
    assert np.isnan(ncc64).sum() > 1000
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ncc64))
    ok = np.isfinite(ncc64)
    assert np.max(np.abs(got - ncc64)[ok]) <= 2 * UNIT[prec] + 1e-6


@pytest.mark.parametrize("W,tw", [(460, 388), (461, 388), (301, 132)])
def test_norm_wide_windows(ocip, W, tw):
    """The denominator pass with window segments of 8 (388 columns) and 4 (132) chunks of 64
    columns, on even and odd image widths."""
    rng = np.random.default_rng(W + tw)
    src = rng.random((40, W), dtype=np.float32)
    tpl = rng.random((3, tw), dtype=np.float32)
    say("wide windows", W, tw)
    ncc64 = orc.prox_f64(orc.CC_NORM, src, tpl)
    got = run_op(ocip, "CrossCorr_Norm", src, tpl, precision=1).astype(np.float64)
    assert np.max(np.abs(got - ncc64)) <= 2 * UNIT["bf16"] + 1e-6


def test_norm_of_an_image_over_one_gib(ocip):
    """A pitched image whose rows span more than 2^30 bytes takes the denominator pass's
    address-select loads instead of its buffer loads; both give the same maps."""
    rng = np.random.default_rng(29)
    big = np.zeros((4100, 65536), np.float32)     # 1.07 GB of rows, 64 columns used
    src = big[:, :64]
    src[...] = rng.random((4100, 64), dtype=np.float32)
    tpl = rng.random((9, 11), dtype=np.float32)
    ncc64 = orc.prox_f64(orc.CC_NORM, np.ascontiguousarray(src), tpl)
    got = run_op(ocip, "CrossCorr_Norm", src, tpl, precision=1).astype(np.float64)
    assert np.max(np.abs(got - ncc64)) <= 2 * UNIT["bf16"] + 1e-6
    small = run_op(ocip, "CrossCorr_Norm", np.ascontiguousarray(src), tpl, precision=1)
    np.testing.assert_allclose(got, small, rtol=1e-6, atol=0)


def test_wide_template_reads_the_integral(ocip):
    """Templates wider than 513 columns take the integral-based denominator pass."""
    rng = np.random.default_rng(23)
    src = rng.random((40, 620), dtype=np.float32)
    tpl = rng.random((3, 530), dtype=np.float32)
    ncc64 = orc.prox_f64(orc.CC_NORM, src, tpl)
    got = run_op(ocip, "CrossCorr_Norm", src, tpl, precision=1).astype(np.float64)
    tol = 2 * UNIT["bf16"] + 2e-6 + 16 * 2.0 ** -53 * integral_kappa(src, tpl)
    assert np.all(np.abs(got - ncc64) <= tol * (1 + np.abs(ncc64)))


# crosscorr_mfma NT = 2, 4; crosscorr_mfma_pair NT = 4, 6; crosscorr_mfma_strip (where its ring
# schedule fits the template; elsewhere the autotuner's choice)
VARIANTS = [2, 4, 14, 16, 24]
VARIANT_CASES = [((150, 260), (1, 1)), ((150, 260), (64, 64)), ((300, 200), (31, 31)),
                 ((131, 97), (97, 5)), ((260, 140), (3, 130)), ((129, 129), (16, 17)),
                 ((40, 620), (530, 3)), ((7, 300), (12, 7))]


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("prec", ["bf16", "f16", "bf16x3"])
def test_kernel_variants(ocip, monkeypatch, prec, variant):
    """Every tile variant the autotuner may pick, forced through OCIP_MFMA_VARIANT: CC within
    the operand-precision bound of the FP64 oracle, NCC within the bound over the exact
    denominator, and an exact crop giving NCC = 1 at its centre."""
    monkeypatch.setenv("OCIP_MFMA_VARIANT", str(variant))
    for k, ((H, W), (tw, th)) in enumerate(VARIANT_CASES):
        rng = np.random.default_rng(100 + k)
        src = rng.random((H, W), dtype=np.float32) - np.float32(0.25)
        tpl = rng.random((th, tw), dtype=np.float32) - np.float32(0.25)
        say("variant", variant, prec, (H, W), (tw, th))
        f64 = orc.prox_f64(orc.CC, src, tpl)
        b = bound(prec, src, tpl)
        got = run_op(ocip, "CrossCorr", src, tpl, precision=PREC[prec]).astype(np.float64)
        assert np.all(np.abs(got - f64) <= b), ((H, W), (tw, th), float(np.max(np.abs(got - f64) / b)))
        ncc64 = orc.prox_f64(orc.CC_NORM, src, tpl)
        s64, t64 = src.astype(np.float64), tpl.astype(np.float64)
        den = np.sqrt(orc.prox_f64(orc.CC, s64 * s64, np.ones_like(t64)) * (t64 * t64).sum())
        got = run_op(ocip, "CrossCorr_Norm", src, tpl, precision=PREC[prec]).astype(np.float64)
        ok = np.isfinite(ncc64)
        cond = 16 * 2.0 ** -53 * integral_kappa(src, tpl) * np.abs(ncc64) + 2.0 ** -22
        assert np.all(np.abs(got - ncc64)[ok] <= (b / den + cond)[ok] * 1.01 + 1e-7), ((H, W), (tw, th))
    # exact crop: NCC = 1 (to the operand rounding) at the crop's centre
    rng = np.random.default_rng(7)
    src = rng.random((300, 330), dtype=np.float32) + np.float32(0.1)
    tpl = np.ascontiguousarray(src[100:164, 150:214])
    got = run_op(ocip, "CrossCorr_Norm", src, tpl, precision=PREC[prec])
    assert np.unravel_index(np.nanargmax(got), got.shape) == (132, 182)
    assert abs(float(got[132, 182]) - 1.0) <= 2 * UNIT[prec]


@pytest.mark.parametrize("prec", ["bf16", "f16", "bf16x3"])
def test_row_pair_variants_are_bit_identical(ocip, monkeypatch, prec):
    """The row-pair variants (14, 16: crosscorr_mfma_pair with 4 or 6 n-tiles; 24: the strip
    kernel, one operand plane, so not for split BF16) add each output's products in the same
    order (template-row pairs in turn, chunks in turn, then the two half sums; for split BF16
    the three products of a chunk in turn), so their maps are equal bit for bit, for CrossCorr
    and for CrossCorr_Norm (the same in-place scale)."""
    rng = np.random.default_rng(77)
    src = rng.random((300, 330), dtype=np.float32)
    tpl = rng.random((40, 52), dtype=np.float32)      # strip schedule fits: 20 steps, KC2 = 5
    variants = (14, 16) if prec == "bf16x3" else (14, 16, 24)
    maps = {}
    for v in variants:
        monkeypatch.setenv("OCIP_MFMA_VARIANT", str(v))
        say("row-pair identity", prec, v)
        maps[v] = (run_op(ocip, "CrossCorr", src, tpl, precision=PREC[prec]),
                   run_op(ocip, "CrossCorr_Norm", src, tpl, precision=PREC[prec]))
    for v in variants[1:]:
        np.testing.assert_array_equal(maps[v][0], maps[14][0])
        np.testing.assert_array_equal(maps[v][1], maps[14][1])


@pytest.mark.parametrize("wgs", [1, 3, 7, 0])
def test_strip_runs(ocip, monkeypatch, wgs):
    """crosscorr_mfma_strip with few workgroups, so that each takes long runs of tiles down its
    strips and across strip boundaries (0: the default count); CC and NCC against FP64."""
    monkeypatch.setenv("OCIP_MFMA_VARIANT", "24")
    if wgs:
        monkeypatch.setenv("OCIP_STRIP_WORKGROUPS", str(wgs))
    rng = np.random.default_rng(40 + wgs)
    for (H, W), (tw, th) in [((700, 150), (20, 24)), ((520, 64), (64, 64)), ((389, 200), (33, 17))]:
        src = rng.random((H, W), dtype=np.float32) - np.float32(0.2)
        tpl = rng.random((th, tw), dtype=np.float32) - np.float32(0.2)
        say("strip", wgs, (H, W), (tw, th))
        f64 = orc.prox_f64(orc.CC, src, tpl)
        got = run_op(ocip, "CrossCorr", src, tpl, precision=1).astype(np.float64)
        assert np.all(np.abs(got - f64) <= bound("bf16", src, tpl)), ((H, W), (tw, th))
        ncc64 = orc.prox_f64(orc.CC_NORM, src, tpl)
        got = run_op(ocip, "CrossCorr_Norm", src, tpl, precision=1).astype(np.float64)
        ok = np.isfinite(ncc64)
        assert np.max(np.abs(got - ncc64)[ok]) <= 2 * UNIT["bf16"] + 1e-6


def test_f32_precision_is_the_plain_path(ocip):
    rng = np.random.default_rng(5)
    src = rng.random((70, 90), dtype=np.float32)
    tpl = rng.random((11, 13), dtype=np.float32)
    a = run_op(ocip, "CrossCorr", src, tpl, precision=0)
    np.testing.assert_array_equal(a, orc.prox_ref(orc.CC, src, tpl))


def test_precision_argument_errors(ocip):
    S = ocip.Image(np.zeros((32, 32), np.uint8))
    T = ocip.Image(np.zeros((4, 4), np.uint8))
    D = ocip.Image(np.zeros((32, 32), np.float32))
    assert ocip.lib.ocipCrossCorrEx(S.handle, T.handle, D.handle, 1) == ocip.CL_INVALID_VALUE
    S2 = ocip.Image(np.zeros((32, 32), np.float32))
    T2 = ocip.Image(np.zeros((4, 4), np.float32))
    assert ocip.lib.ocipCrossCorrEx(S2.handle, T2.handle, D.handle, 9) == ocip.CL_INVALID_VALUE


@pytest.mark.parametrize("prec,limit", [("bf16", 1e-3), ("f16", 2.5e-4), ("bf16x3", 2e-5)])
def test_precision_4096_ncc_error(ocip, prec, limit):
    """The 4096x4096 F32 N(0.5, 0.2) image with a 64x64 crop x1.1: max |NCC - NCC_fp64| over
    the match location and 4096 sampled pixels, against the stated limits (
    1e-3 for BF16)."""
    from openclipp_amd import workloads
    src, tpl, (ox, oy) = workloads.precision_4096()
    say("precision_4096", prec)
    got = run_op(ocip, "CrossCorr_Norm", src, tpl, precision=PREC[prec])
    rng = np.random.default_rng(0)
    ys = np.concatenate([rng.integers(0, 4096, 4095), [oy + 32]])
    xs = np.concatenate([rng.integers(0, 4096, 4095), [ox + 32]])
    P = np.pad(src.astype(np.float64), [(32, 31), (32, 31)])
    t = tpl.astype(np.float64)
    win = np.stack([P[y:y + 64, x:x + 64] for x, y in zip(xs, ys)])
    want = (win * t).sum(axis=(1, 2)) / np.sqrt((win * win).sum(axis=(1, 2)) * (t * t).sum())
    err = np.abs(got[ys, xs] - want)
    say(prec, "max", err.max(), "mean", err.mean())
    assert err.max() <= limit
    assert np.unravel_index(np.nanargmax(got), got.shape) == (oy + 32, ox + 32)

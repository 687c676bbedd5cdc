"""Batched template matching (ocipCrossCorr_NormBatch, ocipSqrDistance_NormBatch).

* Each map of a batch equals the single-template call at the same precision to 1e-6
  relative: the contraction is the same tile code (the per-pixel accumulation order does not
  depend on the tile width the autotuner picks); the single MFMA call forms 1 / sqrt(SS TT)
  from the source image and the batch from the FP64 integral, which differ by float
  rounding.  The F32 batch runs the single-template plain path and is bit-identical.  Every
  map is within the precision bound of the FP64 oracle.
* The batch is recorded into a HIP graph on its first call; later calls replay it.  Replays
  are checked one by one, with host work, new template contents sent between replays, and
  synchronisations in between.
* The 2048x2048 / 256-template workload, CrossCorr_Norm and SqrDistance_Norm: sampled pixels
  against FP64 and the crop location of every 8th template is the arg-max (arg-min) of its
  map.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import integral_kappa, run_op, say

pytestmark = pytest.mark.gpu

PREC = {"f32": 0, "bf16": 1, "f16": 2, "bf16x3": 3}
UNIT = {"f32": 2.0 ** -22, "bf16": 2.0 ** -7 + 2.0 ** -16, "f16": 2.0 ** -10 + 2.0 ** -22,
        "bf16x3": 2.0 ** -14}


@pytest.fixture(scope="module")
def ocip():
    from openclipp_amd import ocip as mod
    ctx = mod.initialize()
    yield mod
    mod.lib.ocipUninitialize(ctx)


def run_batch(ocip, fn, src, tpls, prec, S=None, Ts=None, Ds=None):
    S = S or ocip.Image(src)
    Ts = Ts or [ocip.Image(t) for t in tpls]
    Ds = Ds or [ocip.Image(shape=src.shape, dtype=np.float32) for _ in tpls]
    ocip.check(fn(S.handle, ocip.handles(Ts), ocip.handles(Ds), len(Ts), PREC[prec]), "batch")
    return [d.read().copy() for d in Ds], (S, Ts, Ds)


@pytest.fixture(autouse=True)
def single_calls_use_the_batch_tile_code(monkeypatch, request):
    """The single-template calls compared here run the tile code of the batch kernel at the
    same precision, so that the per-pixel accumulation order is the same: the row-pair tile
    (variant 16) for BF16 and FP16, crosscorr_mfma (variant 4) for split BF16."""
    callspec = getattr(request.node, "callspec", None)
    prec = callspec.params.get("prec") if callspec else None
    monkeypatch.setenv("OCIP_MFMA_VARIANT", "4" if prec == "bf16x3" else "16")


def same_map(batch, single, exact):
    """Batch map against the single-template call: identical NaN positions and values equal
    (exact) or within 1e-6 relative (the denominators' rounding)."""
    if exact:
        np.testing.assert_array_equal(batch, single)
        return
    np.testing.assert_array_equal(np.isnan(batch), np.isnan(single))
    ok = ~np.isnan(single)
    b, s = batch[ok].astype(np.float64), single[ok].astype(np.float64)
    assert np.all(np.abs(b - s) <= 1e-6 * np.abs(s) + 1e-30), float(np.max(np.abs(b - s)))


def make_case(seed, H=190, W=270, n=10, lo=3, hi=40):
    rng = np.random.default_rng(seed)
    src = rng.random((H, W), dtype=np.float32)
    tpls = []
    for _ in range(n):
        tw, th = int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1))
        ox, oy = int(rng.integers(0, W - tw + 1)), int(rng.integers(0, H - th + 1))
        tpls.append((src[oy:oy + th, ox:ox + tw] + rng.normal(0, 0.02, (th, tw))).astype(np.float32))
    return src, tpls


@pytest.mark.parametrize("prec", ["bf16", "f16", "bf16x3", "f32"])
def test_ncc_batch_equals_single_calls(ocip, prec):
    src, tpls = make_case(1)
    say("ncc batch", prec)
    got, _ = run_batch(ocip, ocip.lib.ocipCrossCorr_NormBatch, src, tpls, prec)
    for t, g in zip(tpls, got):
        single = run_op(ocip, "CrossCorr_Norm", src, t,
                        precision=None if prec == "f32" else PREC[prec])
        same_map(g, single, prec == "f32")
        f64 = orc.prox_f64(orc.CC_NORM, src, t)
        kappa = integral_kappa(src, t)
        bound = (UNIT[prec] + 2 * (t.size + 16) * 2.0 ** -24 + 16 * 2.0 ** -53 * kappa) * 1.01
        assert np.all(np.abs(g - f64) <= bound * np.abs(f64) + 1e-6)


@pytest.mark.parametrize("prec", ["bf16", "f16", "bf16x3", "f32"])
def test_ssd_norm_batch(ocip, prec):
    """SqrDistance_Norm: F32 runs the direct (s - t)^2 path; the MFMA precisions expand
    SS - 2 CC + TT, whose error is that of CC times 2 over the denominator."""
    src, tpls = make_case(2, n=6)
    got, _ = run_batch(ocip, ocip.lib.ocipSqrDistance_NormBatch, src, tpls, prec)
    for t, g in zip(tpls, got):
        f64 = orc.prox_f64(orc.SSD_NORM, src, t)
        if prec == "f32":
            np.testing.assert_array_equal(g, run_op(ocip, "SqrDistance_Norm", src, t))
            continue
        s64, t64 = src.astype(np.float64), t.astype(np.float64)
        ss = orc.prox_f64(orc.CC, s64 * s64, np.ones_like(t64))
        tt = (t64 * t64).sum()
        sabs = orc.prox_f64(orc.CC, np.abs(s64), np.abs(t64))
        bound = (2 * (UNIT[prec] + 2 * (t.size + 16) * 2.0 ** -24) * sabs + 2.0 ** -40 * (ss + tt)) \
            / np.sqrt(ss * tt)
        assert np.all(np.abs(g - f64) <= bound * 1.01 + 1e-7)


@pytest.mark.parametrize("prec", ["bf16", "f16", "bf16x3"])
def test_batch_wide_templates(ocip, prec):
    """Templates wider than 97 columns (processed in 64-column blocks) mixed with narrow ones
    in one batch, both ops: each map equals the single-template call to 1e-6 relative."""
    rng = np.random.default_rng(21)
    H, W = 150, 300
    src = rng.random((H, W), dtype=np.float32)
    tpls = [rng.random((th, tw), dtype=np.float32)
            for tw, th in ((100, 9), (131, 20), (12, 33), (98, 5), (64, 64))]
    for op, fn in (("CrossCorr_Norm", ocip.lib.ocipCrossCorr_NormBatch),
                   ("SqrDistance_Norm", ocip.lib.ocipSqrDistance_NormBatch)):
        say("wide batch", prec, op)
        got, _ = run_batch(ocip, fn, src, tpls, prec)
        for t, g in zip(tpls, got):
            if op == "CrossCorr_Norm":
                same_map(g, run_op(ocip, op, src, t, precision=PREC[prec]), False)
            else:
                # the single SqrDistance_Norm call has no MFMA form: against FP64
                f64 = orc.prox_f64(orc.SSD_NORM, src, t)
                s64, t64 = src.astype(np.float64), t.astype(np.float64)
                ss = orc.prox_f64(orc.CC, s64 * s64, np.ones_like(t64))
                tt = (t64 * t64).sum()
                sabs = orc.prox_f64(orc.CC, np.abs(s64), np.abs(t64))
                bound = (2 * (UNIT[prec] + 2 * (t.size + 16) * 2.0 ** -24) * sabs +
                         2.0 ** -40 * (ss + tt)) / np.sqrt(ss * tt)
                assert np.all(np.abs(g - f64) <= bound * 1.01 + 1e-7)


def test_graph_replays_track_new_inputs(ocip):
    """Four calls with the same images: the first records the graph, the next three replay
    it.  Between calls the host rewrites the template data and sends it; every replay must
    match an eager single-template call on the current data."""
    src, tpls = make_case(3, n=5)
    S = ocip.Image(src)
    Ts = [ocip.Image(t) for t in tpls]
    Ds = [ocip.Image(shape=src.shape, dtype=np.float32) for _ in tpls]
    rng = np.random.default_rng(9)
    for rep in range(4):
        say("replay", rep)
        got, _ = run_batch(ocip, ocip.lib.ocipCrossCorr_NormBatch, src, tpls, "bf16", S, Ts, Ds)
        for t, g in zip(tpls, got):
            same_map(g, run_op(ocip, "CrossCorr_Norm", src, t, precision=1), False)
        # host work between replays: new template contents, sent explicitly
        for T in Ts:
            T.arr[...] = rng.random(T.arr.shape, dtype=np.float32)
            T.send()
        ocip.check(ocip.lib.ocipFinish())


def test_batch_argument_errors(ocip):
    src = np.zeros((64, 64), np.float32)
    S = ocip.Image(src)
    T = ocip.Image(np.ones((5, 5), np.float32))
    D = ocip.Image(np.zeros((64, 64), np.float32))
    Dbad = ocip.Image(np.zeros((64, 63), np.float32))
    T8 = ocip.Image(np.ones((5, 5), np.uint8))
    L = ocip.lib
    assert L.ocipCrossCorr_NormBatch(S.handle, ocip.handles([T]), ocip.handles([Dbad]), 1, 1) == ocip.CL_INVALID_IMAGE_SIZE
    assert L.ocipCrossCorr_NormBatch(S.handle, ocip.handles([T8]), ocip.handles([D]), 1, 1) == ocip.CL_INVALID_VALUE
    assert L.ocipCrossCorr_NormBatch(S.handle, ocip.handles([T]), ocip.handles([D]), 1, 7) == ocip.CL_INVALID_VALUE
    assert L.ocipCrossCorr_NormBatch(S.handle, None, None, 0, 1) == ocip.CL_SUCCESS
    S8 = ocip.Image(np.zeros((64, 64), np.uint8))
    assert L.ocipSqrDistance_NormBatch(S8.handle, ocip.handles([T8]), ocip.handles([D]), 1, 0) == ocip.CL_INVALID_VALUE


def _batch_2048_check(got, tpls, offs, ssd):
    """Every 8th template: the crop is the arg-max (NCC) / arg-min (SSD_Norm) of its map, and
    63 random pixels plus the crop's are within tolerance of the FP64 value: 1e-3 for NCC,
    2e-3 for SSD_Norm = (SS - 2 CC + TT) / sqrt(SS TT), whose only inexact term is twice the
    NCC's.  Returns the largest error seen."""
    from openclipp_amd import workloads
    src = workloads.batch_2048()[0]
    rng = np.random.default_rng(0)
    P = np.pad(src.astype(np.float64), 96)
    worst = 0.0
    for k in range(0, len(tpls), 8):
        t, (ox, oy), g = tpls[k], offs[k], got[k]
        th, tw = t.shape
        y, x = oy + th // 2, ox + tw // 2
        at = np.nanargmin(g) if ssd else np.nanargmax(g)
        assert np.unravel_index(at, g.shape) == (y, x), k
        ys = np.concatenate([rng.integers(0, 2048, 63), [y]])
        xs = np.concatenate([rng.integers(0, 2048, 63), [x]])
        t64 = t.astype(np.float64)
        win = np.stack([P[yy + 96 - th // 2: yy + 96 - th // 2 + th, xx + 96 - tw // 2: xx + 96 - tw // 2 + tw]
                        for xx, yy in zip(xs, ys)])
        den = np.sqrt((win * win).sum(axis=(1, 2)) * (t64 * t64).sum())
        if ssd:
            want = ((win - t64) ** 2).sum(axis=(1, 2)) / den
        else:
            want = (win * t64).sum(axis=(1, 2)) / den
        worst = max(worst, float(np.max(np.abs(g[ys, xs] - want))))
    assert worst <= (2e-3 if ssd else 1e-3)
    return worst


def test_batch_2048_256_templates(ocip):
    from openclipp_amd import workloads
    src, tpls, offs = workloads.batch_2048()
    say("batch_2048", len(tpls))
    got, _ = run_batch(ocip, ocip.lib.ocipCrossCorr_NormBatch, src, tpls, "bf16")
    say("batch_2048 max abs NCC error", _batch_2048_check(got, tpls, offs, ssd=False))


def test_batch_2048_256_templates_ssd_norm(ocip):
    from openclipp_amd import workloads
    src, tpls, offs = workloads.batch_2048()
    say("batch_2048 SqrDistance_Norm", len(tpls))
    got, _ = run_batch(ocip, ocip.lib.ocipSqrDistance_NormBatch, src, tpls, "bf16")
    say("batch_2048 max abs SSD_Norm error", _batch_2048_check(got, tpls, offs, ssd=True))

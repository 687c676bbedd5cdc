"""Seeded random shapes through ocipCrossCorrEx / ocipCrossCorr_NormEx against the FP64 oracle
(same bounds as test_gpu_mfma.py): odd sizes, thin images, templates as large as the image,
bands and tiles cut at every kind of edge.  Shapes are drawn once from a fixed seed, so the
cases are the same on every run."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import integral_kappa, run_op, say
from tests.test_gpu_mfma import PREC, UNIT, bound

pytestmark = pytest.mark.gpu


def _cases(n=64, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        H, W = int(rng.integers(1, 400)), int(rng.integers(1, 400))
        th, tw = int(rng.integers(1, min(H, 140) + 1)), int(rng.integers(1, min(W, 140) + 1))
        if W * H * tw * th > 40_000_000:     # keeps each FP64 oracle map under ~0.1 s
            continue
        out.append((H, W, th, tw, str(rng.choice(list(PREC)))))
    return out


CASES = _cases()


@pytest.fixture(scope="module")
def ocip():
    from openclipp_amd import ocip as mod
    ctx = mod.initialize()
    yield mod
    mod.lib.ocipUninitialize(ctx)


@pytest.mark.parametrize("H,W,th,tw,prec", CASES)
def test_random_shapes(ocip, H, W, th, tw, prec):
    rng = np.random.default_rng(H * 7919 + W * 31 + th * 7 + tw)
    src = rng.random((H, W), dtype=np.float32) - np.float32(0.3)
    tpl = rng.random((th, tw), dtype=np.float32) - np.float32(0.3)
    say("random", H, W, th, tw, prec)
    f64 = orc.prox_f64(orc.CC, src, tpl)
    b = bound(prec, src, tpl)
    got = run_op(ocip, "CrossCorr", src, tpl, precision=PREC[prec]).astype(np.float64)
    assert np.all(np.abs(got - f64) <= b)
    ncc64 = orc.prox_f64(orc.CC_NORM, src, tpl)
    s64, t64 = src.astype(np.float64), tpl.astype(np.float64)
    den = np.sqrt(orc.prox_f64(orc.CC, s64 * s64, np.ones_like(t64)) * (t64 * t64).sum())
    got = run_op(ocip, "CrossCorr_Norm", src, tpl, precision=PREC[prec]).astype(np.float64)
    ok = np.isfinite(ncc64)
    np.testing.assert_array_equal(np.isfinite(got), ok)
    cond = 16 * 2.0 ** -53 * integral_kappa(src, tpl) * np.abs(ncc64)
    assert np.all(np.abs(got - ncc64)[ok] <= (b / den + cond)[ok] * 1.01 + 1e-7)

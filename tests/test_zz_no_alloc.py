"""ocipCrossCorr_NormExNoAlloc (the entry point that allocates nothing) against
ocipCrossCorr_NormEx, through the C ABI.

Two calls in a row, the second on a smaller image and template, each after the allocating
entry point has run on the same shape (so both use the same tuned kernel variant): the maps
must be bit-identical.  Then the rejections: F32 precision, a template wider than 513 columns,
and one whose fragments need more than the 8 MiB scratch array.
"""
import numpy as np
import pytest

from tests.gpu_util import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ocip():
    from openclipp_amd import ocip as mod
    ctx = mod.initialize()
    yield mod
    mod.lib.ocipUninitialize(ctx)


def run(ocip, fn, src, tpl, prec):
    S, T = ocip.Image(src), ocip.Image(tpl)
    D = ocip.Image(shape=src.shape, dtype=np.float32)
    code = fn(S.handle, T.handle, D.handle, prec)
    out = D.read().copy() if code == ocip.CL_SUCCESS else None
    for i in (S, T, D):
        i.release()
    return code, out


@pytest.mark.parametrize("prec", [1, 2, 3])
def test_no_alloc_equals_allocating_entry_point(ocip, prec):
    rng = np.random.default_rng(50 + prec)
    # the second call is on a smaller image and template than the first
    for (H, W), (th, tw) in [((700, 900), (64, 64)), ((300, 260), (31, 17))]:
        src = rng.random((H, W), dtype=np.float32)
        y0, x0 = 100, 120
        tpl = (src[y0:y0 + th, x0:x0 + tw] * np.float32(1.1)).copy()
        say("no-alloc", prec, (H, W), (th, tw))
        code, want = run(ocip, ocip.lib.ocipCrossCorr_NormEx, src, tpl, prec)
        ocip.check(code, "ocipCrossCorr_NormEx")
        code, got = run(ocip, ocip.lib.ocipCrossCorr_NormExNoAlloc, src, tpl, prec)
        ocip.check(code, "ocipCrossCorr_NormExNoAlloc")
        np.testing.assert_array_equal(got, want)
        # the crop is found
        assert np.unravel_index(np.nanargmax(got), got.shape) == (y0 + th // 2, x0 + tw // 2)


def test_no_alloc_rejections(ocip):
    rng = np.random.default_rng(60)
    src = rng.random((300, 600), dtype=np.float32)
    f = ocip.lib.ocipCrossCorr_NormExNoAlloc
    assert run(ocip, f, src, src[:16, :16].copy(), 0)[0] == ocip.CL_INVALID_VALUE
    assert run(ocip, f, src, src[:8, :520].copy(), 1)[0] == ocip.CL_INVALID_VALUE
    # 500 x 200 BF16: 8 column blocks x 200 rows x 6 chunks of 1 KiB fragments > 8 MiB
    assert run(ocip, f, src, src[:200, :500].copy(), 1)[0] == ocip.CL_INVALID_BUFFER_SIZE

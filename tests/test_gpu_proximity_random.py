"""Seeded random shapes through the five proximity ops of the plain path: F32 maps bit-exact
against the reference-order oracle (with FMA contraction), 8/16-bit sources exact against the
int64 oracle rounded to float.  Shapes are drawn once from a fixed seed."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import OP_CODE, run_op, say

pytestmark = pytest.mark.gpu

RAW_OPS = ["SqrDistance", "AbsDistance", "CrossCorr"]
DTYPES = [np.float32, np.uint8, np.int16]


def _cases(n=40, seed=77):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        H, W = int(rng.integers(1, 300)), int(rng.integers(1, 300))
        th, tw = int(rng.integers(1, min(H, 80) + 1)), int(rng.integers(1, min(W, 80) + 1))
        if W * H * tw * th > 20_000_000:
            continue
        out.append((H, W, th, tw, int(rng.integers(0, len(DTYPES)))))
    return out


@pytest.fixture(scope="module")
def ocip():
    from openclipp_amd import ocip as mod
    ctx = mod.initialize()
    yield mod
    mod.lib.ocipUninitialize(ctx)


@pytest.mark.parametrize("H,W,th,tw,d", _cases())
def test_random_shapes(ocip, H, W, th, tw, d):
    dtype = DTYPES[d]
    rng = np.random.default_rng(H * 131 + W * 17 + th * 5 + tw + d)
    if dtype == np.float32:
        src = rng.random((H, W), dtype=np.float32)
        tpl = rng.random((th, tw), dtype=np.float32)
    else:
        info = np.iinfo(dtype)
        src = rng.integers(info.min, info.max, (H, W), endpoint=True).astype(dtype)
        tpl = rng.integers(info.min, info.max, (th, tw), endpoint=True).astype(dtype)
    for op in RAW_OPS:
        say("plain random", np.dtype(dtype).name, op, H, W, th, tw)
        got = run_op(ocip, op, src, tpl)
        if dtype == np.float32:
            want = orc.prox_ref(OP_CODE[op], src, tpl, contract=True)
        else:
            want = orc.f64_from_i64(OP_CODE[op], src, tpl).astype(np.float32)
        np.testing.assert_array_equal(got, want, err_msg=op)

"""Shared pieces of the GPU tests: one library context per test session, helpers that run
an op through the C ABI on numpy inputs, and oracle comparisons on row bands."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import oracle as orc

OP_CODE = {"SqrDistance": orc.SSD, "SqrDistance_Norm": orc.SSD_NORM, "AbsDistance": orc.SAD,
           "CrossCorr": orc.CC, "CrossCorr_Norm": orc.CC_NORM}


def has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def real_dtype(arr):
    return np.float64 if arr.dtype == np.float64 else np.float32


def run_op(ocip, op, src, tpl, precision=None):
    """Dest map of ocip<op>(src, tpl) read back to the host."""
    S = ocip.Image(src)
    T = ocip.Image(tpl)
    D = ocip.Image(shape=src.shape[:2], dtype=real_dtype(src),
                   channels=1 if src.ndim == 2 else src.shape[2])
    if precision is None:
        ocip.check(ocip.proximity(op, S, T, D), op)
    else:
        fn = {"CrossCorr": ocip.lib.ocipCrossCorrEx,
              "CrossCorr_Norm": ocip.lib.ocipCrossCorr_NormEx}[op]
        ocip.check(fn(S.handle, T.handle, D.handle, precision), op)
    out = D.read().copy()
    for i in (S, T, D):
        i.release()
    return out


def band_oracle(fn, src, tpl, y0, y1):
    """Apply an oracle map function to output rows [y0, y1) only, using a crop of the source
    that holds every row those windows touch (the rest of the image is irrelevant)."""
    th = tpl.shape[0]
    cy = th // 2
    H = src.shape[0]
    r0 = max(0, y0 - cy)
    r1 = min(H, y1 - cy + th - 1)
    crop = src[r0:r1]
    # a crop starting at r0 > 0 would treat rows < r0 as zero padding; rows of the band only
    # need rows >= y0 - cy >= r0, and rows < r1, so the band is exact.
    out = fn(crop)
    return out[y0 - r0:y1 - r0]


def integral_kappa(src, tpl):
    """Conditioning of a Norm denominator taken from the FP64 integral: total image energy
    over window energy.  The 4-corner window sum has relative error ~eps64 * kappa."""
    s = src.astype(np.float64)
    ss = orc.prox_f64(orc.CC, s * s, np.ones(tpl.shape[:2], np.float64))
    with np.errstate(divide="ignore"):
        return (s * s).sum() / ss

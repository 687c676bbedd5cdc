"""Independent brute-force definitions used to pin the oracle.

Nothing here shares code with oracle/oracle.c: window sums are formed with
numpy's sliding_window_view over an explicitly zero-padded source, integrals
with numpy.cumsum.  The centring convention is the reference's: the template
origin sits at (-tw/2, -th/2) with C integer division
(cl-files/ImageProximity.cl:52-53), pixels outside the source read as 0
(cl-files/ImageProximity.cl:64-67).
"""
import numpy as np
from numpy.lib.stride_tricks import sliding_window_view


def padded(src, tw, th):
    """Zero-pad so that window (y, x) of size th x tw is S[y-th/2 : ..., x-tw/2 : ...]."""
    cx, cy = tw // 2, th // 2
    pad = [(cy, th - 1 - cy), (cx, tw - 1 - cx)] + [(0, 0)] * (src.ndim - 2)
    return np.pad(src, pad)


def windows(src, tw, th):
    """Array of shape (H, W, [C,] th, tw) of the zero-padded windows."""
    P = padded(src, tw, th)
    v = sliding_window_view(P, (th, tw), axis=(0, 1))
    return v  # (H, W, [C,] th, tw)


def _tmpl(tpl):
    # (th, tw, [C]) -> ([C,] th, tw) to line up with the window axes
    return np.moveaxis(tpl, (0, 1), (-2, -1)) if tpl.ndim == 3 else tpl


def maps_f64(src, tpl):
    """Dict of the five proximity maps in float64 (exact for integer inputs < 2**53)."""
    th, tw = tpl.shape[:2]
    win = windows(src.astype(np.float64), tw, th)
    t = _tmpl(tpl.astype(np.float64))
    cc = (win * t).sum(axis=(-2, -1))
    ssd = ((win - t) ** 2).sum(axis=(-2, -1))
    sad = np.abs(win - t).sum(axis=(-2, -1))
    ss = (win * win).sum(axis=(-2, -1))
    tt = (t * t).sum(axis=(-2, -1))
    with np.errstate(divide="ignore", invalid="ignore"):
        den = np.sqrt(ss * tt)
        return {"SqrDistance": ssd, "AbsDistance": sad, "CrossCorr": cc,
                "SqrDistance_Norm": ssd / den, "CrossCorr_Norm": cc / den, "SS": ss}


def maps_i64(src, tpl):
    """Exact integer SSD, SAD, CC, SS for integer inputs."""
    th, tw = tpl.shape[:2]
    win = windows(src.astype(np.int64), tw, th)
    t = _tmpl(tpl.astype(np.int64))
    return {"SqrDistance": ((win - t) ** 2).sum(axis=(-2, -1)),
            "AbsDistance": np.abs(win - t).sum(axis=(-2, -1)),
            "CrossCorr": (win * t).sum(axis=(-2, -1)),
            "SS": (win * win).sum(axis=(-2, -1))}


def integral_i64(src, sqr=False):
    a = src.astype(np.int64)
    if sqr:
        a = a * a
    return np.cumsum(np.cumsum(a, axis=0), axis=1)


def integral_f64(src, sqr=False):
    a = src.astype(np.float64)
    if sqr:
        a = a * a
    return np.cumsum(np.cumsum(a, axis=0), axis=1)

"""Python access to the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker.  The product library
(openclipp_amd) never imports it.

All images are numpy arrays of shape (H, W) or (H, W, C) in one of the eight
SImage element types; pitched images are passed as views (the row stride in
bytes is read from the array).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_BASE_PATH = os.path.join(_HERE, "_build", "libbaseline.so")

# SImage::EDataType (reference include/SImage.h:36-46)
U8, S8, U16, S16, U32, S32, F32, F64 = range(8)
NP_OF_TYPE = {U8: np.uint8, S8: np.int8, U16: np.uint16, S16: np.int16,
              U32: np.uint32, S32: np.int32, F32: np.float32, F64: np.float64}
TYPE_OF_NP = {np.dtype(v): k for k, v in NP_OF_TYPE.items()}

SSD, SSD_NORM, SAD, CC, CC_NORM = range(5)
OP_NAMES = {SSD: "SqrDistance", SSD_NORM: "SqrDistance_Norm", SAD: "AbsDistance",
            CC: "CrossCorr", CC_NORM: "CrossCorr_Norm"}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i32 = ctypes.c_void_p, ctypes.c_int
        L.orc_prox_ref.argtypes = [i32, i32, i32, vp, i32, i32, i32, vp, i32, i32, i32,
                                   vp, i32, i32, i32, i32, i32]
        L.orc_prox_ref_rows.argtypes = [i32, i32, i32, vp, i32, i32, i32, vp, i32, i32, i32,
                                        vp, i32, i32, vp, i32, i32]
        L.orc_prox_f64.argtypes = [i32, i32, i32, vp, i32, i32, i32, vp, i32, i32, i32, vp, i32]
        L.orc_prox_i64.argtypes = [i32, i32, i32, vp, i32, i32, i32, vp, i32, i32, i32, vp, i32]
        L.orc_integral_ref.argtypes = [i32, i32, vp, i32, i32, i32, i32, vp, i32, i32]
        L.orc_integral_exact.argtypes = [i32, i32, vp, i32, i32, i32, vp]
        L.orc_area_sum.argtypes = [vp, i32, i32, i32, i32, vp]
        L.orc_min_max_indx.argtypes = [i32, i32, vp, i32, i32, i32,
                                       ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        for f in ("orc_prox_ref", "orc_prox_ref_rows", "orc_prox_f64", "orc_prox_i64", "orc_integral_ref",
                  "orc_integral_exact", "orc_area_sum", "orc_min_max_indx"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


_base = None


def baseline_lib():
    """The CPU baseline (oracle/baseline.c): timing only, never the checker."""
    global _base
    if _base is None:
        if not os.path.exists(_BASE_PATH):
            build()
        L = ctypes.CDLL(_BASE_PATH)
        vp, i32 = ctypes.c_void_p, ctypes.c_int
        L.base_cc_norm.argtypes = [vp, i32, i32, i32, vp, i32, i32, i32, vp, i32, i32, i32, i32]
        L.base_cc_norm.restype = i32
        L.base_isa.argtypes = []
        L.base_isa.restype = ctypes.c_char_p
        _base = L
    return _base


def baseline_cc_norm(src, tpl, rows=None, nthreads=0, out=None):
    """CrossCorr_Norm of a 1-channel F32 image by the CPU baseline (rows=(y0, y1) restricts
    the output rows computed).  Returns (map, instruction set used)."""
    sp, W, H, C, ss, t = _desc(src)
    tp, tw, th, tc, ts, tt = _desc(tpl)
    assert C == 1 and tc == 1 and t == F32 and tt == F32
    if out is None:
        out = np.zeros((H, W), np.float32)
    y0, y1 = (0, H) if rows is None else rows
    r = baseline_lib().base_cc_norm(sp, W, H, ss, tp, tw, th, ts, out.ctypes.data,
                                    out.strides[0], y0, y1, nthreads)
    assert r == 0
    return out, baseline_lib().base_isa().decode()


def _desc(a):
    """(pointer, width, height, channels, step_bytes, type) of an image array."""
    if a.ndim == 2:
        h, w = a.shape
        c = 1
        if a.strides[1] != a.itemsize:
            raise ValueError("rows must be contiguous")
    elif a.ndim == 3:
        h, w, c = a.shape
        if a.strides[2] != a.itemsize or a.strides[1] != a.itemsize * c:
            raise ValueError("pixels must be contiguous")
    else:
        raise ValueError("image must be 2-D or 3-D")
    return a.ctypes.data, w, h, c, a.strides[0], TYPE_OF_NP[a.dtype]


def prox_ref(op, src, tpl, contract=True, rows=None, nthreads=0):
    """Reference-order proximity map (REAL precision: float32, float64 for F64 sources).

    rows=(y0, y1) restricts the rows computed (others are left 0)."""
    sp, W, H, C, ss, t = _desc(src)
    tp, tw, th, tc, ts, tt = _desc(tpl)
    assert tc == C and tt == t
    real = np.float64 if t == F64 else np.float32
    out = np.zeros((H, W, C) if src.ndim == 3 else (H, W), dtype=real)
    y0, y1 = (0, H) if rows is None else rows
    r = lib().orc_prox_ref(op, t, C, sp, W, H, ss, tp, tw, th, ts, out.ctypes.data,
                           out.strides[0], 1 if contract else 0, y0, y1, nthreads)
    assert r == 0
    return out


def prox_ref_rows(op, src, tpl, rows, contract=True, nthreads=0):
    """prox_ref on the listed output rows only: returns an array of shape (len(rows), W[, C])."""
    sp, W, H, C, ss, t = _desc(src)
    tp, tw, th, tc, ts, tt = _desc(tpl)
    assert tc == C and tt == t
    real = np.float64 if t == F64 else np.float32
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    out = np.zeros((H, W, C) if src.ndim == 3 else (H, W), dtype=real)
    r = lib().orc_prox_ref_rows(op, t, C, sp, W, H, ss, tp, tw, th, ts, out.ctypes.data,
                                out.strides[0], 1 if contract else 0, rows.ctypes.data,
                                len(rows), nthreads)
    assert r == 0
    return out[rows]


def prox_f64(op, src, tpl, nthreads=0):
    """Same window sums accumulated in FP64 (ground truth for float sources)."""
    sp, W, H, C, ss, t = _desc(src)
    tp, tw, th, tc, ts, tt = _desc(tpl)
    assert tc == C and tt == t
    out = np.zeros((H, W, C) if src.ndim == 3 else (H, W), dtype=np.float64)
    r = lib().orc_prox_f64(op, t, C, sp, W, H, ss, tp, tw, th, ts, out.ctypes.data, nthreads)
    assert r == 0
    return out


I64_SSD, I64_SAD, I64_CC, I64_SS, I64_TT = range(5)


def prox_i64(which, src, tpl, nthreads=0):
    """Exact int64 window sums for integer sources (which: I64_*)."""
    sp, W, H, C, ss, t = _desc(src)
    tp, tw, th, tc, ts, tt = _desc(tpl)
    assert tc == C and tt == t and t < F32
    out = np.zeros((H, W, C) if src.ndim == 3 else (H, W), dtype=np.int64)
    r = lib().orc_prox_i64(which, t, C, sp, W, H, ss, tp, tw, th, ts, out.ctypes.data, nthreads)
    assert r == 0
    return out


def integral_ref(src, sqr=False, dtype=np.float32, contract=True):
    """The reference's 5-launch integral scan in its summation order."""
    sp, W, H, C, ss, t = _desc(src)
    assert C == 1
    out = np.zeros((H, W), dtype=dtype)
    r = lib().orc_integral_ref(1 if sqr else 0, t, sp, W, H, ss, TYPE_OF_NP[np.dtype(dtype)],
                               out.ctypes.data, out.strides[0], 1 if contract else 0)
    assert r == 0
    return out


def integral_exact(src, sqr=False):
    """Exact inclusive 2-D prefix sum (or of squares), rounded once to float64."""
    sp, W, H, C, ss, t = _desc(src)
    assert C == 1
    out = np.zeros((H, W), dtype=np.float64)
    r = lib().orc_integral_exact(1 if sqr else 0, t, sp, W, H, ss, out.ctypes.data)
    assert r == 0
    return out


def area_sum(I, w, h):
    """Window sums of the centred w x h window from an inclusive float64 integral."""
    I = np.ascontiguousarray(I, dtype=np.float64)
    H, W = I.shape
    out = np.zeros_like(I)
    lib().orc_area_sum(I.ctypes.data, W, H, w, h, out.ctypes.data)
    return out


def min_max_indx(src, want_max):
    sp, W, H, C, ss, t = _desc(src)
    assert C == 1
    v = ctypes.c_double()
    x = ctypes.c_int()
    y = ctypes.c_int()
    lib().orc_min_max_indx(1 if want_max else 0, t, sp, W, H, ss,
                           ctypes.byref(v), ctypes.byref(x), ctypes.byref(y))
    return v.value, x.value, y.value


def sum_ref(src, sqr=False):
    """Statistics::Sum / SumSqr per channel in the reference's order, for 8-32-bit and F32
    sources (REAL = float).  cl-files/Statistics.cl:47-121 (REDUCE): work-item (lx, ly) of a
    16x16 work-group at (gxg, gyg) sums, in REAL, the pixels (256 gxg + lx + 16 i,
    16 gyg + ly), i = 0..15 inside the image (the first one, then the rest in order; SQR
    squares each in REAL, :202, :311); the 256 work-item sums (lid = 16 ly + lx) are reduced
    by a halving tree, lid += lid + stride when both hold pixels (DO_REDUCE, :28-34);
    :279-288 store_value keeps one float partial per group (gid = gyg * ngx + gxg);
    C++/programs/StatisticsHelpers.cpp:66-76 adds the partials in double in gid order."""
    a = np.asarray(src)
    if a.ndim == 2:
        a = a[:, :, None]
    H, W, C = a.shape
    assert a.dtype != np.float64
    ngx, ngy = -(-W // 256), -(-H // 16)
    P = np.zeros((ngy * 16, ngx * 256, C), np.float32)
    P[:H, :W] = a.astype(np.float32)
    if sqr:
        P = P * P
    valid = np.zeros((ngy * 16, ngx * 256), bool)
    valid[:H, :W] = True
    # column 256 gxg + 16 i + lx -> axes [gyg, ly, gxg, i, lx]
    Q = P.reshape(ngy, 16, ngx, 16, 16, C)
    V = valid.reshape(ngy, 16, ngx, 16, 16)
    res = Q[:, :, :, 0].copy()
    for i in range(1, 16):
        res = np.where(V[:, :, :, i, :, None], res + Q[:, :, :, i], res)
    nb = V.sum(axis=3)                                    # [gyg, ly, gxg, lx]
    buf = res.transpose(0, 2, 1, 3, 4).reshape(ngy, ngx, 256, C)
    cnt = nb.transpose(0, 2, 1, 3).reshape(ngy, ngx, 256)
    stride = 128
    while stride >= 1:
        both = (cnt[:, :, :stride] > 0) & (cnt[:, :, stride:2 * stride] > 0)
        buf[:, :, :stride] = np.where(both[..., None], buf[:, :, :stride] + buf[:, :, stride:2 * stride],
                                      buf[:, :, :stride])
        cnt[:, :, :stride] = np.where(both, cnt[:, :, :stride] + cnt[:, :, stride:2 * stride],
                                      cnt[:, :, :stride])
        stride //= 2
    part = buf[:, :, 0].reshape(ngy * ngx, C).astype(np.float64)
    out = part[0].copy()
    for g in range(1, ngy * ngx):
        out += part[g]
    return out


def sum_exact(src, sqr=False):
    """Exact per-channel sum (or sum of squares) rounded once to double."""
    import math
    from fractions import Fraction
    a = np.asarray(src)
    if a.ndim == 2:
        a = a[:, :, None]
    out = []
    for c in range(a.shape[2]):
        v = a[:, :, c].ravel()
        if np.issubdtype(v.dtype, np.integer):
            x = [int(t) for t in v]
            out.append(float(sum(t * t for t in x) if sqr else sum(x)))
        elif sqr and v.dtype == np.float32:
            # squares of float32 values are exact in double
            out.append(math.fsum(float(t) * float(t) for t in v))
        elif sqr:
            out.append(float(sum(Fraction(float(t)) ** 2 for t in v)))
        else:
            out.append(math.fsum(float(t) for t in v))
    return np.array(out)


def f64_from_i64(op, src, tpl):
    """Exact-integer map of `op` for an integer source, as float64 (Norm in FP64)."""
    if op == SSD:
        return prox_i64(I64_SSD, src, tpl).astype(np.float64)
    if op == SAD:
        return prox_i64(I64_SAD, src, tpl).astype(np.float64)
    if op == CC:
        return prox_i64(I64_CC, src, tpl).astype(np.float64)
    ss = prox_i64(I64_SS, src, tpl).astype(np.float64)
    tt = prox_i64(I64_TT, src, tpl).astype(np.float64)
    num = prox_i64(I64_SSD if op == SSD_NORM else I64_CC, src, tpl).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return num / np.sqrt(ss * tt)

"""Seeded synthetic inputs for the five template-matching workloads.

Each generator returns numpy arrays only; nothing here touches the GPU.
  plain_1024      1024x1024 F32 U[0,1), 16x16 crop + N(0, 0.01^2)
  integer_4096    4096x4096 U8 Gaussian-blurred (sigma 4 px) noise stretched to 0..255,
                  32x32 crop + U{-3..3} clamped
  precision_4096  4096x4096 F32 N(0.5, 0.2^2) clamped to [0,1], 64x64 crop x 1.1
  batch_2048      2048x2048 F32 U[0,1), 256 crops with sides U{8..96} + N(0, 0.02^2)
  diff_1024x16    16 x 1024x1024 F32 U[0,1), 31x31 N(0,1) template, N(0,1) weights
"""
import numpy as np


def _crop(rng, img, tw, th):
    H, W = img.shape[:2]
    ox = int(rng.integers(0, W - tw + 1))
    oy = int(rng.integers(0, H - th + 1))
    return img[oy:oy + th, ox:ox + tw].copy(), (ox, oy)


def plain_1024(seed=1, size=1024, tsize=16):
    rng = np.random.default_rng(seed)
    src = rng.random((size, size), dtype=np.float32)
    tpl, off = _crop(rng, src, tsize, tsize)
    tpl = (tpl + rng.normal(0, 0.01, tpl.shape)).astype(np.float32)
    return src, tpl, off


def blurred_u8(rng, H, W, sigma=4.0):
    import scipy.ndimage
    noise = rng.random((H, W))
    # radius 4 sigma, mirrored border
    field = scipy.ndimage.gaussian_filter(noise, sigma, mode="reflect", truncate=4.0)
    lo, hi = field.min(), field.max()
    return np.rint((field - lo) / (hi - lo) * 255.0).astype(np.uint8)


def integer_4096(seed=2, size=4096, tsize=32):
    rng = np.random.default_rng(seed)
    src = blurred_u8(rng, size, size)
    tpl, off = _crop(rng, src, tsize, tsize)
    noise = rng.integers(-3, 3, size=tpl.shape, endpoint=True)
    tpl = np.clip(tpl.astype(np.int32) + noise, 0, 255).astype(np.uint8)
    return src, tpl, off


def precision_4096(seed=3, size=4096, tsize=64):
    rng = np.random.default_rng(seed)
    src = np.clip(rng.normal(0.5, 0.2, (size, size)), 0.0, 1.0).astype(np.float32)
    tpl, off = _crop(rng, src, tsize, tsize)
    tpl = (tpl * np.float32(1.1)).astype(np.float32)
    return src, tpl, off


def batch_2048(seed=4, size=2048, count=256, lo=8, hi=96):
    rng = np.random.default_rng(seed)
    src = rng.random((size, size), dtype=np.float32)
    tpls, offs = [], []
    for _ in range(count):
        tw = int(rng.integers(lo, hi, endpoint=True))
        th = int(rng.integers(lo, hi, endpoint=True))
        t, off = _crop(rng, src, tw, th)
        tpls.append((t + rng.normal(0, 0.02, t.shape)).astype(np.float32))
        offs.append(off)
    return src, tpls, offs


def diff_1024x16(seed=5, size=1024, count=16, tsize=31):
    rng = np.random.default_rng(seed)
    srcs = rng.random((count, size, size), dtype=np.float32)
    tpl = rng.normal(0, 1, (tsize, tsize)).astype(np.float32)
    w_ncc = rng.normal(0, 1, (count, size, size)).astype(np.float32)
    w_ssd = rng.normal(0, 1, (count, size, size)).astype(np.float32)
    return srcs, tpl, w_ncc, w_ssd

"""Writes the golden fixtures in this folder: small seeded inputs with their expected maps
computed by numpy brute force (tests/brute.py: zero-padded sliding windows in float64 / int64,
numpy.cumsum), independently of the oracle.  The reference ships no golden vectors for this
path, so these are this build's own known answers.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests import brute  # noqa: E402

CASES = [
    # name, dtype, (H, W), (th, tw), seed
    ("f32_odd", np.float32, (13, 17), (5, 3), 1),
    ("f32_even", np.float32, (12, 16), (4, 6), 2),
    ("u8_16x16", np.uint8, (24, 20), (16, 16), 3),
    ("s16_mixed", np.int16, (9, 11), (3, 8), 4),
    ("f64_1x1", np.float64, (7, 5), (1, 1), 5),
]


def make(dtype, shape, tshape, seed):
    rng = np.random.default_rng(seed)
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        src = rng.integers(info.min, info.max, size=shape, endpoint=True).astype(dtype)
        tpl = rng.integers(info.min, info.max, size=tshape, endpoint=True).astype(dtype)
    else:
        src = rng.random(shape).astype(dtype)
        tpl = rng.random(tshape).astype(dtype)
    return src, tpl


def main():
    for name, dtype, shape, tshape, seed in CASES:
        src, tpl = make(dtype, shape, tshape, seed)
        m = brute.maps_f64(src, tpl)
        out = {"src": src, "tpl": tpl, "integral": brute.integral_f64(src),
               "sqr_integral": brute.integral_f64(src, sqr=True)}
        for k in ("SqrDistance", "AbsDistance", "CrossCorr", "SqrDistance_Norm", "CrossCorr_Norm"):
            out[k] = m[k]
        if np.issubdtype(dtype, np.integer):
            mi = brute.maps_i64(src, tpl)
            for k in ("SqrDistance", "AbsDistance", "CrossCorr"):
                out[k + "_i64"] = mi[k]
        np.savez(os.path.join(HERE, name + ".npz"), **out)
        print("wrote", name)


if __name__ == "__main__":
    main()

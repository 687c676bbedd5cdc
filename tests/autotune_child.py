"""Child process of tests/test_gpu_autotune.py: runs ocipCrossCorrEx for a few templates on a
seeded image with OCIP_AUTOTUNE_CACHE set by the parent (the library reads that file once per
process), and saves the maps to an .npz file.

usage: python tests/autotune_child.py OUT.npz
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

# (precision, template width, template height): 64x64 BF16 may tune to the strip kernel, a
# 128-wide template cannot run it; split BF16 cannot run the row-pair kernels
CASES = [(1, 64, 64), (1, 128, 64), (3, 64, 64)]


def inputs():
    rng = np.random.default_rng(77)
    src = rng.random((300, 400), dtype=np.float32)
    tpls = {(tw, th): rng.random((th, tw), dtype=np.float32) for _, tw, th in CASES}
    return src, tpls


def main(out):
    from openclipp_amd import ocip
    src, tpls = inputs()
    ctx = ocip.initialize()
    maps = {}
    try:
        for prec, tw, th in CASES:
            S, T = ocip.Image(src), ocip.Image(tpls[(tw, th)])
            D = ocip.Image(shape=src.shape, dtype=np.float32)
            ocip.check(ocip.lib.ocipCrossCorrEx(S.handle, T.handle, D.handle, prec),
                       "ocipCrossCorrEx")
            maps[f"p{prec}_{tw}x{th}"] = D.read().copy()
            for i in (S, T, D):
                i.release()
    finally:
        ocip.lib.ocipUninitialize(ctx)
    np.savez(out, **maps)
    print("autotune child ok", flush=True)


if __name__ == "__main__":
    main(sys.argv[1])

"""The C drop-in client (tests/c_dropin.c): it compiles with a plain C compiler against
include/OpenCLIPP.h and links against libOpenCLIPP.so (CPU), and on the GPU its results
equal the same calls made through the Python binding and the oracle."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "openclipp_amd")


def build(tmp_path):
    exe = str(tmp_path / "c_dropin")
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-O1",
                           "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "c_dropin.c"),
                           "-L", LIBDIR, "-lOpenCLIPP", "-Wl,-rpath," + LIBDIR, "-o", exe])
    return exe


def test_c_client_compiles_and_links(tmp_path):
    exe = build(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_c_client_matches_python_and_oracle(tmp_path):
    from oracle import oracle as orc
    exe = build(tmp_path)
    W, H = 300, 200
    out = str(tmp_path / "out.bin")
    subprocess.check_call([exe, str(W), str(H), out], timeout=120)
    data = np.fromfile(out, dtype=np.float32)
    ssd, ncc, integ = (data[k * W * H:(k + 1) * W * H].reshape(H, W) for k in range(3))
    best, bx, by = data[3 * W * H:]
    # rebuild the same inputs as the C program
    src = np.empty(W * H, np.float32)
    s = 12345
    for i in range(W * H):
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        src[i] = np.float32((s >> 8) / 16777216.0)
    src = src.reshape(H, W)
    ox, oy = W // 3, H // 4
    tpl = np.ascontiguousarray(src[oy:oy + 16, ox:ox + 16])
    np.testing.assert_array_equal(ssd, orc.prox_ref(orc.SSD, src, tpl))
    np.testing.assert_allclose(ncc, orc.prox_f64(orc.CC_NORM, src, tpl), rtol=1e-5, atol=1e-6)
    exact = orc.integral_exact(src)
    assert np.all(np.abs(integ - exact) <= np.spacing(exact.astype(np.float32)))
    assert (int(bx), int(by)) == (ox + 8, oy + 8) and best == ncc.max()

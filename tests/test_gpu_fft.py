"""GPU parity of the FFT proximity path (ocipPrepareImageProximityFFT, ocipCrossCorrFFT,
ocipCrossCorrFFT_Norm, ocipSqrDistanceFFT, ocipSqrDistanceFFT_Norm) through the C ABI.

Oracle: the FP64 window sums of oracle.prox_f64 (the same maps the direct ops define; the
reference's FFT path computes them with a circular correlation of zero-padded images,
cl-files/ImageProximityFFT.cl:55-212).  The transforms here run in FP64, so every output is
within
    |got - f64| <= 2^-23 |f64| + 2^-40 * sqrt(sum s^2 * sum t^2) * (1 + 1/den)
with den = sqrt(SS TT) for the Norm maps (1 for the others): float rounding of the result plus
an FFT round-off far below float precision.  A wrong anchor, padding, wrap-around or spectrum
separation changes outputs by O(1) and fails this.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import say

pytestmark = pytest.mark.gpu

OPS = {"CrossCorr": orc.CC, "CrossCorr_Norm": orc.CC_NORM, "SqrDistance": orc.SSD,
       "SqrDistance_Norm": orc.SSD_NORM}
FN = {"CrossCorr": "ocipCrossCorrFFT", "CrossCorr_Norm": "ocipCrossCorrFFT_Norm",
      "SqrDistance": "ocipSqrDistanceFFT", "SqrDistance_Norm": "ocipSqrDistanceFFT_Norm"}


@pytest.fixture(scope="module")
def ocip():
    from openclipp_amd import ocip as mod
    ctx = mod.initialize()
    yield mod
    mod.lib.ocipUninitialize(ctx)


def run_fft(ocip, op, src, tpl, prog=None):
    S, T = ocip.Image(src), ocip.Image(tpl)
    D = ocip.Image(shape=src.shape, dtype=np.float32)
    own = prog is None
    if own:
        prog = ctypes.c_void_p()
        ocip.check(ocip.lib.ocipPrepareImageProximityFFT(ctypes.byref(prog), S.handle, T.handle))
    ocip.check(getattr(ocip.lib, FN[op])(prog, S.handle, T.handle, D.handle), op)
    out = D.read().copy()
    if own:
        ocip.check(ocip.lib.ocipReleaseProgram(prog))
    for i in (S, T, D):
        i.release()
    return out


def check_map(op, src, tpl, got):
    f64 = orc.prox_f64(OPS[op], src, tpl)
    s64, t64 = src.astype(np.float64), tpl.astype(np.float64)
    energy = np.sqrt((s64 * s64).sum() * (t64 * t64).sum())
    scale = np.ones_like(f64)
    if op.endswith("_Norm"):
        ss = orc.prox_f64(orc.CC, s64 * s64, np.ones_like(t64))
        scale = 1 + 1 / np.sqrt(ss * (t64 * t64).sum())
    ok = np.isfinite(f64)
    tol = 2.0 ** -23 * np.abs(f64) + 2.0 ** -40 * energy * scale
    err = np.abs(got.astype(np.float64) - f64)
    assert np.array_equal(np.isfinite(got), ok)
    assert np.all(err[ok] <= tol[ok]), (op, float(np.max((err / tol)[ok])))


SHAPES = [((1, 1), (1, 1)), ((37, 53), (5, 7)), ((64, 64), (16, 16)), ((97, 131), (31, 12)),
          ((120, 200), (64, 64)), ((50, 40), (50, 40)), ((33, 250), (4, 129))]


@pytest.mark.parametrize("op", list(OPS))
@pytest.mark.parametrize("shape,tshape", SHAPES)
def test_fft_maps_against_fp64(ocip, op, shape, tshape):
    rng = np.random.default_rng(shape[0] * 1000 + tshape[1])
    src = (rng.random(shape) - 0.3).astype(np.float32)
    oy, ox = (shape[0] - tshape[0]) // 2, (shape[1] - tshape[1]) // 3
    tpl = (src[oy:oy + tshape[0], ox:ox + tshape[1]]
           + rng.normal(0, 0.05, tshape)).astype(np.float32)
    say("fft", op, shape, tshape)
    check_map(op, src, tpl, run_fft(ocip, op, src, tpl))


def test_fft_matches_direct_path_and_finds_the_crop(ocip):
    """512x384 image, 96x80 crop: the FFT maps agree with the direct (plain FP32) maps within
    the direct path's float accumulation, and the crop is the arg-max / arg-min."""
    from tests.gpu_util import run_op
    rng = np.random.default_rng(7)
    src = rng.random((384, 512), dtype=np.float32)
    oy, ox = 101, 233
    tpl = src[oy:oy + 80, ox:ox + 96].copy()
    ncc = run_fft(ocip, "CrossCorr_Norm", src, tpl)
    assert np.unravel_index(np.nanargmax(ncc), ncc.shape) == (oy + 40, ox + 48)
    assert abs(ncc[oy + 40, ox + 48] - 1.0) < 1e-6
    ssd = run_fft(ocip, "SqrDistance", src, tpl)
    assert np.unravel_index(np.argmin(ssd), ssd.shape) == (oy + 40, ox + 48)
    assert abs(ssd[oy + 40, ox + 48]) < 1e-6 * (tpl.astype(np.float64) ** 2).sum()
    direct = run_op(ocip, "CrossCorr_Norm", src, tpl)
    assert np.max(np.abs(direct - ncc)) < 1e-5


def test_program_is_reused_across_sizes(ocip):
    """One program serves calls with other image and template sizes (it re-sizes itself)."""
    rng = np.random.default_rng(8)
    prog = ctypes.c_void_p()
    a = rng.random((70, 90), dtype=np.float32)
    ta = a[10:25, 20:41].copy()
    Sa, Ta = ocip.Image(a), ocip.Image(ta)
    ocip.check(ocip.lib.ocipPrepareImageProximityFFT(ctypes.byref(prog), Sa.handle, Ta.handle))
    for shape, tshape in [((70, 90), (15, 21)), ((128, 60), (9, 9)), ((70, 90), (15, 21))]:
        src = rng.random(shape, dtype=np.float32)
        tpl = rng.random(tshape, dtype=np.float32)
        check_map("CrossCorr_Norm", src, tpl, run_fft(ocip, "CrossCorr_Norm", src, tpl, prog))
    ocip.check(ocip.lib.ocipReleaseProgram(prog))


def test_fft_argument_errors(ocip):
    L = ocip.lib
    assert L.ocipIsFFTAvailable() == b"\x01"
    S = ocip.Image(np.zeros((32, 40), np.float32))
    T = ocip.Image(np.zeros((4, 4), np.float32))
    D = ocip.Image(np.zeros((32, 40), np.float32))
    prog = ctypes.c_void_p()
    ocip.check(L.ocipPrepareImageProximityFFT(ctypes.byref(prog), S.handle, T.handle))
    # a program of another kind, or none
    assert L.ocipCrossCorrFFT(None, S.handle, T.handle, D.handle) == ocip.CL_INVALID_PROGRAM
    # ImageProximityFFT.cpp:117-130 order: float dest, channels, size, template <= source,
    # same type; then F32 1-channel (OpenCLIPP.h:678-688)
    D8 = ocip.Image(np.zeros((32, 40), np.uint8))
    assert L.ocipCrossCorrFFT(prog, S.handle, T.handle, D8.handle) == ocip.CL_INVALID_VALUE
    Dsz = ocip.Image(np.zeros((31, 40), np.float32))
    assert L.ocipCrossCorrFFT(prog, S.handle, T.handle, Dsz.handle) == ocip.CL_INVALID_IMAGE_SIZE
    Tbig = ocip.Image(np.zeros((33, 4), np.float32))
    assert L.ocipCrossCorrFFT(prog, S.handle, Tbig.handle, D.handle) == ocip.CL_IMAGE_FORMAT_NOT_SUPPORTED
    T64 = ocip.Image(np.zeros((4, 4), np.float64))
    assert L.ocipCrossCorrFFT(prog, S.handle, T64.handle, D.handle) == -9  # CL_IMAGE_FORMAT_MISMATCH
    S8 = ocip.Image(np.zeros((32, 40), np.uint8))
    T8 = ocip.Image(np.zeros((4, 4), np.uint8))
    assert L.ocipCrossCorrFFT(prog, S8.handle, T8.handle, D.handle) == ocip.CL_IMAGE_FORMAT_NOT_SUPPORTED
    # a padded width beyond one LDS row of complex doubles
    Sw = ocip.Image(np.zeros((8, 11000), np.float32))
    Dw = ocip.Image(np.zeros((8, 11000), np.float32))
    assert L.ocipCrossCorrFFT(prog, Sw.handle, T.handle, Dw.handle) == ocip.CL_INVALID_IMAGE_SIZE
    ocip.check(L.ocipReleaseProgram(prog))


def test_fft_1024_t128(ocip):
    """1024x1024 F32 image, 128x128 crop: Norm map against FP64 on sampled pixels, the crop is
    the arg-max."""
    rng = np.random.default_rng(9)
    src = rng.random((1024, 1024), dtype=np.float32)
    oy, ox = 300, 611
    tpl = (src[oy:oy + 128, ox:ox + 128] * 1.1).astype(np.float32)
    got = run_fft(ocip, "CrossCorr_Norm", src, tpl)
    assert np.unravel_index(np.nanargmax(got), got.shape) == (oy + 64, ox + 64)
    ys = np.concatenate([rng.integers(0, 1024, 255), [oy + 64]])
    xs = np.concatenate([rng.integers(0, 1024, 255), [ox + 64]])
    P = np.pad(src.astype(np.float64), [(64, 63), (64, 63)])
    t = tpl.astype(np.float64)
    win = np.stack([P[y:y + 128, x:x + 128] for x, y in zip(xs, ys)])
    want = (win * t).sum(axis=(1, 2)) / np.sqrt((win * win).sum(axis=(1, 2)) * (t * t).sum())
    assert np.max(np.abs(got[ys, xs] - want)) < 1e-6


# ------------------------------------------------------------- generic FFT (ocipFFTForward/Inverse)
def complex_image(ocip, H, W, dtype, extra=0):
    """A complex image for a W-wide real image: W/2 + 1 (+ extra) complex values per row, as a
    2-channel image of the real type (real, imaginary interleaved)."""
    return ocip.Image(np.zeros((H, W // 2 + 1 + extra, 2), dtype))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1, 1), (8, 12), (30, 45), (64, 100), (27, 250), (81, 2),
                                   (16, 64), (9, 72)])
def test_fft_forward_inverse_against_numpy(ocip, dtype, shape):
    """Forward = numpy.fft.rfft2 (unnormalised half spectrum along x), inverse =
    W*H*irfft2; FP64 inside, so F64 agrees to ~1e-13 of the sum of |x| and F32 to its own
    rounding of the result."""
    H, W = shape
    rng = np.random.default_rng(H * 100 + W)
    x = (rng.random((H, W)) - 0.4).astype(dtype)
    R, C = ocip.Image(x), complex_image(ocip, H, W, dtype)
    prog = ctypes.c_void_p()
    ocip.check(ocip.lib.ocipPrepareFFT(ctypes.byref(prog), R.handle, C.handle))
    ocip.check(ocip.lib.ocipFFTForward(prog, R.handle, C.handle))
    got = C.read()
    want = np.fft.rfft2(x.astype(np.float64))
    scale = np.abs(x.astype(np.float64)).sum()
    rel = 2.0 ** -23 if dtype == np.float32 else 0.0
    err = np.abs(got[..., 0] + 1j * got[..., 1] - want)
    assert np.all(err <= rel * np.abs(want) + 1e-13 * scale + 1e-300), float(err.max())
    # inverse of the exact spectrum, written to a fresh real image
    spec = np.zeros(got.shape, dtype)
    spec[..., 0], spec[..., 1] = want.real, want.imag     # rounded to the image type
    C2, R2 = ocip.Image(spec), ocip.Image(np.zeros((H, W), dtype))
    ocip.check(ocip.lib.ocipFFTInverse(prog, C2.handle, R2.handle))
    back = R2.read().astype(np.float64)
    ref = np.fft.irfft2(spec[..., 0].astype(np.float64) + 1j * spec[..., 1].astype(np.float64),
                        s=(H, W)) * (W * H)
    tol = (2.0 ** -23 if dtype == np.float32 else 0.0) * np.abs(ref) + 1e-12 * scale * 8
    assert np.all(np.abs(back - ref) <= tol + 1e-300), float(np.abs(back - ref).max())
    ocip.check(ocip.lib.ocipReleaseProgram(prog))


def test_fft_round_trip_and_padded_complex_rows(ocip):
    """inverse(forward(x)) = W*H*x, through a complex image with room for more than W/2 + 1
    values per row."""
    rng = np.random.default_rng(77)
    H, W = 96, 160
    x = rng.random((H, W))
    R, C, R2 = ocip.Image(x), complex_image(ocip, H, W, np.float64, extra=7), ocip.Image(np.zeros((H, W)))
    prog = ctypes.c_void_p()
    ocip.check(ocip.lib.ocipPrepareFFT(ctypes.byref(prog), None, None))
    ocip.check(ocip.lib.ocipFFTForward(prog, R.handle, C.handle))
    ocip.check(ocip.lib.ocipFFTInverse(prog, C.handle, R2.handle))
    np.testing.assert_allclose(R2.read() / (W * H), x, rtol=0, atol=1e-13)
    ocip.check(ocip.lib.ocipReleaseProgram(prog))


def test_fft_argument_errors_generic(ocip):
    L = ocip.lib
    prog = ctypes.c_void_p()
    ocip.check(L.ocipPrepareFFT(ctypes.byref(prog), None, None))
    R = ocip.Image(np.zeros((8, 12), np.float32))
    C = complex_image(ocip, 8, 12, np.float32)
    assert L.ocipFFTForward(None, R.handle, C.handle) == ocip.CL_INVALID_PROGRAM
    # FFT.cpp:140-168 order: same type, F32/F64, 1-channel real, 2-3-5 lengths, complex width
    C64 = complex_image(ocip, 8, 12, np.float64)
    assert L.ocipFFTForward(prog, R.handle, C64.handle) == ocip.CL_IMAGE_FORMAT_NOT_SUPPORTED
    R8, C8 = ocip.Image(np.zeros((8, 12), np.uint8)), ocip.Image(np.zeros((8, 14), np.uint8))
    assert L.ocipFFTForward(prog, R8.handle, C8.handle) == ocip.CL_IMAGE_FORMAT_NOT_SUPPORTED
    R7 = ocip.Image(np.zeros((8, 7), np.float32))
    assert L.ocipFFTForward(prog, R7.handle, C.handle) == ocip.CL_INVALID_IMAGE_SIZE
    Cn = ocip.Image(np.zeros((8, 6, 2), np.float32))    # 6 < 12/2 + 1 complex values per row
    assert L.ocipFFTForward(prog, R.handle, Cn.handle) == ocip.CL_IMAGE_FORMAT_NOT_SUPPORTED
    ocip.check(L.ocipReleaseProgram(prog))

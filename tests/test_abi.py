"""CPU checks of the C ABI: the library loads, exports every declared entry point, and the
argument checks follow the reference's order and error codes (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import pytest

from openclipp_amd import ocip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "OpenCLIPP.h")


def declared_entry_points():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ocip[A-Z]\w*)\s*\(", text)))


def test_header_declares_the_reference_path_entry_points():
    names = declared_entry_points()
    for n in ["ocipInitialize", "ocipUninitialize", "ocipChangeContext", "ocipSetCLFilesPath",
              "ocipGetErrorName", "ocipGetDeviceName", "ocipFinish", "ocipCreateImage",
              "ocipSendImage", "ocipReadImage", "ocipReleaseImage", "ocipReleaseProgram",
              "ocipPrepareProximity", "ocipSqrDistance_Norm", "ocipSqrDistance",
              "ocipAbsDistance", "ocipCrossCorr", "ocipCrossCorr_Norm", "ocipPrepareIntegral",
              "ocipIntegral", "ocipSqrIntegral"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    out = subprocess.check_output(["nm", "-D", "--defined-only", ocip.LIB_PATH], text=True)
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    missing = [n for n in declared_entry_points() if n not in exported]
    assert not missing, missing
    # the binding knows every declared entry point with its argument list
    assert set(declared_entry_points()) == set(ocip.SIGNATURES)
    # nothing but the ABI leaks out of the library
    assert all(n.startswith("ocip") or not n[0].isalpha() or n.startswith("_") for n in exported
               if not n.startswith("__")), sorted(exported)[:20]


def test_error_names():
    assert ocip.lib.ocipGetErrorName(0) == b"SUCCESS"
    assert ocip.lib.ocipGetErrorName(-30) == b"CL_INVALID_VALUE"
    assert ocip.lib.ocipGetErrorName(-34) == b"CL_INVALID_CONTEXT"
    assert ocip.lib.ocipGetErrorName(-40) == b"CL_INVALID_IMAGE_SIZE"
    assert ocip.lib.ocipGetErrorName(12345) == b"Unknown error"


def test_no_context_errors():
    """Entry points that need the current context return CL_INVALID_CONTEXT before touching a
    device (OpenCLIPP/OpenCLIPP.cpp:99-175, 717-727)."""
    L = ocip.lib
    h = ctypes.c_void_p()
    img = ocip.SImage(4, 4, 16, 1, ocip.F32)
    assert L.ocipCreateImage(ctypes.byref(h), img, None, ocip.CL_MEM_READ_WRITE) == ocip.CL_INVALID_CONTEXT
    assert L.ocipFinish() == ocip.CL_INVALID_CONTEXT
    assert L.ocipGetDeviceName(ctypes.create_string_buffer(64), 64) == ocip.CL_INVALID_CONTEXT
    assert L.ocipUninitialize(None) == ocip.CL_INVALID_CONTEXT
    assert L.ocipChangeContext(None) == ocip.CL_INVALID_CONTEXT
    assert L.ocipCrossCorr_Norm(None, None, None) == ocip.CL_INVALID_CONTEXT
    assert L.ocipPrepareIntegral(ctypes.byref(h), None) == ocip.CL_INVALID_CONTEXT
    assert L.ocipSum(None, None, None) == ocip.CL_INVALID_CONTEXT
    assert L.ocipCrossCorrFFT_Norm(None, None, None, None) == ocip.CL_INVALID_CONTEXT
    assert L.ocipPrepareImageProximityFFT(ctypes.byref(h), None, None) == ocip.CL_INVALID_CONTEXT
    assert L.ocipIsFFTAvailable() == b"\x01"
    assert L.ocipSumSqr(None, None, None) == ocip.CL_INVALID_CONTEXT
    assert L.ocipReleaseProgram(None) == ocip.CL_INVALID_PROGRAM
    assert L.ocipSendImage(None) == ocip.CL_INVALID_MEM_OBJECT
    assert L.ocipReadImage(None) == ocip.CL_INVALID_MEM_OBJECT
    L.ocipSetCLFilesPath(b"/nonexistent")  # accepted and ignored


def test_initialize_without_gpu_or_with_cpu_type():
    ctx = ctypes.c_void_p()
    # only GPU devices exist for this library
    assert ocip.lib.ocipInitialize(ctypes.byref(ctx), None, ocip.CL_DEVICE_TYPE_CPU) == ocip.CL_DEVICE_NOT_FOUND
    try:
        import torch
        has_gpu = torch.cuda.device_count() > 0
    except Exception:
        has_gpu = False
    if not has_gpu:
        assert ocip.lib.ocipInitialize(ctypes.byref(ctx), None, ocip.CL_DEVICE_TYPE_ALL) == ocip.CL_DEVICE_NOT_FOUND


# --------------------------------------------------------------------------- argument checks
@pytest.fixture(scope="module")
def vc(tmp_path_factory):
    out = tmp_path_factory.mktemp("vc") / "libvc.so"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-shared", "-fPIC",
                           "-I", os.path.join(ROOT, "include"),
                           "-I", os.path.join(ROOT, "openclipp_amd", "csrc"),
                           os.path.join(ROOT, "tests", "validation_shim.cpp"), "-o", str(out)])
    L = ctypes.CDLL(str(out))
    for f in ("vc_create_image", "vc_proximity", "vc_integral", "vc_precision"):
        getattr(L, f).restype = ctypes.c_int
    return L


def S(w, h, t=ocip.F32, c=1, step=None):
    esz = {0: 1, 1: 1, 2: 2, 3: 2, 4: 4, 5: 4, 6: 4, 7: 8}[t]
    return ocip.SImage(w, h, step if step is not None else w * c * esz, c, t)


def test_create_image_checks(vc):
    assert vc.vc_create_image(S(8, 8)) == 0
    assert vc.vc_create_image(S(8, 8, c=3)) == 0   # a colour image, 4 channels on the device
    assert vc.vc_create_image(S(8, 8, c=5)) == ocip.CL_IMAGE_FORMAT_NOT_SUPPORTED
    assert vc.vc_create_image(ocip.SImage(8, 8, 32, 1, 8)) == ocip.CL_IMAGE_FORMAT_NOT_SUPPORTED
    assert vc.vc_create_image(S(8, 0)) == ocip.CL_INVALID_BUFFER_SIZE
    assert vc.vc_create_image(S(8, 8, step=16)) == ocip.CL_INVALID_VALUE
    assert vc.vc_create_image(S(8, 8, t=ocip.S16, step=17)) == ocip.CL_INVALID_VALUE
    assert vc.vc_create_image(S(8, 8, step=64)) == 0  # pitched


def test_proximity_check_order(vc):
    """ImageProximity.cpp:36-44: size of template first, then type, then size, channels,
    float dest."""
    ok = vc.vc_proximity(S(32, 32), S(16, 16), S(32, 32))
    assert ok == 0
    # template larger than the source wins over every later failure
    assert vc.vc_proximity(S(8, 8), S(16, 4, t=ocip.U8), S(4, 4, t=ocip.U8)) == ocip.CL_INVALID_VALUE
    # type mismatch (float vs int, signedness, depth, channels) before the size check
    for t in (S(4, 4, t=ocip.S32), S(4, 4, t=ocip.F64), S(4, 4, c=2)):
        assert vc.vc_proximity(S(32, 32), t, S(7, 7)) == ocip.CL_INVALID_VALUE
    assert vc.vc_proximity(S(32, 32, t=ocip.U16), S(4, 4, t=ocip.S16), S(32, 32)) == ocip.CL_INVALID_VALUE
    # dest size then channels then float dest
    assert vc.vc_proximity(S(32, 32), S(4, 4), S(31, 32)) == ocip.CL_INVALID_IMAGE_SIZE
    assert vc.vc_proximity(S(32, 32), S(4, 4), S(32, 32, c=2)) == ocip.CL_INVALID_VALUE
    assert vc.vc_proximity(S(32, 32, t=ocip.U8), S(4, 4, t=ocip.U8), S(32, 32, t=ocip.U8)) == ocip.CL_INVALID_VALUE
    # REAL of the kernels: F32 dest for non-F64 sources, F64 dest for F64 sources
    assert vc.vc_proximity(S(32, 32, t=ocip.U8), S(4, 4, t=ocip.U8), S(32, 32)) == 0
    assert vc.vc_proximity(S(32, 32), S(4, 4), S(32, 32, t=ocip.F64)) == ocip.CL_INVALID_VALUE
    assert vc.vc_proximity(S(32, 32, t=ocip.F64), S(4, 4, t=ocip.F64), S(32, 32, t=ocip.F64)) == 0
    # template equal to the source is allowed
    assert vc.vc_proximity(S(16, 16), S(16, 16), S(16, 16)) == 0


def test_integral_checks(vc):
    assert vc.vc_integral(S(10, 10, t=ocip.U8), S(10, 10, t=ocip.F64)) == 0
    assert vc.vc_integral(S(10, 10, t=ocip.U8), S(10, 10, t=ocip.F32)) == 0
    # dest type is checked before the size (Integral.cpp:97-118 then :145)
    assert vc.vc_integral(S(10, 10, t=ocip.U8), S(11, 10, t=ocip.S32)) == ocip.CL_INVALID_VALUE
    assert vc.vc_integral(S(10, 10, t=ocip.U8), S(11, 10, t=ocip.F32)) == ocip.CL_INVALID_IMAGE_SIZE
    assert vc.vc_integral(S(10, 10, c=2), S(10, 10)) == ocip.CL_INVALID_VALUE


def test_precision_checks(vc):
    assert vc.vc_precision(S(8, 8), ocip.PREC_BF16) == 0
    assert vc.vc_precision(S(8, 8, t=ocip.U8), ocip.PREC_F32) == 0
    assert vc.vc_precision(S(8, 8, t=ocip.U8), ocip.PREC_BF16) == ocip.CL_INVALID_VALUE
    assert vc.vc_precision(S(8, 8), 7) == ocip.CL_INVALID_VALUE

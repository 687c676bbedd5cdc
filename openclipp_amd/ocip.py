"""ctypes binding of libOpenCLIPP.so (the drop-in C ABI in include/OpenCLIPP.h).

The functions keep the C names and argument order; each returns the ocipError code
(0 = CL_SUCCESS), exactly as a C caller sees them.  `Image` is a small convenience
that owns an ocipImage created from a numpy array (the array is the host side of
Send/Read, as the ImageData pointer of ocipCreateImage is in C).

There is no fallback: if the shared library is missing or fails to load, importing
this module raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libOpenCLIPP.so")

# SImage::EDataType
U8, S8, U16, S16, U32, S32, F32, F64 = range(8)
NP_OF_TYPE = {U8: np.uint8, S8: np.int8, U16: np.uint16, S16: np.int16,
              U32: np.uint32, S32: np.int32, F32: np.float32, F64: np.float64}
TYPE_OF_NP = {np.dtype(v): k for k, v in NP_OF_TYPE.items()}

# Khronos error codes used by the ABI
CL_SUCCESS = 0
CL_DEVICE_NOT_FOUND = -1
CL_MEM_OBJECT_ALLOCATION_FAILURE = -4
CL_OUT_OF_RESOURCES = -5
CL_IMAGE_FORMAT_NOT_SUPPORTED = -10
CL_INVALID_VALUE = -30
CL_INVALID_CONTEXT = -34
CL_INVALID_MEM_OBJECT = -38
CL_INVALID_IMAGE_SIZE = -40
CL_INVALID_PROGRAM = -44
CL_INVALID_OPERATION = -59
CL_INVALID_BUFFER_SIZE = -61

CL_DEVICE_TYPE_CPU = 1 << 1
CL_DEVICE_TYPE_GPU = 1 << 2
CL_DEVICE_TYPE_ALL = 0xFFFFFFFF
CL_MEM_READ_WRITE = 1 << 0
CL_MEM_WRITE_ONLY = 1 << 1
CL_MEM_READ_ONLY = 1 << 2

PREC_F32, PREC_BF16, PREC_F16, PREC_BF16X3 = 0, 1, 2, 3


class SImage(ctypes.Structure):
    _fields_ = [("Width", ctypes.c_uint), ("Height", ctypes.c_uint), ("Step", ctypes.c_uint),
                ("Channels", ctypes.c_uint), ("Type", ctypes.c_int)]


# (name, argtypes) for every entry point declared in include/OpenCLIPP.h
_vp = ctypes.c_void_p
_ptr_vp = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = {
    "ocipInitialize": [_ptr_vp, ctypes.c_char_p, ctypes.c_uint64],
    "ocipUninitialize": [_vp],
    "ocipChangeContext": [_vp],
    "ocipSetCLFilesPath": [ctypes.c_char_p],
    "ocipGetErrorName": [ctypes.c_int],
    "ocipGetDeviceName": [ctypes.c_char_p, ctypes.c_uint],
    "ocipFinish": [],
    "ocipCreateImage": [_ptr_vp, SImage, _vp, ctypes.c_uint64],
    "ocipSendImage": [_vp],
    "ocipReadImage": [_vp],
    "ocipReleaseImage": [_vp],
    "ocipReleaseProgram": [_vp],
    "ocipPrepareIntegral": [_ptr_vp, _vp],
    "ocipIntegral": [_vp, _vp, _vp],
    "ocipSqrIntegral": [_vp, _vp, _vp],
    "ocipPrepareProximity": [_vp],
    "ocipSqrDistance_Norm": [_vp, _vp, _vp],
    "ocipSqrDistance": [_vp, _vp, _vp],
    "ocipAbsDistance": [_vp, _vp, _vp],
    "ocipCrossCorr": [_vp, _vp, _vp],
    "ocipCrossCorr_Norm": [_vp, _vp, _vp],
    "ocipPrepareStatistics": [_ptr_vp, _vp],
    "ocipMinIndx": [_vp, _vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                    ctypes.POINTER(ctypes.c_int)],
    "ocipMaxIndx": [_vp, _vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                    ctypes.POINTER(ctypes.c_int)],
    "ocipSum": [_vp, _vp, ctypes.POINTER(ctypes.c_double)],
    "ocipSumSqr": [_vp, _vp, ctypes.POINTER(ctypes.c_double)],
    "ocipIsFFTAvailable": [],
    "ocipPrepareFFT": [_ptr_vp, _vp, _vp],
    "ocipFFTForward": [_vp, _vp, _vp],
    "ocipFFTInverse": [_vp, _vp, _vp],
    "ocipPrepareImageProximityFFT": [_ptr_vp, _vp, _vp],
    "ocipSqrDistanceFFT": [_vp, _vp, _vp, _vp],
    "ocipSqrDistanceFFT_Norm": [_vp, _vp, _vp, _vp],
    "ocipCrossCorrFFT": [_vp, _vp, _vp, _vp],
    "ocipCrossCorrFFT_Norm": [_vp, _vp, _vp, _vp],
    "ocipCrossCorrEx": [_vp, _vp, _vp, ctypes.c_int],
    "ocipCrossCorr_NormEx": [_vp, _vp, _vp, ctypes.c_int],
    "ocipCrossCorr_NormExNoAlloc": [_vp, _vp, _vp, ctypes.c_int],
    "ocipCrossCorr_NormBatch": [_vp, _ptr_vp, _ptr_vp, ctypes.c_uint, ctypes.c_int],
    "ocipSqrDistance_NormBatch": [_vp, _ptr_vp, _ptr_vp, ctypes.c_uint, ctypes.c_int],
    "ocipCrossCorr_NormBackward": [_vp, _vp, _vp, _vp, _vp],
    "ocipSqrDistanceBackward": [_vp, _vp, _vp, _vp, _vp],
    "ocipProximityBackwardBatch": [_ptr_vp, _vp, _ptr_vp, _ptr_vp, _ptr_vp, _vp, ctypes.c_uint],
    "ocipEnableKernelTiming": [ctypes.c_char],
    "ocipGetKernelTime": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_uint)],
}

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
lib = ctypes.CDLL(LIB_PATH)
for _name, _args in SIGNATURES.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = ctypes.c_char_p if _name == "ocipGetErrorName" else (
        None if _name == "ocipSetCLFilesPath" else (
            ctypes.c_char if _name == "ocipIsFFTAvailable" else ctypes.c_int))


class OcipError(RuntimeError):
    def __init__(self, code, what):
        name = lib.ocipGetErrorName(code).decode()
        super().__init__(f"{what}: {name} ({code})")
        self.code = code


def check(code, what="ocip call"):
    if code != CL_SUCCESS:
        raise OcipError(code, what)
    return code


def initialize(device_type=CL_DEVICE_TYPE_ALL):
    ctx = ctypes.c_void_p()
    check(lib.ocipInitialize(ctypes.byref(ctx), None, device_type), "ocipInitialize")
    return ctx


def simage(arr):
    """SImage describing a numpy image of shape (H, W) or (H, W, C)."""
    if arr.ndim == 2:
        h, w = arr.shape
        c = 1
    else:
        h, w, c = arr.shape
    return SImage(w, h, arr.strides[0], c, TYPE_OF_NP[arr.dtype])


class Image:
    """An ocipImage whose host side is a numpy array (kept alive by this object)."""

    def __init__(self, arr=None, shape=None, dtype=None, channels=1, flags=CL_MEM_READ_WRITE,
                 device_only=False):
        if arr is None:
            shp = tuple(shape) if channels == 1 else tuple(shape) + (channels,)
            arr = np.zeros(shp, dtype=dtype)
        self.arr = arr
        self.handle = ctypes.c_void_p()
        data = None if device_only else arr.ctypes.data_as(ctypes.c_void_p)
        check(lib.ocipCreateImage(ctypes.byref(self.handle), simage(arr), data, flags),
              "ocipCreateImage")

    def send(self):
        check(lib.ocipSendImage(self.handle), "ocipSendImage")
        return self

    def read(self):
        check(lib.ocipReadImage(self.handle), "ocipReadImage")
        return self.arr

    def release(self):
        if self.handle:
            check(lib.ocipReleaseImage(self.handle), "ocipReleaseImage")
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            if self.handle:
                lib.ocipReleaseImage(self.handle)
        except Exception:
            pass


PROXIMITY_OPS = ("SqrDistance", "SqrDistance_Norm", "AbsDistance", "CrossCorr", "CrossCorr_Norm")


def proximity(op, src, tpl, dst):
    """Run ocip<op>(src, tpl, dst) on Image objects; returns the error code."""
    return getattr(lib, "ocip" + op)(src.handle, tpl.handle, dst.handle)


def kernel_time(prefix):
    """(total_ms, count) of the library's kernels whose name starts with `prefix`."""
    ms, n = ctypes.c_double(), ctypes.c_uint()
    check(lib.ocipGetKernelTime(prefix.encode(), ctypes.byref(ms), ctypes.byref(n)),
          "ocipGetKernelTime")
    return ms.value, n.value


def handles(images):
    arr = (ctypes.c_void_p * len(images))(*[i.handle.value for i in images])
    return arr

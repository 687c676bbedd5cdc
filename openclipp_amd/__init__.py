"""MI355X-native template matching behind OpenCLIPP's C interface.

The product is the shared library libOpenCLIPP.so built from openclipp_amd/csrc
(hand-written HIP for gfx950 behind the C ABI declared in include/OpenCLIPP.h).
`openclipp_amd.ocip` is its ctypes binding; `openclipp_amd.build` compiles it.
"""

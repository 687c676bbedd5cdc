// validation.h -- argument checks of the template-matching entry points, as pure functions
// over SImage descriptors (no device calls), in the reference's order.  Returns CL_SUCCESS or
// the first failing check's error code.  Header-only so that a CPU test can compile it alone.
#pragma once

#include "OpenCLIPP.h"

namespace ocip {
namespace check {

inline int elem_bytes(int t)
{
   switch (t) {
   case SImage::U8: case SImage::S8: return 1;
   case SImage::U16: case SImage::S16: return 2;
   case SImage::U32: case SImage::S32: case SImage::F32: return 4;
   case SImage::F64: return 8;
   default: return 0;
   }
}
inline bool floating(int t) { return t == SImage::F32 || t == SImage::F64; }
inline bool unsigned_(int t) { return t == SImage::U8 || t == SImage::U16 || t == SImage::U32; }

// C++/Image.cpp:147-165 (channels 1..4, 3 being a colour image held as 4 channels on the
// device, OpenCLIPP/OpenCLIPP.cpp:185-188; type in range), plus the buffer checks of this
// build: a non-empty buffer, a Step that holds a row and is a multiple of the element size.
inline cl_int create_image(const SImage &i)
{
   if (i.Channels < 1 || i.Channels > 4) return CL_IMAGE_FORMAT_NOT_SUPPORTED;
   if ((int)i.Type < 0 || (int)i.Type >= (int)SImage::NbDataTypes) return CL_IMAGE_FORMAT_NOT_SUPPORTED;
   if ((unsigned long long)i.Height * i.Step == 0) return -61 /* CL_INVALID_BUFFER_SIZE */;
   if ((unsigned long long)i.Step < (unsigned long long)i.Width * i.Channels * elem_bytes(i.Type))
      return CL_INVALID_VALUE;
   if (i.Step % elem_bytes(i.Type)) return CL_INVALID_VALUE;
   return CL_SUCCESS;
}

// C++/programs/Program.cpp:440-446 SameType
inline bool same_type(const SImage &a, const SImage &b)
{
   return floating(a.Type) == floating(b.Type) && unsigned_(a.Type) == unsigned_(b.Type) &&
          elem_bytes(a.Type) == elem_bytes(b.Type) && a.Channels == b.Channels;
}

// C++/programs/ImageProximity.cpp:36-44: template <= source, SameType(S, T),
// CheckSameSize(S, D), CheckSameNbChannels(S, D), CheckFloat(D); then this build's check
// that D's depth is the kernels' REAL (float, double for F64 sources,
// cl-files/Buffers.h:58-63) -- the reference accepts a mismatched depth and writes it with
// the wrong stride.
inline cl_int proximity(const SImage &S, const SImage &T, const SImage &D)
{
   if (T.Width > S.Width || T.Height > S.Height) return CL_INVALID_VALUE;
   if (!same_type(S, T)) return CL_INVALID_VALUE;
   if (S.Width != D.Width || S.Height != D.Height) return CL_INVALID_IMAGE_SIZE;
   if (S.Channels != D.Channels) return CL_INVALID_VALUE;
   if (!floating(D.Type)) return CL_INVALID_VALUE;
   const int real = S.Type == SImage::F64 ? SImage::F64 : SImage::F32;
   if ((int)D.Type != real) return CL_INVALID_VALUE;
   return CL_SUCCESS;
}

// C++/programs/ImageProximityFFT.cpp:117-130 (each of the four ops): CheckFloat(D),
// CheckSameNbChannels(S, T) and (S, D), CheckSameSize(S, D), template <= source
// (CL_IMAGE_FORMAT_NOT_SUPPORTED), SameType(S, T) (CL_IMAGE_FORMAT_MISMATCH); then the
// contract of include/OpenCLIPP.h:678-688, "Images must be F32 - 1 channel".
inline cl_int proximity_fft(const SImage &S, const SImage &T, const SImage &D)
{
   if (!floating(D.Type)) return CL_INVALID_VALUE;
   if (S.Channels != T.Channels || S.Channels != D.Channels) return CL_INVALID_VALUE;
   if (S.Width != D.Width || S.Height != D.Height) return CL_INVALID_IMAGE_SIZE;
   if (T.Width > S.Width || T.Height > S.Height) return CL_IMAGE_FORMAT_NOT_SUPPORTED;
   if (!same_type(S, T)) return -9 /* CL_IMAGE_FORMAT_MISMATCH */;
   if (S.Type != SImage::F32 || S.Channels != 1 || D.Type != SImage::F32)
      return CL_IMAGE_FORMAT_NOT_SUPPORTED;
   return CL_SUCCESS;
}

// C++/programs/FFT.cpp:140-168 (PrepareFor(Real, Complex)): same data type, F32 or F64, a
// 1-channel real image, 2-3-5-smooth width and height (CL_INVALID_IMAGE_SIZE, checked by the
// caller with the plan), room for W/2 + 1 complex values per complex row; then this build's
// check that the complex image has the real image's height (the reference would go out of
// bounds).
inline cl_int fft(const SImage &R, const SImage &C)
{
   if (R.Type != C.Type) return CL_IMAGE_FORMAT_NOT_SUPPORTED;
   if (C.Type != SImage::F32 && C.Type != SImage::F64) return CL_IMAGE_FORMAT_NOT_SUPPORTED;
   if (R.Channels != 1) return CL_IMAGE_FORMAT_NOT_SUPPORTED;
   return CL_SUCCESS;
}
inline cl_int fft_layout(const SImage &R, const SImage &C)
{
   if (C.Step / elem_bytes(C.Type) / 2 < R.Width / 2 + 1) return CL_IMAGE_FORMAT_NOT_SUPPORTED;
   if (C.Height < R.Height) return CL_INVALID_IMAGE_SIZE;
   return CL_SUCCESS;
}

// C++/programs/Integral.cpp:97-146: Dest F32/F64, then CheckSameSize.  One channel on both
// sides (include/OpenCLIPP.h:642-645; the reference does not check it).
inline cl_int integral(const SImage &S, const SImage &D)
{
   if (D.Type != SImage::F32 && D.Type != SImage::F64) return CL_INVALID_VALUE;
   if (S.Channels != 1 || D.Channels != 1) return CL_INVALID_VALUE;
   if (S.Width != D.Width || S.Height != D.Height) return CL_INVALID_IMAGE_SIZE;
   return CL_SUCCESS;
}

// Extensions: MFMA precisions, batch and backward need F32 1-channel images.
inline cl_int precision(const SImage &S, int p)
{
   if (p < ocipPrecisionF32 || p > ocipPrecisionBF16x3) return CL_INVALID_VALUE;
   if (p != ocipPrecisionF32 && (S.Type != SImage::F32 || S.Channels != 1)) return CL_INVALID_VALUE;
   return CL_SUCCESS;
}

}  // namespace check
}  // namespace ocip

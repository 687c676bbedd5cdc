// Device helpers shared by the kernels: typed element loads and wavefront scans.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ocip_internal.h"

namespace ocip {

template <int TYPE> struct ElemT;
template <> struct ElemT<SImage::U8>  { typedef uint8_t  type; };
template <> struct ElemT<SImage::S8>  { typedef int8_t   type; };
template <> struct ElemT<SImage::U16> { typedef uint16_t type; };
template <> struct ElemT<SImage::S16> { typedef int16_t  type; };
template <> struct ElemT<SImage::U32> { typedef uint32_t type; };
template <> struct ElemT<SImage::S32> { typedef int32_t  type; };
template <> struct ElemT<SImage::F32> { typedef float    type; };
template <> struct ElemT<SImage::F64> { typedef double   type; };

// element i of the row starting at `row` (byte address)
template <int TYPE>
__device__ __forceinline__ typename ElemT<TYPE>::type ld(const char *row, int i)
{
   return reinterpret_cast<const typename ElemT<TYPE>::type *>(row)[i];
}

// Dispatch a runtime SImage type to a template parameter.
#define OCIP_TYPE_SWITCH(t, TT, ...)                                     \
   switch (t) {                                                          \
   case SImage::U8:  { constexpr int TT = SImage::U8;  __VA_ARGS__; } break;  \
   case SImage::S8:  { constexpr int TT = SImage::S8;  __VA_ARGS__; } break;  \
   case SImage::U16: { constexpr int TT = SImage::U16; __VA_ARGS__; } break;  \
   case SImage::S16: { constexpr int TT = SImage::S16; __VA_ARGS__; } break;  \
   case SImage::U32: { constexpr int TT = SImage::U32; __VA_ARGS__; } break;  \
   case SImage::S32: { constexpr int TT = SImage::S32; __VA_ARGS__; } break;  \
   case SImage::F32: { constexpr int TT = SImage::F32; __VA_ARGS__; } break;  \
   case SImage::F64: { constexpr int TT = SImage::F64; __VA_ARGS__; } break;  \
   default: fail(CL_IMAGE_FORMAT_NOT_SUPPORTED, "type");                \
   }

// Inclusive scan across the 64 lanes of a wavefront.
template <class T> __device__ __forceinline__ T wave_inclusive_scan(T v)
{
   const int lane = threadIdx.x & 63;
#pragma unroll
   for (int d = 1; d < 64; d <<= 1) {
      T o = __shfl_up(v, d, 64);
      if (lane >= d) v += o;
   }
   return v;
}

// Inclusive scans across the 64 lanes of N doubles at once, by DPP moves of both 32-bit halves
// (no LDS round trips; the N scans interleaved): within rows of 16 lanes, then rows 1 and 3
// take the last lane of rows 0 and 2, rows 2 and 3 the last lane of row 1.  A lane without a
// source adds -0.0, which leaves every value (signed zeros included) unchanged.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v)
{
   const unsigned long long b = __double_as_longlong(v);
   const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROWS, 0xf, false);
   const int hi = __builtin_amdgcn_update_dpp((int)0x80000000u, (int)(unsigned)(b >> 32), CTRL,
                                              ROWS, 0xf, false);
   return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

template <int N>
__device__ __forceinline__ void wave_incl_scan_f64(double (&x)[N])
{
#pragma unroll
   for (int k = 0; k < N; ++k) x[k] += dpp_f64<0x111, 0xf>(x[k]);   // row_shr:1
#pragma unroll
   for (int k = 0; k < N; ++k) x[k] += dpp_f64<0x112, 0xf>(x[k]);   // row_shr:2
#pragma unroll
   for (int k = 0; k < N; ++k) x[k] += dpp_f64<0x114, 0xf>(x[k]);   // row_shr:4
#pragma unroll
   for (int k = 0; k < N; ++k) x[k] += dpp_f64<0x118, 0xf>(x[k]);   // row_shr:8
#pragma unroll
   for (int k = 0; k < N; ++k) x[k] += dpp_f64<0x142, 0xa>(x[k]);   // row_bcast:15
#pragma unroll
   for (int k = 0; k < N; ++k) x[k] += dpp_f64<0x143, 0xc>(x[k]);   // row_bcast:31
}

template <class T> __device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
   for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
   return v;
}

inline void launch_check(const char *what) { hip_check(hipGetLastError(), what); }

// XCD-aware tile order.  The 8 XCDs of an MI355X each have their own L2 and receive
// workgroups round-robin in dispatch order (workgroup b runs on XCD b % 8).  Remapping
// workgroup b to tile (b % 8) * (n / 8) + b / 8 gives every XCD one contiguous run of tiles
// (row-major), so tiles that share halo rows and integral rows share an L2.  Bijective on
// [0, n): the last n % 8 workgroups keep their own index.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int n)
{
   const int q = n / kXcds;
   if (b >= q * kXcds) return b;
   return (b % kXcds) * q + b / kXcds;
}

// Sum over columns [xa, xb) and rows [ya, yb), clamped to the image, from the internal
// exclusive FP64 integral (layout: excl_pitch / kExclOffset in ocip_internal.h).
__device__ __forceinline__ double rect_sum(const double *__restrict__ I, int W, int H, int xa,
                                           int xb, int ya, int yb)
{
   xa = min(max(xa, 0), W); xb = min(max(xb, 0), W);
   ya = min(max(ya, 0), H); yb = min(max(yb, 0), H);
   const size_t P = excl_pitch(W);
   const double *r0 = I + ya * P + kExclOffset, *r1 = I + yb * P + kExclOffset;
   return (r1[xb] - r0[xb]) - (r1[xa] - r0[xa]);
}

// 1 / sqrt(p) for the Norm epilogues: p (= window sum s^2 * sum t^2) is formed in FP64; the
// reciprocal square root is taken in FP32 (v_rsq_f32, about 1 ulp) when p is in float range,
// in FP64 otherwise.  p = 0 gives +inf, as the reference's x / sqrt(0) does.
__device__ __forceinline__ float inv_sqrt(double p)
{
   if (p >= 1e-30 && p <= 1e30) return __builtin_amdgcn_rsqf((float)p);
   return (float)(1.0 / sqrt(p));
}

// Window sum of the Norm denominators: the template window of output (x, y) with origin at
// (-tw/2, -th/2) (cl-files/ImageProximityFFT.cl:55-84 gives the same window).
__device__ __forceinline__ double window_sqsum(const double *__restrict__ I, int W, int H,
                                               int x, int y, int tw, int th)
{
   return rect_sum(I, W, H, x - tw / 2, x - tw / 2 + tw, y - th / 2, y - th / 2 + th);
}

}  // namespace ocip

// mfma_kernels.h -- cross-correlation as a contraction on gfx950 matrix cores.
//
// Semantics: reference CrossCorr / CrossCorr_Norm / SqrDistance_Norm
// (cl-files/ImageProximity.cl:122-158, 290-466):
//   C(x,y) = sum_v sum_u s(x+u-tw/2, y+v-th/2) * t(u,v), s = 0 outside the image.
// Arithmetic: products of BF16 or FP16 operands (or split BF16: s = s_hi + s_lo,
// t = t_hi + t_lo, s*t ~ s_hi t_hi + s_hi t_lo + s_lo t_hi), accumulated in FP32 by
// v_mfma_f32_32x32x16_{bf16,f16}.  Epilogues (FP64): CrossCorr_Norm = C / sqrt(SS TT),
// SqrDistance_Norm = (SS - 2 C + TT) / sqrt(SS TT), with SS the window sum of s^2 from the
// FP64 integral of the exact F32 source and TT the FP64 sum of t^2.
//
// Contraction (Toeplitz along x).  With P the zero-padded source and template row v:
//   C[y][x0 + 32j + n] += sum_kc sum_k P[y+v][x0 + 32j + 16kc + k] * B_kc[k][n],
//   B_kc[k][n] = t(16kc + k - n, v)   (0 outside the template row)
// so one MFMA takes M = 32 output rows, N = 32 output columns and K = 16 source columns.
// KC = ceil((tw + 31) / 16) chunks cover a template row; B_kc depends only on (v, kc), so the
// KC B fragments of a template row sit in registers while a wave sweeps its NT output
// column tiles.  The A fragment at source column 16c serves every (j, kc) with 2j + kc = c.
// Issued MACs per useful MAC: 16 KC / tw (1.5 for tw = 64).
//
// Tiling: a workgroup (4 waves) owns 128 output rows x 32 NT output columns; wave w owns rows
// 32w..32w+31.  The source tile plus halo is staged in LDS as 16-bit values (two planes, hi
// and lo, for split BF16), one band of template rows at a time.  Templates wider than 97
// columns are cut into 64-column blocks (KC = 6 each), staged one block at a time.  The B
// fragments of template row v+1 are fetched from L2 while row v is multiplied.
#pragma once

#include "device_common.h"

namespace ocip {
namespace mfma {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr int kWaves = 4;               // waves per workgroup, stacked in y
constexpr int kTileRows = 32 * kWaves;  // 128
constexpr int kBlockW = 64;             // template column block for wide templates
// dynamic LDS budget: the 160 KiB of a CU less room for the kernels' static LDS
constexpr int kLdsMax = 160 * 1024 - 256;

enum { kBF16 = 1, kF16 = 2, kBF16x3 = 3 };
// kModeCC: the CrossCorr map (the single-template call; CrossCorr_Norm scales it in place
// afterwards); kModeNCC / kModeSSDN: the batch kernel's epilogues from the FP64 integral;
// kModeInt: the row-pair kernel on 8-bit integer images, exact integer sums (any of the
// four cross-term ops, TileArgs::op)
enum { kModeCC = 0, kModeNCC = 1, kModeSSDN = 2, kModeInt = 3 };

// Arguments of one output tile computation.
struct TileArgs {
   const float *src;
   int W, H, sstep;
   int tw, th, nub;
   const u16x8 *frag0, *frag1;   // B fragments (plane 1: lo parts for split BF16)
   const double *tsums;          // [0] = sum t^2
   const double *integ;          // internal exclusive FP64 integral of s^2 (Norm modes)
   float *dst;
   int dstep;
   int vb;                       // template rows per LDS band
   int vec;                      // 16-byte staging loads allowed
   // kModeInt (exact sums of 8-bit integer images): element type of source and template
   // (kSrcU8 / kSrcS8), the op (ProxOp) and the steps between the int32 flushes
   int src8 = 0, op = 0, flush = 0;
};

// Several sources of one size and template with crosscorr_mfma in one launch: the source, its
// row pitch and the output of image z < n (n = 0: the one source of TileArgs).  Passed by
// value among the kernel arguments.
constexpr int kZImages = 16;
struct ZImages {
   const float *src[kZImages];
   float *dst[kZImages];
   int sstep[kZImages];
   int n;
};

// One template of a batch (device table).
struct BatchItem {
   int tw, th, nub, vb;
   const u16x8 *frag0, *frag1;
   const double *tsums;
   float *dst;
   int dstep;
};

// Template staging of a batch (device table): B fragments and sums of one template.
struct BatchDesc {
   const char *tptr;
   int tw, th, tstep, KC, nub;
   u16x8 *frag0, *frag1;
   double *sums;
   int src8 = 0;                 // template elements: F32 (0), U8 (kSrcU8), S8 (kSrcS8)
};

// 8-bit integer element types of kModeInt (their values and products are exact in BF16 and
// FP32)
enum { kSrcU8 = 1, kSrcS8 = 2 };

// Element u of a template row as float: F32, or an 8-bit integer (exact)
__host__ __device__ inline float tmpl_elem(const char *row, int u, int src8)
{
   if (src8 == kSrcU8) return (float)reinterpret_cast<const unsigned char *>(row)[u];
   if (src8 == kSrcS8) return (float)reinterpret_cast<const signed char *>(row)[u];
   return reinterpret_cast<const float *>(row)[u];
}

__host__ __device__ constexpr int lcols_for(int KC, int NT) { return 16 * (2 * (NT - 1) + KC); }

#ifdef __HIPCC__

__device__ __forceinline__ unsigned short to_bf16(float f)
{
   return __builtin_bit_cast(unsigned short, (__bf16)f);
}
__device__ __forceinline__ float from_bf16(unsigned short h)
{
   return (float)__builtin_bit_cast(__bf16, h);
}
__device__ __forceinline__ unsigned short to_f16(float f)
{
   return __builtin_bit_cast(unsigned short, (_Float16)f);
}

// 16-bit encodings of a float for plane 0 (and plane 1 for split BF16)
template <int P> __device__ __forceinline__ void encode(float f, unsigned short &p0,
                                                        unsigned short &p1)
{
   if (P == kF16) {
      p0 = to_f16(f);
      p1 = 0;
   } else {
      p0 = to_bf16(f);
      p1 = P == kBF16x3 ? to_bf16(f - from_bf16(p0)) : (unsigned short)0;
   }
}

// B fragment f = (ub * th + v) * KC + kc of a template (row pitch tstep bytes, F32):
// lane holds B_kc[k][n] for n = lane % 32, k = 8 * (lane / 32) + i.
template <int P>
__device__ __forceinline__ void build_fragment(const char *tptr, int tstep, int tw, int th,
                                               int KC, int nub, int f, int lane, u16x8 *frag0,
                                               u16x8 *frag1)
{
   const int kc = f % KC, v = (f / KC) % th, ub = f / (KC * th);
   const int n = lane & 31, kb = 8 * (lane >> 5);
   const int u0 = nub > 1 ? ub * kBlockW : 0;
   const int wb = nub > 1 ? min(kBlockW, tw - u0) : tw;
   const float *trow = reinterpret_cast<const float *>(tptr + (size_t)v * tstep);
   u16x8 a, b;
#pragma unroll
   for (int i = 0; i < 8; ++i) {
      const int u = 16 * kc + kb + i - n;
      const float t = (u >= 0 && u < wb) ? trow[u0 + u] : 0.f;
      unsigned short h0, h1;
      encode<P>(t, h0, h1);
      a[i] = h0;
      b[i] = h1;
   }
   frag0[(size_t)f * 64 + lane] = a;
   if (P == kBF16x3) frag1[(size_t)f * 64 + lane] = b;
}

// FP64 sum of t^2 over the template into d.sums[0], by one 256-thread block: rows padded to 16
// columns, thread k takes elements k, k + 256, ... in that order (eight loads in flight), then
// the wave sums and the four waves in order.  Every prep kernel uses it, so batch and single
// calls give identical sums.
__device__ __forceinline__ void template_sqsum(const BatchDesc &d, double *red)
{
   const int pitch = (d.tw + 15) / 16 * 16;
   const int total = pitch * d.th;
   double s2 = 0;
   // thread k sums elements k, k + 256, k + 512, ... in that order, 16 loads in flight
   for (int k0 = threadIdx.x; k0 < total; k0 += 256 * 16) {
      float t[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
         const int k = k0 + 256 * i;
         const int v = k / pitch, u = k - v * pitch;
         t[i] = (k < total && u < d.tw) ? tmpl_elem(d.tptr + (size_t)v * d.tstep, u, d.src8) : 0.f;
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) s2 += (double)t[i] * (double)t[i];
   }
   s2 = wave_sum(s2);
   if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s2;
   __syncthreads();
   if (threadIdx.x == 0) d.sums[0] = red[0] + red[1] + red[2] + red[3];
}

// Single template straight from its image: one fragment per wavefront, the FP64 sum of
// squares (same order as batch_prep) in the last block, alongside.
template <int P>
__global__ __launch_bounds__(256) void single_prep(BatchDesc d)
{
   __shared__ double red[4];
   if (blockIdx.x == gridDim.x - 1) {
      template_sqsum(d, red);
      return;
   }
   const int nfrag = d.nub * d.th * d.KC;
   const int f = blockIdx.x * 4 + (threadIdx.x >> 6);
   if (f < nfrag)
      build_fragment<P>(d.tptr, d.tstep, d.tw, d.th, d.KC, d.nub, f, threadIdx.x & 63, d.frag0,
                        d.frag1);
}

template <int P>
__device__ __forceinline__ f32x16 mma(u16x8 a, u16x8 b, f32x16 c)
{
   if (P == kF16)
      return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a),
                                                    __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
   return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                  __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// The fragments are in device (global) memory.  Their pointers may come from a table in memory
// (the batch kernels), where the compiler cannot infer the address space and would emit flat
// loads; flat loads count against the LDS wait counter too, so waiting for the A fragments'
// LDS reads would also wait for these prefetches.  Reading through global-address-space
// pointers gives global loads, counted by vmcnt only.
typedef const u16x8 __attribute__((address_space(1))) *gfrag_ptr;
// The same for the output maps and the template sums, whose pointers the batch kernels also
// read from a table: stores through a flat pointer count against the LDS wait counter too, so
// an epilogue's LDS reads would wait for every earlier store of the thread.
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) float gmem_f32;
typedef __attribute__((address_space(1))) f32x4_t gmem_f32x4;
typedef __attribute__((address_space(1))) char gmem_char;
typedef __attribute__((address_space(1))) int gmem_i32;
typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) i32x4_t gmem_i32x4;
__device__ __forceinline__ gmem_f32 *out_row(const TileArgs &a, int y)
{
   return reinterpret_cast<gmem_f32 *>((gmem_char *)a.dst + (size_t)y * a.dstep);
}
__device__ __forceinline__ double tmpl_sqsum(const TileArgs &a)
{
   return *(const __attribute__((address_space(1))) double *)a.tsums;
}

template <int NPL, int KC>
__device__ __forceinline__ void load_frags(const u16x8 *__restrict__ f0, const u16x8 *__restrict__ f1,
                                           size_t base, u16x8 (&b0)[KC], u16x8 (&b1)[KC])
{
   const gfrag_ptr g0 = (gfrag_ptr)f0, g1 = (gfrag_ptr)f1;
#pragma unroll
   for (int kc = 0; kc < KC; ++kc) {
      b0[kc] = g0[base + (size_t)kc * 64];
      if (NPL == 2) b1[kc] = g1[base + (size_t)kc * 64];
   }
}

// One template row: A fragments from the LDS row `ab` (16 columns apart), KC B fragments.
template <int P, int NPL, int KC, int NT>
__device__ __forceinline__ void row_mfma(f32x16 (&acc)[NT], const unsigned short *ab,
                                         size_t plane, const u16x8 (&b0)[KC],
                                         const u16x8 (&b1)[KC])
{
   if constexpr (NPL == 1) {
      // chunk-major: consecutive MFMAs go to different accumulators; each accumulator still
      // takes its chunks in the same order
      constexpr int NA = 2 * (NT - 1) + KC;
      u16x8 a0[NA];
#pragma unroll
      for (int c = 0; c < NA; ++c) a0[c] = *reinterpret_cast<const u16x8 *>(ab + 16 * c);
#pragma unroll
      for (int kc = 0; kc < KC; ++kc)
#pragma unroll
         for (int j = 0; j < NT; ++j) acc[j] = mma<P>(a0[2 * j + kc], b0[kc], acc[j]);
   } else {
      // split BF16 (three products per chunk, registers for one A column pair at a time)
#pragma unroll
      for (int c = 0; c < 2 * (NT - 1) + KC; ++c) {
         const u16x8 a0 = *reinterpret_cast<const u16x8 *>(ab + 16 * c);
         const u16x8 a1 = *reinterpret_cast<const u16x8 *>(ab + plane + 16 * c);
#pragma unroll
         for (int j = 0; j < NT; ++j) {
            const int kc = c - 2 * j;
            if (kc < 0 || kc >= KC) continue;
            acc[j] = mma<P>(a0, b0[kc], acc[j]);
            acc[j] = mma<P>(a0, b1[kc], acc[j]);
            acc[j] = mma<P>(a1, b0[kc], acc[j]);
         }
      }
   }
}

// Stage source rows [row0, row0 + nrows) x columns [col0, col0 + LCOLS) into LDS as 16-bit
// values (zero outside the image), 8 columns per item, four items per thread in flight.
// SRC8: the source holds 8-bit integers (a.src8 gives the sign), exact in BF16.
template <int P, int LCOLS, bool SRC8 = false>
__device__ __forceinline__ void stage_rows(const TileArgs &a, unsigned short *lds, int ldsp,
                                           size_t plane, int row0, int col0, int nrows)
{
   constexpr int NPL = P == kBF16x3 ? 2 : 1;
   constexpr int groups = LCOLS / 8;
   const int W = a.W, H = a.H;
   const int total = nrows * groups;
   if constexpr (SRC8) {
      // 8-bit sources: eight items per thread in flight, each held as its raw bytes (two
      // dwords) until every load of the batch is out
      constexpr int K = 8;
      for (int it0 = threadIdx.x; it0 < total; it0 += 256 * K) {
         uint32_t lo[K], hi[K];
#pragma unroll
         for (int k = 0; k < K; ++k) {
            const int it = it0 + 256 * k;
            const int r = it / groups, g = it - r * groups;
            const int sy = row0 + r, sx = col0 + 8 * g;
            const bool rowin = it < total && sy >= 0 && sy < H;
            const unsigned char *brow = reinterpret_cast<const unsigned char *>(a.src) +
                                        (size_t)(rowin ? sy : 0) * a.sstep;
            if (a.vec && rowin && sx >= 0 && sx + 8 <= W) {
               // a.vec: 8-byte aligned rows and staged columns starting on multiples of 8
               const uint2 q = *reinterpret_cast<const uint2 *>(brow + sx);
               lo[k] = q.x;
               hi[k] = q.y;
            } else {
               lo[k] = hi[k] = 0;
#pragma unroll
               for (int i = 0; i < 8; ++i) {
                  const int x = sx + i;
                  const uint32_t b = (rowin && x >= 0 && x < W) ? brow[x] : 0u;
                  if (i < 4) lo[k] |= b << (8 * i); else hi[k] |= b << (8 * (i - 4));
               }
            }
         }
#pragma unroll
         for (int k = 0; k < K; ++k) {
            const int it = it0 + 256 * k;
            if (it >= total) break;
            const int r = it / groups, g = it - r * groups;
            u16x8 h0, h1;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
               const unsigned b = ((i < 4 ? lo[k] : hi[k]) >> (8 * (i & 3))) & 0xffu;
               const float f = a.src8 == kSrcS8 ? (float)(int)(signed char)b : (float)b;
               unsigned short e0, e1;
               encode<P>(f, e0, e1);
               h0[i] = e0;
               h1[i] = e1;
            }
            *reinterpret_cast<u16x8 *>(lds + (size_t)r * ldsp + 8 * g) = h0;
            if (NPL == 2) *reinterpret_cast<u16x8 *>(lds + plane + (size_t)r * ldsp + 8 * g) = h1;
         }
      }
      return;
   }
   for (int it0 = threadIdx.x; it0 < total; it0 += 256 * 4) {
      float f[4][8];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
         const int it = it0 + 256 * k;
         const int r = it / groups, g = it - r * groups;
         const int sy = row0 + r, sx = col0 + 8 * g;
         const bool rowin = it < total && sy >= 0 && sy < H;
         const float *srow = reinterpret_cast<const float *>(
             reinterpret_cast<const char *>(a.src) + (size_t)(rowin ? sy : 0) * a.sstep);
         if (a.vec && rowin && sx >= 0 && sx + 8 <= W) {
            const float4 q0 = *reinterpret_cast<const float4 *>(srow + sx);
            const float4 q1 = *reinterpret_cast<const float4 *>(srow + sx + 4);
            f[k][0] = q0.x; f[k][1] = q0.y; f[k][2] = q0.z; f[k][3] = q0.w;
            f[k][4] = q1.x; f[k][5] = q1.y; f[k][6] = q1.z; f[k][7] = q1.w;
         } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
               const int x = sx + i;
               f[k][i] = (rowin && x >= 0 && x < W) ? srow[x] : 0.f;
            }
         }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
         const int it = it0 + 256 * k;
         if (it >= total) break;
         const int r = it / groups, g = it - r * groups;
         u16x8 h0, h1;
#pragma unroll
         for (int i = 0; i < 8; ++i) {
            unsigned short e0, e1;
            encode<P>(f[k][i], e0, e1);
            h0[i] = e0;
            h1[i] = e1;
         }
         *reinterpret_cast<u16x8 *>(lds + (size_t)r * ldsp + 8 * g) = h0;
         if (NPL == 2) *reinterpret_cast<u16x8 *>(lds + plane + (size_t)r * ldsp + 8 * g) = h1;
      }
   }
}

// One 128 x 32NT output tile at tile coordinates (bx, by).  ldsp: LDS row pitch in 16-bit
// elements.  Starts with a barrier, so consecutive calls may reuse the LDS.
template <int P, int KC, int NT, int MODE>
__device__ __forceinline__ void mfma_tile(const TileArgs &a, int bx, int by, unsigned short *lds,
                                          int ldsp)
{
   constexpr int NPL = P == kBF16x3 ? 2 : 1;
   constexpr int lcols = lcols_for(KC, NT);
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int x0 = bx * 32 * NT, y0 = by * kTileRows;
   const int cx = a.tw / 2, cy = a.th / 2;
   const int W = a.W, H = a.H;

   f32x16 acc[NT];
#pragma unroll
   for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[j][i] = 0.f;

   // template column blocks (one block unless tw > 97), then bands of template rows
   for (int ub = 0; ub < a.nub; ++ub) {
      for (int v0 = 0; v0 < a.th; v0 += a.vb) {
         const int nv = min(a.vb, a.th - v0);
         const int nrows = kTileRows + nv - 1;
         const size_t plane = (size_t)nrows * ldsp;
         __syncthreads();
         stage_rows<P, lcols>(a, lds, ldsp, plane, y0 - cy + v0, x0 - cx + ub * kBlockW, nrows);
         __syncthreads();

         // ping-pong: the B fragments of template row v+1 load while row v multiplies
         const size_t fb0 = ((size_t)(ub * a.th + v0) * KC) * 64 + lane;
         u16x8 pa0[KC], pa1[KC], pb0[KC], pb1[KC];
         load_frags<NPL, KC>(a.frag0, a.frag1, fb0, pa0, pa1);
         const unsigned short *arow = lds + (size_t)(wave * 32 + (lane & 31)) * ldsp + 8 * (lane >> 5);
         int v = 0;
         for (; v + 1 < nv; v += 2) {
            // keep each prefetch ahead of the MFMAs of the row before it
            load_frags<NPL, KC>(a.frag0, a.frag1, fb0 + (size_t)(v + 1) * KC * 64, pb0, pb1);
            __builtin_amdgcn_sched_barrier(0);
            row_mfma<P, NPL, KC, NT>(acc, arow + (size_t)v * ldsp, plane, pa0, pa1);
            const int vn = v + 2 < nv ? v + 2 : v + 1;
            load_frags<NPL, KC>(a.frag0, a.frag1, fb0 + (size_t)vn * KC * 64, pa0, pa1);
            __builtin_amdgcn_sched_barrier(0);
            row_mfma<P, NPL, KC, NT>(acc, arow + (size_t)(v + 1) * ldsp, plane, pb0, pb1);
         }
         if (v < nv) row_mfma<P, NPL, KC, NT>(acc, arow + (size_t)v * ldsp, plane, pa0, pa1);
      }
   }

   // C fragment: lane holds column n = lane % 32 of each n-tile, rows
   // m = 8 * (i / 4) + 4 * (lane / 32) + i % 4.
   const int n = lane & 31, mb = 4 * (lane >> 5);
   const int ybase = y0 + wave * 32 + mb;
   double tt = 0.0;
   if (MODE == kModeNCC || MODE == kModeSSDN) tt = tmpl_sqsum(a);
#pragma unroll
   for (int j = 0; j < NT; ++j) {
      const int x = x0 + 32 * j + n;
      float val[16];
      if (MODE != kModeCC) {
         // four outputs (16 integral corners) at a time keeps the epilogue within the
         // registers of two waves per SIMD
#pragma unroll
         for (int i0 = 0; i0 < 16; i0 += 4) {
            double ss[4];
#pragma unroll
            for (int i = i0; i < i0 + 4; ++i)
               ss[i - i0] = window_sqsum(a.integ, W, H, x, ybase + 8 * (i / 4) + (i % 4), a.tw, a.th);
#pragma unroll
            for (int i = i0; i < i0 + 4; ++i) {
               const double num = MODE == kModeNCC ? (double)acc[j][i]
                                                   : ss[i - i0] - 2.0 * (double)acc[j][i] + tt;
               val[i] = (float)num * inv_sqrt(ss[i - i0] * tt);
            }
         }
      } else {
#pragma unroll
         for (int i = 0; i < 16; ++i) val[i] = acc[j][i];
      }
      if (x >= W) continue;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
         const int y = ybase + 8 * (i / 4) + (i % 4);
         if (y < H)
            out_row(a, y)[x] = val[i];
      }
   }
}

// LDS use limits these kernels to two workgroups per CU (two waves per SIMD), so the compiler
// may spend registers on keeping two template rows of B fragments in flight.
// One-dimensional grid of tiles_x * tiles_y workgroups per image in XCD-aware order (images
// z.n of them, one after the other, or the one image of `a`).
template <int P, int KC, int NT, int MODE>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(NT == 2 ? 3 : 1, NT == 2 ? 3 : 2)))
void crosscorr_mfma(TileArgs a, int ldsp, int tiles_x, int tiles, ZImages z)
{
   extern __shared__ u16x8 lds_v[];
   const int zi = blockIdx.x / tiles;
   const int t = xcd_tile(blockIdx.x - zi * tiles, tiles);
   if (z.n) {
      a.src = z.src[zi];
      a.dst = z.dst[zi];
      a.sstep = z.sstep[zi];
   }
   mfma_tile<P, KC, NT, MODE>(a, t % tiles_x, t / tiles_x, reinterpret_cast<unsigned short *>(lds_v),
                              ldsp);
}

// Persistent batch kernel: workgroups take (template, tile) items from `counter` until
// `nitems * tiles` items are handed out.  Items are ordered template-major with the
// templates sorted by decreasing cost, so the largest templates start first.  Registers are
// held to two waves per SIMD, the two workgroups per CU that the LDS allows.
template <int P, int KC, int NT, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
void crosscorr_mfma_batch(const float *src, int W, int H, int sstep, int vec,
                          const double *integ, const BatchItem *__restrict__ items, int nitems,
                          int tiles_x, int tiles_y, unsigned *counter, int ldsp)
{
   extern __shared__ u16x8 lds_v[];
   __shared__ int s_item;
   const int tiles = tiles_x * tiles_y;
   const int total = nitems * tiles;
   for (;;) {
      __syncthreads();
      if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
      __syncthreads();
      // uniform across the workgroup: the item's table entry and every address derived from
      // it stay in scalar registers
      const int it = __builtin_amdgcn_readfirstlane(s_item);
      if (it >= total) break;
      const BatchItem b = items[it / tiles];
      const int tile = it - (it / tiles) * tiles;
      // 16-byte staging loads: aligned rows (vec) and tw/2 a multiple of 4
      TileArgs a{src, W, H, sstep, b.tw, b.th, b.nub, b.frag0, b.frag1, b.tsums, integ,
                 b.dst, b.dstep, b.vb, (vec && (b.tw / 2) % 4 == 0) ? 1 : 0};
      mfma_tile<P, KC, NT, MODE>(a, tile % tiles_x, tile / tiles_x,
                                 reinterpret_cast<unsigned short *>(lds_v), ldsp);
   }
}

// ---------------------------------------------------------------------------------------
// Row-pair variant (one 16-bit plane: BF16 or FP16).
//
// The 32 MFMA columns n = 16 dy + dx hold 16 output columns dx for two output rows one
// apart.  Wave w of a workgroup whose outputs are rows [Y0, Y0 + 127) accumulates, in MFMA
// row m and column half dy, output row y = Y0 - 1 + 32 w + m + dy.  At step s its A rows are
// source rows Y0 - cy + 32 w + m + 2 s, so one A fragment feeds template row 2 s + 1 to half
// dy = 0 and template row 2 s to half dy = 1 (rows at or past th are zero):
//   B_kc[k][16 dy + dx] = t(16 kc + k - dx, 2 s + 1 - dy),
// with KC = ceil((tw + 15) / 16) chunks of 16 source columns for 16 output columns.  After
// the last step, half dy = 1 of row m and half dy = 0 of row m + 1 hold the even and the odd
// template rows of one output row; their sum goes through LDS.  The 4 x 32 rows of each
// half pair up into 127 complete output rows per workgroup.  Issued MACs per useful MAC:
// 16 KC / tw x 128 / 127 (1.26 for tw = 64, against 1.5 for the kernel above), with
// NT + KC - 1 A fragments and KC B fragments per KC NT MFMAs.
constexpr int kPairRows = kTileRows - 1;   // 127 output rows per workgroup

__host__ __device__ constexpr int pair_lcols(int KC, int NT) { return 16 * (NT + KC - 1); }
__host__ __device__ constexpr int pair_tile_pitch(int NT) { return 16 * NT + 4; }

// Steps per template column block in the fragment layout of the row-pair kernels: ceil(th / 2),
// rounded up to even for one-block templates (the strip kernel's step count; the extra step
// of an odd count has all-zero fragments).
__host__ __device__ constexpr int pair_frag_steps(int th, int nub)
{
   return nub > 1 ? (th + 1) / 2 : ((th + 1) / 2 + 1) / 2 * 2;
}

// B fragment f = (ub * S + s) * KC + kc, S = pair_frag_steps(th, nub): lane holds B_kc[k][n]
// for n = lane % 32, k = 8 * (lane / 32) + i.
template <int P>
__device__ __forceinline__ void build_pair_fragment(const char *tptr, int tstep, int tw, int th,
                                                    int KC, int nub, int f, int lane, u16x8 *frag,
                                                    u16x8 *frag1, int src8)
{
   const int S = pair_frag_steps(th, nub);
   const int kc = f % KC, s = (f / KC) % S, ub = f / (KC * S);
   const int n = lane & 31, dy = n >> 4, dx = n & 15, kb = 8 * (lane >> 5);
   const int v = 2 * s + 1 - dy;
   const int u0 = nub > 1 ? ub * kBlockW : 0;
   const int wb = nub > 1 ? min(kBlockW, tw - u0) : tw;
   const char *trow = tptr + (size_t)(v < th ? v : 0) * tstep;
   u16x8 h, l;
#pragma unroll
   for (int i = 0; i < 8; ++i) {
      const int u = 16 * kc + kb + i - dx;
      const float t = (v < th && u >= 0 && u < wb) ? tmpl_elem(trow, u0 + u, src8) : 0.f;
      unsigned short h0, h1;
      encode<P>(t, h0, h1);
      h[i] = h0;
      l[i] = h1;
   }
   frag[(size_t)f * 64 + lane] = h;
   if (P == kBF16x3) frag1[(size_t)f * 64 + lane] = l;   // split BF16: the lo plane
}

// Fragments of the row-pair kernel, one per wavefront, the FP64 sum of squares (same order
// as single_prep) in the last block, alongside.
template <int P>
__global__ __launch_bounds__(256) void pair_prep(BatchDesc d)
{
   __shared__ double red[4];
   if (blockIdx.x == gridDim.x - 1) {
      template_sqsum(d, red);
      return;
   }
   const int nfrag = d.nub * pair_frag_steps(d.th, d.nub) * d.KC;
   const int f = blockIdx.x * 4 + (threadIdx.x >> 6);
   if (f < nfrag)
      build_pair_fragment<P>(d.tptr, d.tstep, d.tw, d.th, d.KC, d.nub, f, threadIdx.x & 63,
                             d.frag0, d.frag1, d.src8);
}

// Batch: for template blockIdx.y, workgroups 0 .. gridDim.x - 2 build its fragments (in the
// row-pair layout when PAIR), the last one its FP64 sum of squares.
constexpr int kPrepBlocks = 17;
template <int P, bool PAIR>
__global__ __launch_bounds__(256) void batch_prep(const BatchDesc *__restrict__ descs)
{
   __shared__ double red[4];
   const BatchDesc d = descs[blockIdx.y];
   if (blockIdx.x == gridDim.x - 1) {
      template_sqsum(d, red);
      return;
   }
   const int nfrag = d.nub * (PAIR ? pair_frag_steps(d.th, d.nub) : d.th) * d.KC;
   for (int k = blockIdx.x * 256 + threadIdx.x; k < nfrag * 64; k += (gridDim.x - 1) * 256) {
      if constexpr (PAIR)
         build_pair_fragment<P>(d.tptr, d.tstep, d.tw, d.th, d.KC, d.nub, k >> 6, k & 63,
                                d.frag0, d.frag1, d.src8);
      else
         build_fragment<P>(d.tptr, d.tstep, d.tw, d.th, d.KC, d.nub, k >> 6, k & 63, d.frag0,
                           d.frag1);
   }
}

// One step: NT + KC - 1 A fragments from the LDS row `ab`, KC B fragments.  Split BF16: the
// lo planes of A (at `plane` past the hi plane) and B too, three products per chunk (hi hi,
// hi lo, lo hi, in that order for each accumulator), consecutive MFMAs on different
// accumulators.
template <int P, int KC, int NT>
__device__ __forceinline__ void pair_mfma(f32x16 (&acc)[NT], const unsigned short *ab,
                                          size_t plane, const u16x8 (&b)[KC],
                                          const u16x8 (&b1)[KC])
{
   constexpr int NA = NT + KC - 1;
   u16x8 a0[NA];
#pragma unroll
   for (int c = 0; c < NA; ++c) a0[c] = *reinterpret_cast<const u16x8 *>(ab + 16 * c);
   if constexpr (P != kBF16x3) {
#pragma unroll
      for (int kc = 0; kc < KC; ++kc)
#pragma unroll
         for (int j = 0; j < NT; ++j) acc[j] = mma<P>(a0[j + kc], b[kc], acc[j]);
   } else {
      u16x8 a1[NA];
#pragma unroll
      for (int c = 0; c < NA; ++c) a1[c] = *reinterpret_cast<const u16x8 *>(ab + plane + 16 * c);
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) {
#pragma unroll
         for (int j = 0; j < NT; ++j) acc[j] = mma<P>(a0[j + kc], b[kc], acc[j]);
#pragma unroll
         for (int j = 0; j < NT; ++j) acc[j] = mma<P>(a0[j + kc], b1[kc], acc[j]);
#pragma unroll
         for (int j = 0; j < NT; ++j) acc[j] = mma<P>(a1[j + kc], b[kc], acc[j]);
      }
   }
}

// Output tile (bx, by): 127 rows x 16 NT columns; a.vb = steps (template row pairs) per LDS
// band.  Starts with a barrier, so consecutive calls may reuse the LDS.  kModeCC stores the
// CrossCorr map; kModeNCC / kModeSSDN store CrossCorr_Norm / SqrDistance_Norm with the window
// sums of s^2 from the FP64 integral a.integ (the batch kernel's epilogue, the same
// expressions as mfma_tile).
//
// kModeInt (8-bit integer images, BF16 operands): every value and product is an integer that
// BF16 and FP32 hold exactly, and so is every partial sum below 2^24.  Each step adds one
// template row to each accumulator, at most tw max|s t| in magnitude, so after a.flush steps
// (a.flush tw max|s t| < 2^24, set by the host) the FP32 accumulators are converted to and
// added into int32 accumulators and cleared: the sums are exact (below 2^31, also checked by
// the host).  The epilogue pairs the int32 halves and forms the op from the exact cross term
// (and the window sum of squares, exact in int32, which ss_box8 wrote into the map before)
// with the same FP64 expressions as prox_tile_int, so the maps equal that kernel's bit for bit.
template <int P, int KC, int NT, int MODE>
__device__ __forceinline__ void pair_tile(const TileArgs &a, int bx, int by, unsigned short *lds,
                                          int ldsp)
{
   static_assert(MODE == kModeCC || MODE == kModeNCC || MODE == kModeSSDN ||
                     (MODE == kModeInt && P == kBF16 && NT == 4),
                 "float maps, or exact integer sums (BF16, 4 n-tiles)");
   constexpr bool INT = MODE == kModeInt;
   // operand planes: two for split BF16 (hi and lo), the lo plane of the staged rows at
   // `plane` 16-bit elements past the hi plane (the rows a band of a.vb steps stages)
   constexpr int NPL = P == kBF16x3 ? 2 : 1;
   const size_t plane = NPL == 2 ? (size_t)(kTileRows - 2 + 2 * a.vb) * ldsp : 0;
   constexpr int lcols = pair_lcols(KC, NT);
   constexpr int C = 16 * NT;
   constexpr int TP = pair_tile_pitch(NT);
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int x0 = bx * C, Y0 = by * kPairRows;
   const int cx = a.tw / 2, cy = a.th / 2;
   const int S = (a.th + 1) / 2;

   f32x16 acc[NT];
#pragma unroll
   for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[j][i] = 0.f;
   // kModeInt: the int32 sums, and the steps since the last flush
   int iacc[INT ? NT : 1][16];
   int since = 0;
   auto flush = [&]() {
      if constexpr (INT) {
#pragma unroll
         for (int j = 0; j < NT; ++j)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
               iacc[j][i] += (int)acc[j][i];
               acc[j][i] = 0.f;
            }
         since = 0;
      }
   };
   auto stepped = [&]() {
      if constexpr (INT) {
         if (++since == a.flush) flush();
      }
   };
   if constexpr (INT) {
#pragma unroll
      for (int j = 0; j < NT; ++j)
#pragma unroll
         for (int i = 0; i < 16; ++i) iacc[j][i] = 0;
   }

   for (int ub = 0; ub < a.nub; ++ub) {
      for (int s0 = 0; s0 < S; s0 += a.vb) {
         const int ns = min(a.vb, S - s0);
         __syncthreads();
         // LDS row r = source row Y0 - cy + 2 s0 + r; wave w reads rows 32 w + m + 2 (s - s0)
         stage_rows<P, lcols, INT>(a, lds, ldsp, plane, Y0 - cy + 2 * s0, x0 - cx + ub * kBlockW,
                                   kTileRows - 2 + 2 * ns);
         __syncthreads();

         // a wave whose output rows (from Y0 - 1 + 32 wave on) all lie below the image skips
         // its products: its half rows are never stored
         if (Y0 - 1 + 32 * wave >= a.H) continue;
         // ping-pong: the B fragments of step s+1 load while step s multiplies
         const size_t fb0 = ((size_t)(ub * pair_frag_steps(a.th, a.nub) + s0) * KC) * 64 + lane;
         u16x8 pa[KC], pb[KC], pa1[KC], pb1[KC];   // (pa1, pb1: the lo planes, split BF16)
         load_frags<NPL, KC>(a.frag0, a.frag1, fb0, pa, pa1);
         const unsigned short *arow = lds + (size_t)(wave * 32 + (lane & 31)) * ldsp + 8 * (lane >> 5);
         int q = 0;
         for (; q + 1 < ns; q += 2) {
            load_frags<NPL, KC>(a.frag0, a.frag1, fb0 + (size_t)(q + 1) * KC * 64, pb, pb1);
            __builtin_amdgcn_sched_barrier(0);
            pair_mfma<P, KC, NT>(acc, arow + (size_t)(2 * q) * ldsp, plane, pa, pa1);
            stepped();
            const int qn = q + 2 < ns ? q + 2 : q + 1;
            load_frags<NPL, KC>(a.frag0, a.frag1, fb0 + (size_t)qn * KC * 64, pa, pa1);
            __builtin_amdgcn_sched_barrier(0);
            pair_mfma<P, KC, NT>(acc, arow + (size_t)(2 * q + 2) * ldsp, plane, pb, pb1);
            stepped();
         }
         if (q < ns) {
            pair_mfma<P, KC, NT>(acc, arow + (size_t)(2 * q) * ldsp, plane, pa, pa1);
            stepped();
         }
      }
   }
   flush();

   // Norm modes: groups of four consecutive outputs of one row per thread (kNormIters of them,
   // kEpiGroups per batch).  A group's window sums of s^2 are differences of the FP64
   // integral at rows ya, yb (shared by the four) and columns xa + e, xb + e, all clamped to
   // the image, summed in the order of rect_sum.  The loads of the first batch go out before
   // the halves are paired in LDS, and those of each next batch before a batch is used.
   constexpr int kNormGroups = kPairRows * (C / 4);
   constexpr int kNormIters = (kNormGroups + 255) / 256;
   constexpr int kEpiGroups = 2;
   constexpr bool NORM = MODE == kModeNCC || MODE == kModeSSDN;
   typedef double Corners[kEpiGroups][4][4];   // [group][corner][e]
   auto load_corners = [&](int i0, Corners &q) {
      const size_t ip = excl_pitch(a.W);
#pragma unroll
      for (int u = 0; u < kEpiGroups; ++u) {
         const int k = threadIdx.x + 256 * (i0 + u);
         const int r = k / (C / 4), x = x0 + 4 * (k - r * (C / 4));
         const int y = Y0 + r;
         const int ya = min(max(y - cy, 0), a.H), yb = min(max(y - cy + a.th, 0), a.H);
         const double *r0 = a.integ + (size_t)ya * ip + kExclOffset;
         const double *r1 = a.integ + (size_t)yb * ip + kExclOffset;
#pragma unroll
         for (int e = 0; e < 4; ++e) {
            const int xa = min(max(x + e - cx, 0), a.W), xb = min(max(x + e - cx + a.tw, 0), a.W);
            q[u][0][e] = r0[xa];
            q[u][1][e] = r1[xa];
            q[u][2][e] = r0[xb];
            q[u][3][e] = r1[xb];
         }
      }
   };
   Corners qa, qb;
   if constexpr (NORM) load_corners(0, qa);
   // kModeInt: the window sums of s^2 that ss_box8 left in the map (all ops but CrossCorr),
   // the thread's groups of four outputs in the epilogue's order, loaded before the halves
   // are paired so that their latency passes under it
   constexpr int kIntIters = (kPairRows * (C / 4) + 255) / 256;
   int4 ssq[INT ? kIntIters : 1];
   if constexpr (INT) {
      const bool v4 = a.dstep % 16 == 0 && reinterpret_cast<uintptr_t>(a.dst) % 16 == 0;
#pragma unroll
      for (int i = 0; i < kIntIters; ++i) {
         const int k = threadIdx.x + 256 * i;
         const int r = k / (C / 4), c = 4 * (k - r * (C / 4));
         const int y = Y0 + r, x = x0 + c;
         ssq[i] = make_int4(0, 0, 0, 0);
         if (a.op == kCC || k >= kPairRows * (C / 4) || y >= a.H || x >= a.W) continue;
         const gmem_i32 *oi = reinterpret_cast<const gmem_i32 *>(out_row(a, y) + x);
         if (v4 && x + 4 <= a.W) {
            const i32x4_t q = *reinterpret_cast<const gmem_i32x4 *>(oi);
            ssq[i] = make_int4(q[0], q[1], q[2], q[3]);
         } else {
            int e4[4] = {0, 0, 0, 0};
#pragma unroll
            for (int e = 0; e < 4; ++e)
               if (x + e < a.W) e4[e] = oi[e];
            ssq[i] = make_int4(e4[0], e4[1], e4[2], e4[3]);
         }
      }
   }

   // Pair the halves in LDS: tile row r = 32 w + m + dy holds output row Y0 - 1 + r.
   // C fragment: lane holds column n = lane % 32 of each n-tile, rows
   // m = 8 * (i / 4) + 4 * (lane / 32) + i % 4.  (kModeInt: int32 halves, exact sums.)
   float *tile = reinterpret_cast<float *>(lds);
   int *itile = reinterpret_cast<int *>(lds);
   const int n = lane & 31, dy = n >> 4, dx = n & 15;
   const int rb = 32 * wave + 4 * (lane >> 5) + dy;
   __syncthreads();
   if (dy) {
#pragma unroll
      for (int j = 0; j < NT; ++j)
#pragma unroll
         for (int i = 0; i < 16; ++i) {
            const int at = (rb + 8 * (i / 4) + (i % 4)) * TP + 16 * j + dx;
            if constexpr (INT) itile[at] = iacc[j][i]; else tile[at] = acc[j][i];
         }
   }
   __syncthreads();
   if (!dy) {
#pragma unroll
      for (int j = 0; j < NT; ++j)
#pragma unroll
         for (int i = 0; i < 16; ++i) {
            const int at = (rb + 8 * (i / 4) + (i % 4)) * TP + 16 * j + dx;
            if constexpr (INT) itile[at] = iacc[j][i] + itile[at];
            else tile[at] = acc[j][i] + tile[at];
         }
   }
   __syncthreads();
   if constexpr (INT) {
      // each output's op from its exact cross term CC (and for all but CrossCorr the window's
      // sum of squares SS, an exact int32 that ss_box8 left in the map), as prox_tile_int forms
      // it; four consecutive columns per thread, one 16-byte store where the row allows
      const double tt = a.tsums[0];
      const bool v4 = a.dstep % 16 == 0 && reinterpret_cast<uintptr_t>(a.dst) % 16 == 0;
#pragma unroll
      for (int i = 0; i < kIntIters; ++i) {
         const int k = threadIdx.x + 256 * i;
         const int r = k / (C / 4), c = 4 * (k - r * (C / 4));
         const int y = Y0 + r, x = x0 + c;
         if (k >= kPairRows * (C / 4) || y >= a.H || x >= a.W) continue;
         const int4 cc = *reinterpret_cast<const int4 *>(&itile[(r + 1) * TP + c]);
         const int ci[4] = {cc.x, cc.y, cc.z, cc.w};
         gmem_f32 *o = out_row(a, y) + x;
         const int si[4] = {ssq[i].x, ssq[i].y, ssq[i].z, ssq[i].w};
         float v[4];
#pragma unroll
         for (int e = 0; e < 4; ++e) {
            if (a.op == kCC) {
               v[e] = (float)ci[e];
            } else {
               const double ss = (double)si[e];
               const double cd = (double)ci[e];
               const double d = (a.op == kSSD || a.op == kSSDNorm) ? ss - 2.0 * cd + tt : cd;
               v[e] = a.op == kSSD ? (float)d : (float)(d / sqrt(ss * tt));
            }
         }
         if (v4 && x + 4 <= a.W) {
            *reinterpret_cast<gmem_f32x4 *>(o) = f32x4_t{v[0], v[1], v[2], v[3]};
         } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
               if (x + e < a.W) o[e] = v[e];
         }
      }
      return;
   }
   // rows 1..127 are whole.  16-byte aligned output rows: four consecutive columns per
   // thread (the tile pitch keeps them 16-byte aligned in LDS too), one at a time at the
   // right edge; otherwise consecutive threads take consecutive columns
   static_assert(TP % 4 == 0 && C % 4 == 0, "16-byte groups of the LDS tile");
   if constexpr (NORM) {
      const double tt = tmpl_sqsum(a);
      const bool v4 = a.dstep % 16 == 0 && reinterpret_cast<uintptr_t>(a.dst) % 16 == 0;
      auto finish = [&](int i0, const Corners &q) {
#pragma unroll
         for (int u = 0; u < kEpiGroups; ++u) {
            const int k = threadIdx.x + 256 * (i0 + u);
            const int r = k / (C / 4), c = 4 * (k - r * (C / 4));
            const int y = Y0 + r, x = x0 + c;
            if (k >= kNormGroups || y >= a.H || x >= a.W) continue;
            const float4 t4 = *reinterpret_cast<const float4 *>(&tile[(r + 1) * TP + c]);
            const float cc[4] = {t4.x, t4.y, t4.z, t4.w};
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
               const double ss = (q[u][3][e] - q[u][2][e]) - (q[u][1][e] - q[u][0][e]);
               const double num = MODE == kModeNCC ? (double)cc[e] : ss - 2.0 * (double)cc[e] + tt;
               v[e] = (float)num * inv_sqrt(ss * tt);
            }
            gmem_f32 *o = out_row(a, y) + x;
            if (v4 && x + 4 <= a.W) {
               *reinterpret_cast<gmem_f32x4 *>(o) = f32x4_t{v[0], v[1], v[2], v[3]};
            } else {
#pragma unroll
               for (int e = 0; e < 4; ++e)
                  if (x + e < a.W) o[e] = v[e];
            }
         }
      };
      // two register sets alternate: the loads of batch i0 + kEpiGroups go out before batch i0
      // is finished
#pragma unroll
      for (int i0 = 0; i0 < kNormIters; i0 += 2 * kEpiGroups) {
         if (i0 + kEpiGroups < kNormIters) load_corners(i0 + kEpiGroups, qb);
         finish(i0, qa);
         if (i0 + kEpiGroups < kNormIters) {
            if (i0 + 2 * kEpiGroups < kNormIters) load_corners(i0 + 2 * kEpiGroups, qa);
            finish(i0 + kEpiGroups, qb);
         }
      }
      return;
   }
   if (a.dstep % 16 == 0 && reinterpret_cast<uintptr_t>(a.dst) % 16 == 0) {
      for (int k = threadIdx.x; k < kPairRows * (C / 4); k += 256) {
         const int r = k / (C / 4), c = 4 * (k - r * (C / 4));
         const int y = Y0 + r, x = x0 + c;
         if (y >= a.H || x >= a.W) continue;
         gmem_f32 *o = out_row(a, y) + x;
         const float4 q = *reinterpret_cast<const float4 *>(&tile[(r + 1) * TP + c]);
         if (x + 4 <= a.W) {
            *reinterpret_cast<gmem_f32x4 *>(o) = f32x4_t{q.x, q.y, q.z, q.w};
         } else {
            o[0] = q.x;
            if (x + 1 < a.W) o[1] = q.y;
            if (x + 2 < a.W) o[2] = q.z;
         }
      }
   } else {
      for (int k = threadIdx.x; k < kPairRows * C; k += 256) {
         const int r = k / C, c = k - r * C;
         const int y = Y0 + r, x = x0 + c;
         if (y >= a.H || x >= a.W) continue;
         out_row(a, y)[x] = tile[(r + 1) * TP + c];
      }
   }
}

// Waves per SIMD the row-pair kernel's registers are sized for: 3 for the narrow one-plane
// CrossCorr tile, 1 for the widest split-BF16 tiles (whose two staged planes leave LDS for one
// workgroup per CU at these template widths, and whose three fragment sets would spill at 2).
__host__ __device__ constexpr int pair_waves(int P, int KC, int NT, int MODE)
{
   return NT <= 4 && MODE == kModeCC && P != kBF16x3 ? 3 : P == kBF16x3 && NT == 6 && KC >= 7 ? 1 : 2;
}

// One 127 x 16 NT tile per workgroup, XCD-aware order (the single-template call): tiles_x *
// tiles_y workgroups per image, images z.n of them one after the other (or the one image of
// `a`).
template <int P, int KC, int NT, int MODE>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(pair_waves(P, KC, NT, MODE), pair_waves(P, KC, NT, MODE))))
void crosscorr_mfma_pair(TileArgs a, int ldsp, int tiles_x, int tiles, ZImages z)
{
   extern __shared__ u16x8 lds_v[];
   const int zi = blockIdx.x / tiles;
   const int t = xcd_tile(blockIdx.x - zi * tiles, tiles);
   if (z.n) {
      a.src = z.src[zi];
      a.dst = z.dst[zi];
      a.sstep = z.sstep[zi];
   }
   pair_tile<P, KC, NT, MODE>(a, t % tiles_x, t / tiles_x, reinterpret_cast<unsigned short *>(lds_v),
                              ldsp);
}

// Persistent batch kernel on the row-pair tile (the batch of BF16 / FP16 templates).  The
// tiles (row-major) are cut into kXcds * kSubBands bands of consecutive tiles, kSubBands per
// XCD, each band with its own work-queue counter (counter[band]) over (template, tile of the
// band) items, template-major with the templates by decreasing cost.  Workgroup b works
// through the bands of XCD b % kXcds (the XCD that round-robin dispatch puts it on) in order,
// so that an XCD's L2 holds the source rows, integral rows and template fragments of one
// band's few current templates; then it takes items of the other XCDs' bands.  Two
// workgroups per CU.
constexpr int kSubBands = 4;
constexpr int kPairBatchBands = kXcds * kSubBands;

template <int P, int KC, int NT, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
void crosscorr_mfma_pair_batch(const float *src, int W, int H, int sstep, int vec,
                               const double *integ, const BatchItem *__restrict__ items,
                               int nitems, int tiles_x, int tiles_y, unsigned *counter, int ldsp)
{
   extern __shared__ u16x8 lds_v[];
   __shared__ int s_item;
   const int tiles = tiles_x * tiles_y;
   const int home = (int)(blockIdx.x % kXcds) * kSubBands;
   for (int rr = 0; rr < kPairBatchBands; ++rr) {
      const int band = (home + rr) % kPairBatchBands;
      const int t0 = (int)((long long)tiles * band / kPairBatchBands);
      const int bt = (int)((long long)tiles * (band + 1) / kPairBatchBands) - t0;
      const int total = nitems * bt;
      for (;;) {
         __syncthreads();
         if (threadIdx.x == 0) s_item = total > 0 ? (int)atomicAdd(counter + band, 1u) : 0;
         __syncthreads();
         // uniform across the workgroup: the item's table entry and every address derived
         // from it stay in scalar registers
         const int it = __builtin_amdgcn_readfirstlane(s_item);
         if (it >= total) break;
         const BatchItem b = items[it / bt];
         const int tile = t0 + (it - (it / bt) * bt);
         // 16-byte staging loads: aligned rows (vec) and tw/2 a multiple of 4
         TileArgs a{src, W, H, sstep, b.tw, b.th, b.nub, b.frag0, nullptr, b.tsums, integ,
                    b.dst, b.dstep, b.vb, (vec && (b.tw / 2) % 4 == 0) ? 1 : 0};
         pair_tile<P, KC, NT, MODE>(a, tile % tiles_x, tile / tiles_x,
                                    reinterpret_cast<unsigned short *>(lds_v), ldsp);
      }
   }
}

// ---------------------------------------------------------------------------------------
// Strip variant of the row-pair kernel: persistent workgroups, each taking a contiguous run of
// 127 x 16 NT tiles down a strip of 16 NT output columns, with the source rows streamed
// through an LDS ring of kRingRows rows.
//
// Ring row g (slot g % kRingRows) holds source row Y00 - cy + g, Y00 the first tile's top
// output row; tile k of the run reads ring rows 127 k + 32 w + m + 2 s at step s.  Rows
// [0, kRingFill) are staged before the first step; after that every step (one barrier each)
// first writes to LDS the rows whose loads were issued kRingDepth steps before, then issues
// loads for at most q_rows further rows, never past the capacity limit (the first row still
// read at this step) + kRingRows.  The host accepts (S, q_rows) only when a simulation of this
// schedule shows that every row read at a step was written at an earlier step.  Every source
// row of a run is loaded once, and the loads of the next tile's rows overlap this tile's
// MFMAs.  The two halves of a tile's C fragments are paired with lane shuffles; the top row
// of waves 1..3 takes the bottom partial row of the wave above through a small LDS buffer,
// read after the next barrier.
constexpr int kRingRows = 256;
constexpr int kRingDepth = 2;
constexpr int kRingFill = kTileRows + 2 * kRingDepth;   // 132

// ring loads of one step: one 8-column item per thread
struct RingBatch {
   float f[8];
   int lo, cnt;
};

// Offset that puts a buffer load out of range (the hardware returns 0), for positions outside
// the image; the strip kernel runs on images of at most kRingOob bytes.
constexpr unsigned kRingOob = 0x40000000u;

// Byte offsets of the 8 columns of this thread's ring item (column group threadIdx.x % groups,
// first column col0 + 8 g): VEC, two 16-byte groups, each wholly inside the image or wholly
// kRingOob (the caller guarantees that no group straddles an image edge); otherwise one per
// column.
template <int LCOLS, bool VEC>
__device__ __forceinline__ void ring_columns(const TileArgs &a, int col0, unsigned (&coff)[VEC ? 2 : 8])
{
   constexpr int groups = LCOLS / 8;
   const int sx = col0 + 8 * (int)(threadIdx.x % groups);
#pragma unroll
   for (int i = 0; i < (VEC ? 2 : 8); ++i) {
      const int x = sx + (VEC ? 4 : 1) * i;
      coff[i] = (x >= 0 && x < a.W) ? (unsigned)(4 * x) : kRingOob;
   }
}

// Ring loads of rows [lo, lo + cnt): unconditional buffer loads of this thread's item, zero
// outside the image and for items past the batch.
template <int LCOLS, bool VEC>
__device__ __forceinline__ void ring_issue(__amdgpu_buffer_rsrc_t rsrc, const TileArgs &a, RingBatch &b,
                                           int lo, int cnt, int row0,
                                           const unsigned (&coff)[VEC ? 2 : 8])
{
   constexpr int groups = LCOLS / 8;
   b.lo = lo;
   b.cnt = cnt;
   const int r = (int)(threadIdx.x / groups);
   const int sy = row0 + lo + r;
   const unsigned roff = (r < cnt && sy >= 0 && sy < a.H) ? (unsigned)(sy * a.sstep) : kRingOob;
   if constexpr (VEC) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
         const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)(roff + coff[q]), 0, 0);
#pragma unroll
         for (int i = 0; i < 4; ++i) b.f[4 * q + i] = __uint_as_float(v[i]);
      }
   } else {
#pragma unroll
      for (int i = 0; i < 8; ++i)
         b.f[i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)(roff + coff[i]), 0, 0));
   }
}

template <int P, int LCOLS>
__device__ __forceinline__ void ring_write(const RingBatch &b, unsigned short *ring, int ldsp)
{
   constexpr int groups = LCOLS / 8;
   const int it = threadIdx.x;
   const int r = it / groups, g = it - r * groups;
   if (r >= b.cnt) return;
   u16x8 h;
#pragma unroll
   for (int i = 0; i < 8; ++i) {
      unsigned short e0, e1;
      encode<P>(b.f[i], e0, e1);
      h[i] = e0;
   }
   *reinterpret_cast<u16x8 *>(ring + (size_t)((b.lo + r) & (kRingRows - 1)) * ldsp + 8 * g) = h;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// State of one run of crosscorr_mfma_strip (uniform except `keep`).
template <int NT>
struct StripRun {
   int k, s, ih, grows, row0, col0, x0, Y00;
   float keep[NT];      // own half (dy = 0) of the top row of the previous tile (lanes 0..15)
   int keep_y, keep_x0;
};

// Buffer resources of the strip kernel's source and output: loads at an out-of-range offset
// (kRingOob) return 0 and stores there are dropped, so no lane branches around them.
struct StripBufs {
   __amdgpu_buffer_rsrc_t src, dst;
};

__device__ __forceinline__ unsigned strip_off(bool ok, int y, int pitch, int x)
{
   return ok ? (unsigned)(y * pitch + 4 * x) : kRingOob;
}

// Value of lane + 16 in rows 0 and 2 (rows of 16 lanes): v_permlane16_swap exchanges the odd
// rows of its first operand with the even rows of its second.
__device__ __forceinline__ float lane_plus16(float v)
{
   const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
   return __uint_as_float(r[1]);
}

__device__ __forceinline__ void strip_store(const TileArgs &a, const StripBufs &bufs, bool ok, int y,
                                            int x, float v)
{
   __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), bufs.dst, (int)strip_off(ok, y, a.dstep, x), 0, 0);
}

template <int NT>
__device__ __forceinline__ void strip_flush(const TileArgs &a, const StripBufs &bufs, StripRun<NT> &st,
                                            const float *xbuf, int wave, int lane)
{
   constexpr int C = 16 * NT;
   if (st.keep_y < 0) return;
   const int dx = lane & 15;
   const bool rowok = wave > 0 && lane < 16 && st.keep_y < a.H;
#pragma unroll
   for (int j = 0; j < NT; ++j) {
      const int x = st.keep_x0 + 16 * j + dx;
      const float val = st.keep[j] + xbuf[(wave > 0 ? wave - 1 : 0) * C + 16 * j + dx];
      strip_store(a, bufs, rowok && x < a.W, st.keep_y, x, val);
   }
   st.keep_y = -1;
}

// Epilogue of tile st.k: output row of (m, dy) is Y0 - 1 + 32 wave + m + dy.  Lane
// L = 16 dy + dx + 32 h holds rows m = 8 (i / 4) + 4 h + i % 4 of column dx in register i; the
// partner (m - 1, dy = 1) of (m, dy = 0) is register i - 1 of lane L + 16 (i % 4 != 0), register
// i - 1 of lane L + 48 (i % 4 == 0, h = 0) or register i + 3 of lane L - 16 (i % 4 == 0, h = 1),
// brought over with permlane swaps.  The bottom partial row of the wave goes to xbuf for the
// wave below; the wave's own top partial row (i = 0, h = 0) waits in st.keep.
template <int NT>
__device__ __forceinline__ void strip_epilogue(const TileArgs &a, const StripBufs &bufs,
                                               StripRun<NT> &st, f32x16 (&acc)[NT], float *xbuf,
                                               int wave, int lane)
{
   constexpr int C = 16 * NT;
   const int dy = (lane >> 4) & 1, dx = lane & 15, h = lane >> 5;
   const int Y0 = st.Y00 + kPairRows * st.k;
   if (lane >= 48) {
#pragma unroll
      for (int j = 0; j < NT; ++j) xbuf[wave * C + 16 * j + dx] = acc[j][15];
   }
   const int ybase = Y0 - 1 + 32 * wave + 4 * h;
#pragma unroll
   for (int j = 0; j < NT; ++j) {
      __builtin_amdgcn_sched_barrier(0);
      const int x = st.x0 + 16 * j + dx;
      const bool colok = dy == 0 && x < a.W;
      st.keep[j] = acc[j][0];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
         float p;
         if (i % 4 != 0) {
            p = lane_plus16(acc[j][i - 1]);
         } else {
            // lanes 16..31 of z: register i - 1 of lanes 48..63; lanes 48..63 of z: register
            // i + 3 of lanes 16..31 (v_permlane32_swap exchanges the upper half of its first
            // operand with the lower half of its second)
            const unsigned bv = __float_as_uint(i >= 4 ? acc[j][i - 1] : 0.f);
            const auto q = __builtin_amdgcn_permlane32_swap(bv, __float_as_uint(acc[j][i + 3]), false, false);
            const float z = __uint_as_float(h ? q[0] : q[1]);
            p = lane_plus16(z);
         }
         const int y = ybase + 8 * (i / 4) + (i % 4);
         const bool ok = colok && y < a.H && !(i == 0 && h == 0);
         strip_store(a, bufs, ok, y, x, acc[j][i] + p);
      }
   }
   st.keep_y = Y0 - 1 + 32 * wave;
   st.keep_x0 = st.x0;
}

template <int P, int KC, int NT>
__device__ __forceinline__ void strip_mfma(f32x16 (&acc)[NT], const unsigned short *ab,
                                           const u32x4 (&b)[KC])
{
   constexpr int NA = NT + KC - 1;
   u16x8 a0[NA];
#pragma unroll
   for (int c = 0; c < NA; ++c) a0[c] = *reinterpret_cast<const u16x8 *>(ab + 16 * c);
#pragma unroll
   for (int kc = 0; kc < KC; ++kc)
#pragma unroll
      for (int j = 0; j < NT; ++j)
         acc[j] = mma<P>(a0[j + kc], __builtin_bit_cast(u16x8, b[kc]), acc[j]);
}

// One step s of tile st.k: barrier, ring write of the batch issued kRingDepth steps ago (same
// parity), B fragments of the next step, new ring loads, MFMAs.
template <int P, int KC, int NT, int MODE, bool VEC>
__device__ __forceinline__ void strip_step(const TileArgs &a, StripRun<NT> &st, f32x16 (&acc)[NT],
                                           RingBatch &mine, const u32x4 (&bcur)[KC],
                                           u32x4 (&bnext)[KC], unsigned short *ring, float *xbuf,
                                           const StripBufs &bufs, const unsigned (&coff)[VEC ? 2 : 8],
                                           int ldsp, int q_rows, int S, int s, int wave, int lane)
{
   constexpr int lcols = pair_lcols(KC, NT);
   const int live = kPairRows * st.k + 2 * s;
   __syncthreads();
   strip_flush<NT>(a, bufs, st, xbuf, wave, lane);
   ring_write<P, lcols>(mine, ring, ldsp);
   const int sn = s + 1 < S ? s + 1 : 0;
   const u32x4 *fr = reinterpret_cast<const u32x4 *>(a.frag0) + (size_t)sn * KC * 64 + lane;
#pragma unroll
   for (int kc = 0; kc < KC; ++kc) bnext[kc] = fr[kc * 64];
   const int cap = min(st.grows, live + kRingRows);
   const int cnt = max(0, min(q_rows, cap - st.ih));
   ring_issue<lcols, VEC>(bufs.src, a, mine, st.ih, cnt, st.row0, coff);
   st.ih += cnt;
   __builtin_amdgcn_sched_barrier(0);
   const int slot = (live + wave * 32 + (lane & 31)) & (kRingRows - 1);
   strip_mfma<P, KC, NT>(acc, ring + (size_t)slot * ldsp + 8 * (lane >> 5), bcur);
   __builtin_amdgcn_sched_barrier(0);
}

// The kernel runs S = pair_frag_steps(th, 1) steps per tile (an even count: for odd
// ceil(th / 2) the last step's fragments are zero), so that the two register sets of the ring
// batches and B fragments take the same roles in every tile.

template <int P, int KC, int NT, int MODE, bool VEC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
void crosscorr_mfma_strip(TileArgs a, int ldsp, int tiles_x, int tiles_y, int q_rows)
{
   static_assert(P != kBF16x3, "one operand plane");
   static_assert(MODE == kModeCC, "the CrossCorr map");
   constexpr int lcols = pair_lcols(KC, NT);
   constexpr int C = 16 * NT;
   extern __shared__ u16x8 lds_v[];
   unsigned short *ring = reinterpret_cast<unsigned short *>(lds_v);
   float *xbuf = reinterpret_cast<float *>(ring + (size_t)kRingRows * ldsp);   // [4][C]
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int S = pair_frag_steps(a.th, 1);
   StripBufs bufs;
   bufs.src = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(a.src), (short)0, a.H * a.sstep,
                                                0x00020000);
   bufs.dst = __builtin_amdgcn_make_buffer_rsrc(a.dst, (short)0, a.H * a.dstep, 0x00020000);

   // this workgroup's run of tiles, strip-major, consecutive runs on one XCD
   const int total = tiles_x * tiles_y;
   const int v = xcd_tile(blockIdx.x, gridDim.x);
   const int t_end = (int)((long long)(v + 1) * total / gridDim.x);
   StripRun<NT> st;
   st.keep_y = -1;
   for (int t = (int)((long long)v * total / gridDim.x); t < t_end;) {
      const int bx = t / tiles_y, by0 = t - bx * tiles_y;
      const int K = min(tiles_y - by0, t_end - t);
      t += K;
      st.x0 = bx * C;
      st.Y00 = by0 * kPairRows;
      st.row0 = st.Y00 - a.th / 2;
      st.col0 = st.x0 - a.tw / 2;
      st.grows = kPairRows * (K - 1) + 2 * S + kTileRows - 2;

      __syncthreads();   // the previous run's reads of the ring and xbuf are done
      strip_flush<NT>(a, bufs, st, xbuf, wave, lane);
      const int fill = min(st.grows, kRingFill);
      stage_rows<P, lcols>(a, ring, ldsp, 0, st.row0, st.col0, fill);
      unsigned coff[VEC ? 2 : 8];
      ring_columns<lcols, VEC>(a, st.col0, coff);
      st.ih = fill;
      RingBatch b0, b1;
      b0.cnt = 0;
      b1.cnt = 0;
      u32x4 f0[KC], f1[KC];
      const u32x4 *fr = reinterpret_cast<const u32x4 *>(a.frag0) + lane;
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) f0[kc] = fr[kc * 64];
      __syncthreads();   // the cold rows are in LDS
      for (st.k = 0; st.k < K; ++st.k) {
         f32x16 acc[NT];
#pragma unroll
         for (int j = 0; j < NT; ++j)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[j][i] = 0.f;
         for (int s = 0; s < S; s += 2) {
            strip_step<P, KC, NT, MODE, VEC>(a, st, acc, b0, f0, f1, ring, xbuf, bufs, coff, ldsp,
                                             q_rows, S, s, wave, lane);
            strip_step<P, KC, NT, MODE, VEC>(a, st, acc, b1, f1, f0, ring, xbuf, bufs, coff, ldsp,
                                             q_rows, S, s + 1, wave, lane);
         }
         strip_epilogue<NT>(a, bufs, st, acc, xbuf, wave, lane);
      }
   }
   __syncthreads();
   strip_flush<NT>(a, bufs, st, xbuf, wave, lane);
}

#endif  // __HIPCC__

// ---------------------------------------------------------------------------------------
// Host-side launch descriptions, implemented per precision (mfma_bf16.hip, mfma_f16.hip,
// mfma_bf16x3.hip) so the kernel instances compile in parallel.
struct SingleLaunch {
   TileArgs a;
   int KC, NT, mode, ldsp;
   size_t lds;
   dim3 grid;
   ZImages z{};   // crosscorr_mfma over several sources (z.n > 0; grid per image)
};

struct BatchLaunch {
   const float *src;
   int W, H, sstep, vec;
   const double *integ;
   const BatchItem *items;
   int nitems, tiles_x, tiles_y;
   unsigned *counter;
   int KC, NT, mode, ldsp;
   size_t lds;
   int grid;
};

template <int P> void launch_single(hipStream_t s, const SingleLaunch &L);
template <int P> void launch_batch(hipStream_t s, const BatchLaunch &L);
template <int P> void launch_single_prep(hipStream_t s, const BatchDesc &d);
template <int P> void launch_pair(hipStream_t s, const SingleLaunch &L);
template <int P> void launch_pair_prep(hipStream_t s, const BatchDesc &d);
template <int P> void launch_strip(hipStream_t s, const SingleLaunch &L, int tiles_y, int q_rows);
template <int P> void launch_batch_prep(hipStream_t s, const BatchDesc *descs, int n);
// the batch on the row-pair tile (BF16, FP16): its kernel and its fragments
template <int P> void launch_pair_batch(hipStream_t s, const BatchLaunch &L);
template <int P> void launch_pair_batch_prep(hipStream_t s, const BatchDesc *descs, int n);

}  // namespace mfma
}  // namespace ocip

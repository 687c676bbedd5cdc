// mfma.hip -- cross-correlation (and its Norm) as a contraction on gfx950 matrix cores.
//
// Semantics: reference CrossCorr / CrossCorr_Norm (cl-files/ImageProximity.cl:290-466):
//   D(x,y) = sum_v sum_u s(x+u-tw/2, y+v-th/2) * t(u,v), s = 0 outside the image.
// Arithmetic: products of BF16 or FP16 operands (or split BF16: s = s_hi + s_lo,
// t = t_hi + t_lo, s*t ~ s_hi t_hi + s_hi t_lo + s_lo t_hi), accumulated in FP32 by
// v_mfma_f32_32x32x16_{bf16,f16}.  Norm: the FP32 sum divided in FP64 by
// sqrt(window sum s^2 * sum t^2), the window sum taken from the FP64 integral of the
// exact F32 source.
//
// Contraction (Toeplitz along x).  With P the zero-padded source and template row v:
//   D[y][x0 + 32j + n] += sum_kc sum_k P[y+v][x0 + 32j + 16kc + k] * B_kc[k][n],
//   B_kc[k][n] = t(16kc + k - n, v)   (0 outside the template row)
// so one MFMA takes M = 32 output rows, N = 32 output columns and K = 16 source columns.
// KC = ceil((tw + 31) / 16) chunks cover a template row; B_kc depends only on (v, kc), so
// the KC B fragments of a template row sit in registers while a wave sweeps its output
// columns.  The A fragment at source column 16c serves every (j, kc) with 2j + kc = c.
// Issued MACs per useful MAC: 16*KC / tw (1.5 for tw = 64).
//
// Tiling: a workgroup (4 waves) owns 128 output rows x 128 output columns; wave w owns rows
// 32w..32w+31 and four 32-column n-tiles.  The source tile plus halo is staged in LDS as
// 16-bit values (two planes, hi and lo, for split BF16), one band of template rows at a
// time.  Templates wider than 97 columns are cut into 64-column blocks (KC = 6 each), staged
// one block at a time.
#include "device_common.h"

namespace ocip {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr int kNT = 4;                  // 32-column n-tiles per wave
constexpr int kWaves = 4;               // waves per workgroup, stacked in y
constexpr int kTileRows = 32 * kWaves;  // 128
constexpr int kTileCols = 32 * kNT;     // 128
constexpr int kBlockW = 64;             // template column block for wide templates
constexpr int kLdsMax = 160 * 1024;

enum { kBF16 = 1, kF16 = 2, kBF16x3 = 3 };

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

// B fragments: frag[plane][(ub * th + v) * KC + kc][lane] = 8 values B_kc[k][n] for
// n = lane % 32, k = 8 * (lane / 32) + i.
template <int P>
__global__ __launch_bounds__(256) void toeplitz_prep(const float *__restrict__ tstage,
                                                     int tpitch, int tw, int th, int KC,
                                                     int nub, u16x8 *frag0, u16x8 *frag1)
{
   const int nfrag = nub * th * KC;
   const int f = blockIdx.x * 4 + (threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   if (f >= nfrag) return;
   const int kc = f % KC, v = (f / KC) % th, ub = f / (KC * th);
   const int n = lane & 31, kb = 8 * (lane >> 5);
   const int u0 = nub > 1 ? ub * kBlockW : 0;
   const int wb = nub > 1 ? min(kBlockW, tw - u0) : tw;
   u16x8 a, b;
#pragma unroll
   for (int i = 0; i < 8; ++i) {
      const int u = 16 * kc + kb + i - n;
      const float t = (u >= 0 && u < wb) ? tstage[(size_t)v * tpitch + u0 + u] : 0.f;
      unsigned short h0, h1;
      encode<P>(t, h0, h1);
      a[i] = h0;
      b[i] = h1;
   }
   frag0[(size_t)f * 64 + lane] = a;
   if (P == kBF16x3) frag1[(size_t)f * 64 + lane] = b;
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


template <int NPL, int KC>
__device__ __forceinline__ void load_frags(const u16x8 *__restrict__ f0, const u16x8 *__restrict__ f1,
                                           size_t base, u16x8 (&b0)[KC], u16x8 (&b1)[KC])
{
#pragma unroll
   for (int kc = 0; kc < KC; ++kc) {
      b0[kc] = f0[base + (size_t)kc * 64];
      if (NPL == 2) b1[kc] = f1[base + (size_t)kc * 64];
   }
}

// One template row: A fragments from the LDS row `ab` (16 columns apart), KC B fragments.
template <int P, int NPL, int KC>
__device__ __forceinline__ void row_mfma(f32x16 (&acc)[kNT], const unsigned short *ab,
                                         size_t plane, const u16x8 (&b0)[KC],
                                         const u16x8 (&b1)[KC])
{
#pragma unroll
   for (int c = 0; c < 2 * (kNT - 1) + KC; ++c) {
      const u16x8 a0 = *reinterpret_cast<const u16x8 *>(ab + 16 * c);
      u16x8 a1;
      if (NPL == 2) a1 = *reinterpret_cast<const u16x8 *>(ab + plane + 16 * c);
#pragma unroll
      for (int j = 0; j < kNT; ++j) {
         const int kc = c - 2 * j;
         if (kc < 0 || kc >= KC) continue;
         acc[j] = mma<P>(a0, b0[kc], acc[j]);
         if (NPL == 2) {
            acc[j] = mma<P>(a0, b1[kc], acc[j]);
            acc[j] = mma<P>(a1, b0[kc], acc[j]);
         }
      }
   }
}

// ldsp: LDS row pitch in 16-bit elements (multiple of 8); lcols: staged columns (multiple
// of 8); vb: template rows per LDS band.
// LDS use limits the kernel to two workgroups per CU (two waves per SIMD), so the compiler
// may spend registers on keeping two template rows of B fragments in flight.
template <int P, int KC, bool NORM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void crosscorr_mfma(const float *__restrict__ src, int W,
                                                      int H, int sstep, int tw, int th, int nub,
                                                      const u16x8 *__restrict__ frag0,
                                                      const u16x8 *__restrict__ frag1,
                                                      const double *__restrict__ tsums,
                                                      const double *__restrict__ integ,
                                                      float *dst, int dstep, int vb, int ldsp,
                                                      int lcols, bool vec)
{
   extern __shared__ u16x8 lds_v[];
   unsigned short *lds = reinterpret_cast<unsigned short *>(lds_v);
   constexpr int NPL = P == kBF16x3 ? 2 : 1;
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int x0 = blockIdx.x * kTileCols, y0 = blockIdx.y * kTileRows;
   const int cx = tw / 2, cy = th / 2;

   f32x16 acc[kNT];
#pragma unroll
   for (int j = 0; j < kNT; ++j)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[j][i] = 0.f;

   // template column blocks (one block unless tw > 97), then bands of template rows
   for (int ub = 0; ub < nub; ++ub) {
      for (int v0 = 0; v0 < th; v0 += vb) {
         const int nv = min(vb, th - v0);
         const int nrows = kTileRows + nv - 1;
         const size_t plane = (size_t)nrows * ldsp;
         __syncthreads();
         // stage rows [y0 - cy + v0, +nrows) x columns [x0 - cx + 64 ub, +lcols), 8 per item,
         // four items per thread in flight
         const int groups = lcols / 8;
         const int total = nrows * groups;
         for (int it0 = threadIdx.x; it0 < total; it0 += 256 * 4) {
            float f[4][8];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
               const int it = it0 + 256 * k;
               const int r = it / groups, g = it - r * groups;
               const int sy = y0 - cy + v0 + r, sx = x0 - cx + ub * kBlockW + 8 * g;
               const bool rowin = it < total && sy >= 0 && sy < H;
               const float *srow = reinterpret_cast<const float *>(
                   reinterpret_cast<const char *>(src) + (size_t)(rowin ? sy : 0) * sstep);
               if (vec && rowin && sx >= 0 && sx + 8 <= W) {
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
                  unsigned short a, b;
                  encode<P>(f[k][i], a, b);
                  h0[i] = a;
                  h1[i] = b;
               }
               *reinterpret_cast<u16x8 *>(lds + (size_t)r * ldsp + 8 * g) = h0;
               if (NPL == 2) *reinterpret_cast<u16x8 *>(lds + plane + (size_t)r * ldsp + 8 * g) = h1;
            }
         }
         __syncthreads();

         // ping-pong: the B fragments of template row v+1 load while row v computes
         const size_t fb0 = ((size_t)(ub * th + v0) * KC) * 64 + lane;
         u16x8 pa0[KC], pa1[KC], pb0[KC], pb1[KC];
         load_frags<NPL, KC>(frag0, frag1, fb0, pa0, pa1);
         const unsigned short *arow = lds + (size_t)(wave * 32 + (lane & 31)) * ldsp + 8 * (lane >> 5);
         int v = 0;
         for (; v + 1 < nv; v += 2) {
            // keep each prefetch ahead of the MFMAs of the row before it
            load_frags<NPL, KC>(frag0, frag1, fb0 + (size_t)(v + 1) * KC * 64, pb0, pb1);
            __builtin_amdgcn_sched_barrier(0);
            row_mfma<P, NPL, KC>(acc, arow + (size_t)v * ldsp, plane, pa0, pa1);
            const int vn = v + 2 < nv ? v + 2 : v + 1;
            load_frags<NPL, KC>(frag0, frag1, fb0 + (size_t)vn * KC * 64, pa0, pa1);
            __builtin_amdgcn_sched_barrier(0);
            row_mfma<P, NPL, KC>(acc, arow + (size_t)(v + 1) * ldsp, plane, pb0, pb1);
         }
         if (v < nv) row_mfma<P, NPL, KC>(acc, arow + (size_t)v * ldsp, plane, pa0, pa1);
      }
   }

   // C fragment: lane holds column n = lane % 32 of each n-tile, rows
   // m = 8 * (i / 4) + 4 * (lane / 32) + i % 4
   const double tt = NORM ? tsums[0] : 0.0;
   const int n = lane & 31, mb = 4 * (lane >> 5);
#pragma unroll
   for (int j = 0; j < kNT; ++j) {
      const int x = x0 + 32 * j + n;
      if (x >= W) continue;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
         const int y = y0 + wave * 32 + 8 * (i / 4) + mb + (i % 4);
         if (y >= H) continue;
         float val = acc[j][i];
         if (NORM) val = (float)((double)val / sqrt(window_sqsum(integ, W, H, x, y, tw, th) * tt));
         reinterpret_cast<float *>(reinterpret_cast<char *>(dst) + (size_t)y * dstep)[x] = val;
      }
   }
}

struct Geometry {
   int KC, nub, lcols, ldsp, vb;
   size_t lds_bytes;
};

Geometry geometry(int precision, int tw, int th)
{
   Geometry g;
   if (tw <= 97) {
      g.KC = (tw + 31 + 15) / 16;
      if (g.KC < 2) g.KC = 2;
      g.nub = 1;
   } else {
      g.KC = 6;
      g.nub = (tw + kBlockW - 1) / kBlockW;
   }
   // A reads reach column 16*(2*(kNT-1) + KC - 1) + 15 of the staged block
   g.lcols = 16 * (2 * (kNT - 1) + g.KC);
   // row pitch: 16-byte aligned rows, shifted by 16 bytes per row modulo 128 bytes
   g.ldsp = (g.lcols + 7) / 8 * 8;
   while ((g.ldsp * 2) % 128 != 16) g.ldsp += 8;
   const int planes = precision == kBF16x3 ? 2 : 1;
   const long rows_fit = (long)kLdsMax / ((long)planes * g.ldsp * 2);
   g.vb = (int)std::min<long>(th, rows_fit - (kTileRows - 1));
   g.lds_bytes = g.vb >= 1 ? (size_t)planes * (kTileRows + g.vb - 1) * g.ldsp * 2 : 0;
   return g;
}

template <int P, int KC, bool NORM>
void launch_main(hipStream_t s, const View &S, int tw, int th, const Geometry &g,
                 const u16x8 *f0, const u16x8 *f1, const double *tsums, const double *integ,
                 float *dst, int dstep)
{
   auto kern = crosscorr_mfma<P, KC, NORM>;
   // staged columns start at x0 - cx + 64 ub + 8 g with x0 a multiple of 128: 16-byte
   // aligned when cx is a multiple of 4 and rows are 16-byte aligned
   const bool vec = (tw / 2) % 4 == 0 && S.step % 16 == 0 &&
                    reinterpret_cast<uintptr_t>(S.ptr) % 16 == 0;
   hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)g.lds_bytes),
             "lds attribute");
   dim3 grid((S.width + kTileCols - 1) / kTileCols, (S.height + kTileRows - 1) / kTileRows);
   kern<<<grid, 256, g.lds_bytes, s>>>(static_cast<const float *>(S.ptr), S.width, S.height,
                                       S.step, tw, th, g.nub, f0, f1, tsums, integ, dst, dstep,
                                       g.vb, g.ldsp, g.lcols, vec);
}

template <int P, bool NORM>
void launch_kc(hipStream_t s, const View &S, int tw, int th, const Geometry &g,
               const u16x8 *f0, const u16x8 *f1, const double *tsums, const double *integ,
               float *dst, int dstep)
{
   switch (g.KC) {
   case 2: launch_main<P, 2, NORM>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep); break;
   case 3: launch_main<P, 3, NORM>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep); break;
   case 4: launch_main<P, 4, NORM>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep); break;
   case 5: launch_main<P, 5, NORM>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep); break;
   case 6: launch_main<P, 6, NORM>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep); break;
   case 7: launch_main<P, 7, NORM>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep); break;
   case 8: launch_main<P, 8, NORM>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep); break;
   default: fail(CL_INVALID_VALUE, "template width");
   }
}

size_t frag_bytes(int precision, int tw, int th)
{
   Geometry g = geometry(precision, tw, th);
   const int planes = precision == kBF16x3 ? 2 : 1;
   return (size_t)planes * g.nub * th * g.KC * 64 * sizeof(u16x8);
}

}  // namespace

size_t crosscorr_mfma_work_bytes(int precision, int, int, int tw, int th)
{
   return frag_bytes(precision, tw, th);
}

void launch_crosscorr_mfma(hipStream_t s, int precision, bool norm, const View &S, int tw, int th,
                           const float *tstage, int tpitch, const double *tsums,
                           const double *integ, float *dst, int dstep, void *work)
{
   if (tw <= 0 || th <= 0) fail(CL_INVALID_VALUE, "empty template");
   Geometry g = geometry(precision, tw, th);
   if (g.vb < 1) fail(CL_INVALID_VALUE, "template too wide for the MFMA tile");
   const size_t nfrag = (size_t)g.nub * th * g.KC;
   u16x8 *f0 = static_cast<u16x8 *>(work);
   u16x8 *f1 = precision == kBF16x3 ? f0 + nfrag * 64 : nullptr;
   {
      KernelTimer kt(s, "toeplitz_prep");
      const int blocks = (int)((nfrag + 3) / 4);
      if (precision == kBF16)
         toeplitz_prep<kBF16><<<blocks, 256, 0, s>>>(tstage, tpitch, tw, th, g.KC, g.nub, f0, f1);
      else if (precision == kF16)
         toeplitz_prep<kF16><<<blocks, 256, 0, s>>>(tstage, tpitch, tw, th, g.KC, g.nub, f0, f1);
      else
         toeplitz_prep<kBF16x3><<<blocks, 256, 0, s>>>(tstage, tpitch, tw, th, g.KC, g.nub, f0, f1);
      launch_check("toeplitz_prep");
   }
   KernelTimer kt(s, "crosscorr_mfma");
   if (precision == kBF16) {
      if (norm) launch_kc<kBF16, true>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep);
      else launch_kc<kBF16, false>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep);
   } else if (precision == kF16) {
      if (norm) launch_kc<kF16, true>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep);
      else launch_kc<kF16, false>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep);
   } else {
      if (norm) launch_kc<kBF16x3, true>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep);
      else launch_kc<kBF16x3, false>(s, S, tw, th, g, f0, f1, tsums, integ, dst, dstep);
   }
   launch_check("crosscorr_mfma");
}

}  // namespace ocip

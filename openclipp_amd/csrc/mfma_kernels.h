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
// kModeNCCMap: CrossCorr_Norm with 1 / sqrt(SS TT) precomputed per output (ncc_inv_den)
enum { kModeCC = 0, kModeNCC = 1, kModeSSDN = 2, kModeNCCMap = 3 };

// Arguments of one output tile computation.
struct TileArgs {
   const float *src;
   int W, H, sstep;
   int tw, th, nub;
   const u16x8 *frag0, *frag1;   // B fragments (plane 1: lo parts for split BF16)
   const double *tsums;          // [0] = sum t^2
   const double *integ;          // internal exclusive FP64 integral of s^2 (Norm modes)
   const float *invden;          // 1 / sqrt(SS TT) per output, pitch W (kModeNCCMap)
   float *dst;
   int dstep;
   int vb;                       // template rows per LDS band
   int vec;                      // 16-byte staging loads allowed
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
};

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

// Batch: one workgroup per template builds its fragments and its FP64 sum of squares.
template <int P>
__global__ __launch_bounds__(256) void batch_prep(const BatchDesc *__restrict__ descs)
{
   __shared__ double red[4];
   const BatchDesc d = descs[blockIdx.x];
   const int nfrag = d.nub * d.th * d.KC;
   for (int k = threadIdx.x; k < nfrag * 64; k += 256)
      build_fragment<P>(d.tptr, d.tstep, d.tw, d.th, d.KC, d.nub, k >> 6, k & 63, d.frag0,
                        d.frag1);
   // same summation order as the single-template staging (rows padded to 16 columns,
   // 256 threads striding), so batch and single calls give identical sums
   const int pitch = (d.tw + 15) / 16 * 16;
   double s2 = 0;
   for (int k = threadIdx.x; k < pitch * d.th; k += 256) {
      const int v = k / pitch, u = k - v * pitch;
      const double t = u < d.tw ? reinterpret_cast<const float *>(d.tptr + (size_t)v * d.tstep)[u] : 0.0;
      s2 += t * t;
   }
   s2 = wave_sum(s2);
   if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s2;
   __syncthreads();
   if (threadIdx.x == 0) d.sums[0] = red[0] + red[1] + red[2] + red[3];
}

// Single template straight from its image: fragments in every block, the FP64 sum of
// squares in block 0 (same order as batch_prep).
template <int P>
__global__ __launch_bounds__(256) void single_prep(BatchDesc d)
{
   __shared__ double red[4];
   const int nfrag = d.nub * d.th * d.KC;
   const int f = blockIdx.x * 4 + (threadIdx.x >> 6);
   if (f < nfrag)
      build_fragment<P>(d.tptr, d.tstep, d.tw, d.th, d.KC, d.nub, f, threadIdx.x & 63, d.frag0,
                        d.frag1);
   if (blockIdx.x != 0) return;
   const int pitch = (d.tw + 15) / 16 * 16;
   double s2 = 0;
   for (int k = threadIdx.x; k < pitch * d.th; k += 256) {
      const int v = k / pitch, u = k - v * pitch;
      const double t = u < d.tw ? reinterpret_cast<const float *>(d.tptr + (size_t)v * d.tstep)[u] : 0.0;
      s2 += t * t;
   }
   s2 = wave_sum(s2);
   if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s2;
   __syncthreads();
   if (threadIdx.x == 0) d.sums[0] = red[0] + red[1] + red[2] + red[3];
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
         // stage rows [y0 - cy + v0, +nrows) x columns [x0 - cx + 64 ub, +lcols), 8 per item,
         // four items per thread in flight
         constexpr int groups = lcols / 8;
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
   if (MODE == kModeNCC || MODE == kModeSSDN) tt = a.tsums[0];
#pragma unroll
   for (int j = 0; j < NT; ++j) {
      const int x = x0 + 32 * j + n;
      float val[16];
      if (MODE == kModeNCCMap) {
         const int xc = x < W ? x : W - 1;
#pragma unroll
         for (int i = 0; i < 16; ++i) {
            const int y = ybase + 8 * (i / 4) + (i % 4);
            val[i] = acc[j][i] * a.invden[(size_t)(y < H ? y : H - 1) * W + xc];
         }
      } else if (MODE != kModeCC) {
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
            reinterpret_cast<float *>(reinterpret_cast<char *>(a.dst) + (size_t)y * a.dstep)[x] = val[i];
      }
   }
}

// LDS use limits these kernels to two workgroups per CU (two waves per SIMD), so the compiler
// may spend registers on keeping two template rows of B fragments in flight.
// One-dimensional grid of tiles_x * tiles_y workgroups in XCD-aware order.
template <int P, int KC, int NT, int MODE>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(NT == 2 ? 3 : 1, NT == 2 ? 3 : 2)))
void crosscorr_mfma(TileArgs a, int ldsp, int tiles_x, int tiles)
{
   extern __shared__ u16x8 lds_v[];
   const int t = xcd_tile(blockIdx.x, tiles);
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
      const int it = s_item;
      if (it >= total) break;
      const BatchItem b = items[it / tiles];
      const int tile = it - (it / tiles) * tiles;
      // 16-byte staging loads: aligned rows (vec) and tw/2 a multiple of 4
      TileArgs a{src, W, H, sstep, b.tw, b.th, b.nub, b.frag0, b.frag1, b.tsums, integ,
                 nullptr, b.dst, b.dstep, b.vb, (vec && (b.tw / 2) % 4 == 0) ? 1 : 0};
      mfma_tile<P, KC, NT, MODE>(a, tile % tiles_x, tile / tiles_x,
                                 reinterpret_cast<unsigned short *>(lds_v), ldsp);
   }
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
template <int P> void launch_batch_prep(hipStream_t s, const BatchDesc *descs, int n);

}  // namespace mfma
}  // namespace ocip

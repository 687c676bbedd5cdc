// backward.hip -- gradients of CrossCorr_Norm and SqrDistance (F32, 1 channel).
//
// The reference has no gradients; the forward maps are cl-files/ImageProximity.cl:44-73
// (SqrDistance) and :375-412 (CrossCorr_Norm).  With the template origin at (-cx, -cy),
// P the zero-padded source, G the gradient of the loss with respect to the output map,
// C = sum P t, SS = window sum of P^2 (in FP64), TT = sum t^2,
// D = sqrt(SS TT) and N = C / D:
//
//   CrossCorr_Norm   dL/dt(u,v) = sum_{x,y} (G/D)(x,y) P(x+u-cx, y+v-cy) - (sum G N) t(u,v) / TT
//                    dL/ds(j,i) = sum_{u,v} (G/D)(j+cx-u, i+cy-v) t(u,v)
//                                 - s(j,i) * sum_{u,v} (G N / SS)(j+cx-u, i+cy-v)
//   SqrDistance      dL/dt(u,v) = -2 sum_{x,y} G(x,y) P(x+u-cx, y+v-cy) + 2 t(u,v) sum G
//                    dL/ds(j,i) = 2 s(j,i) * sum_{u,v} G(j+cx-u, i+cy-v)
//                                 - 2 sum_{u,v} G(j+cx-u, i+cy-v) t(u,v)
// (map values outside the image are 0; outputs whose denominator D is 0 -- NaN in the
// forward map -- contribute nothing.)
//
// Building blocks: the template-gradient contraction (sum over all output pixels of a map
// times the shifted source, one value per template tap: tgrad_mfma on split-BF16 MFMA for
// templates up to 32 x 32, weight_corr in FP32 otherwise; partials reduced in FP64), the FP32
// correlation of a map with the flipped template (the plain tiled kernel with anchor
// (tw-1-cx, th-1-cy)), and window sums of maps in FP64 (box_f64; from an FP64 integral for
// templates too large for its tile).
#include <algorithm>

#include "device_common.h"

namespace ocip {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ double block_sum_d(double v, double *red)
{
   v = wave_sum(v);
   if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
   __syncthreads();
   double t = red[0] + red[1] + red[2] + red[3];
   __syncthreads();
   return t;
}

// Per-image arguments of the batched kernels (blockIdx.y = image of a chunk): a pointer and a
// row pitch per image, passed by value among the kernel arguments.
struct ZPtrs {
   const char *p[kSplitImages];
   int step[kSplitImages];
};

// Per-pixel maps of the gradient of L = sum G1 NCC + sum G2 SSD (has1 / has2: that term
// present), with D = sqrt(SS TT) and N = C / D (the first terms 0 where D = 0):
//   M  = G1 / D - 2 G2          (dT: its correlation with the source; dS: its correlation
//                                with the flipped template)
//   Eb = 2 G2 - G1 N / SS       (dS: s times its flipped window sums)
// each formed in FP64 and rounded once; a1 and a2 gather G1 N and G2 for the block partials.
__device__ __forceinline__ void grad_pixel(bool has1, bool has2, double c, double ss, double tt,
                                           double g1, double g2, float &m_out, float &e_out,
                                           double &a1, double &a2)
{
   double m = 0.0, e = 0.0;
   if (has1) {
      const double d = sqrt(ss * tt);
      if (d > 0) {
         const double nn = c / d;
         m = g1 / d;
         e = -(g1 * nn / ss);
         a1 += g1 * nn;
      }
   }
   if (has2) {
      m -= 2.0 * g2;
      e += 2.0 * g2;
      a2 += g2;
   }
   m_out = (float)m;
   e_out = (float)e;
}

// Window sums of a map (or of its squares) in FP64, sum over v < th, u < tw of
// f(m(x + u - ax, y + v - ay)), f = identity or square, zero outside the image, used at once
// by the epilogue EPI:
//   kBoxMaps   (the windows' sums of s^2 of the sources, anchor (tw/2, th/2)): the gradient
//              maps M and Eb of every pixel (grad_pixel) and the workgroup's partial sums;
//   kBoxSource (the flipped window sums of Eb, anchor (tw-1-tw/2, th-1-th/2)): the source
//              gradient dS = F + s * box, in FP64, rounded once (a null dS: nothing written).
// One workgroup per 32 x 64 outputs of image blockIdx.y: the map region (31 + th rows by
// 63 + tw columns) is staged in LDS as F32; a vertical pass (thread = column) keeps two FP64
// running sums down the rows from the region's top, `lead` taking each row as it enters the
// window and `lag` each row as it leaves, so the column sum V = lead - lag is a difference of
// two prefixes of one sequence (exactly 0 for a column whose window rows are all zero); V goes
// to LDS in FP64; a horizontal pass (thread = output row and 8 columns) does the same along
// the row.  Every value of f is exact in FP64 (a float or the square of a float), and the
// prefixes start at the tile, so the relative error stays near 2^-53 times (tile energy /
// window energy).
constexpr int kBxRows = 32, kBxCols = 64;
enum { kBoxMaps = 1, kBoxSource = 2 };

// Epilogue arguments of box_f64: image z's maps C (= F), M, Eb at z * zstride elements from
// their bases (W x H, rows W apart); the gradient tables (kBoxMaps), the sources and the dS
// outputs (kBoxSource); partial sums of image z at sncc / sssd + z * gridDim.x.
struct BoxEpi {
   const float *C;
   float *M, *Eb;
   size_t zstride;
   ZPtrs g1, g2, src, ds;
   int has1, has2;
   const double *tsums;
   double *sncc, *sssd;
};

size_t box_f64_lds(int tw, int th)
{
   const size_t ncols = kBxCols - 1 + tw, nrows = kBxRows - 1 + th;
   return (size_t)kBxRows * (ncols | 1) * sizeof(double) + nrows * ncols * sizeof(float);
}

template <int EPI>
__global__ __launch_bounds__(kThreads) void box_f64(ZPtrs maps, int W, int H, int tw, int th,
                                                     int ax, int ay, int sqr, BoxEpi ep,
                                                     int tiles_x)
{
   extern __shared__ double bx_lds[];
   __shared__ double red[4];
   const int z = blockIdx.y;
   const char *m = maps.p[z];
   const int mstep = maps.step[z];
   const int ncols = kBxCols - 1 + tw, nrows = kBxRows - 1 + th;
   // column sums V at an odd pitch of doubles: the horizontal pass's lanes (eight rows by
   // eight 8-column groups per wavefront) then read 16 distinct pairs of LDS banks, not 8
   const int vp = ncols | 1;
   double *vs = bx_lds;                                            // kBxRows x vp
   float *ms = reinterpret_cast<float *>(vs + kBxRows * vp);      // nrows x ncols
   const int x0 = (blockIdx.x % tiles_x) * kBxCols, y0 = (blockIdx.x / tiles_x) * kBxRows;
   const int c0 = x0 - ax, r0 = y0 - ay;
   const int total = nrows * ncols;
   // staging: up to 32 loads of each thread in flight at a time (a 31 x 31 template's tile
   // takes 23 per thread)
   for (int k0 = threadIdx.x; k0 < total; k0 += 32 * kThreads) {
      float f[32];
#pragma unroll
      for (int j = 0; j < 32; ++j) {
         const int k = k0 + j * kThreads;
         const int r = k / ncols, c = k - r * ncols;
         const int y = r0 + r, x = c0 + c;
         f[j] = (k < total && y >= 0 && y < H && x >= 0 && x < W)
                    ? reinterpret_cast<const float *>(m + (size_t)y * mstep)[x]
                    : 0.f;
      }
#pragma unroll
      for (int j = 0; j < 32; ++j) {
         const int k = k0 + j * kThreads;
         if (k < total) ms[k] = f[j];
      }
   }
   __syncthreads();
   auto fv = [&](int r, int c) {
      const double v = ms[r * ncols + c];
      return sqr ? v * v : v;
   };
   for (int c = threadIdx.x; c < ncols; c += kThreads) {
      double lead = 0.0, lag = 0.0;
      for (int i = 0; i < th - 1; ++i) lead += fv(i, c);
      for (int r = 0; r < kBxRows; ++r) {
         lead += fv(r + th - 1, c);
         vs[r * vp + c] = lead - lag;
         lag += fv(r, c);
      }
   }
   __syncthreads();
   const int r = threadIdx.x >> 3, q = (threadIdx.x & 7) * 8;
   const int y = y0 + r;
   const size_t zo = (size_t)z * ep.zstride;
   double a1 = 0.0, a2 = 0.0;
   if (y < H) {
      const double *vrow = vs + r * vp + q;
      double lead = 0.0, lag = 0.0;
      for (int k = 0; k < tw - 1; ++k) lead += vrow[k];
      const size_t ro = zo + (size_t)y * W;
      if constexpr (EPI == kBoxMaps) {
         const double tt = ep.tsums[0];
         const float *g1 = ep.has1 ? reinterpret_cast<const float *>(ep.g1.p[z] + (size_t)y * ep.g1.step[z]) : nullptr;
         const float *g2 = ep.has2 ? reinterpret_cast<const float *>(ep.g2.p[z] + (size_t)y * ep.g2.step[z]) : nullptr;
#pragma unroll
         for (int j = 0; j < 8; ++j) {
            lead += vrow[j + tw - 1];
            const int x = x0 + q + j;
            if (x < W)
               grad_pixel(ep.has1, ep.has2, ep.C[ro + x], lead - lag, tt, g1 ? (double)g1[x] : 0.0,
                          g2 ? (double)g2[x] : 0.0, ep.M[ro + x], ep.Eb[ro + x], a1, a2);
            lag += vrow[j];
         }
      } else {
         char *dsrow = const_cast<char *>(ep.ds.p[z]);
         if (dsrow) {
            float *ds = reinterpret_cast<float *>(dsrow + (size_t)y * ep.ds.step[z]);
            const float *srow = reinterpret_cast<const float *>(ep.src.p[z] + (size_t)y * ep.src.step[z]);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
               lead += vrow[j + tw - 1];
               const int x = x0 + q + j;
               if (x < W)
                  ds[x] = (float)((double)ep.C[ro + x] + (double)srow[x] * (lead - lag));
               lag += vrow[j];
            }
         }
      }
   }
   if constexpr (EPI == kBoxMaps) {
      // the workgroup's partial sums (every thread takes part in the reductions)
      if (ep.has1) {
         a1 = block_sum_d(a1, red);
         if (threadIdx.x == 0) ep.sncc[(size_t)z * gridDim.x + blockIdx.x] = a1;
      }
      if (ep.has2) {
         a2 = block_sum_d(a2, red);
         if (threadIdx.x == 0) ep.sssd[(size_t)z * gridDim.x + blockIdx.x] = a2;
      }
   }
}

int box_tiles(int W, int H) { return ((W + kBxCols - 1) / kBxCols) * ((H + kBxRows - 1) / kBxRows); }

// box_f64<EPI> over n images (maps.p[z], z < n)
template <int EPI>
void launch_box_f64(hipStream_t s, const ZPtrs &maps, int n, int W, int H, int tw, int th,
                    int ax, int ay, bool sqr, const BoxEpi &ep)
{
   const int tiles_x = (W + kBxCols - 1) / kBxCols;
   const size_t lds = box_f64_lds(tw, th);
   if (lds > 64 * 1024)
      hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(box_f64<EPI>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                "lds attribute");
   box_f64<EPI><<<dim3(box_tiles(W, H), n), kThreads, lds, s>>>(maps, W, H, tw, th, ax, ay,
                                                               sqr ? 1 : 0, ep, tiles_x);
   launch_check("box_f64");
}

// The gradient maps (grad_pixel) of every pixel with SS from the exclusive FP64 integral
// `integ` (templates whose box tile does not fit LDS), or without SS (has1 = 0), and
// per-block partial sums.  Image z = blockIdx.y: C, M and Eb at z * zstride elements from their
// bases (W x H, rows W apart), G1 / G2 from the tables, partial sums at sncc / sssd +
// z * gridDim.x.
__global__ __launch_bounds__(kThreads) void grad_maps(const float *Cs, ZPtrs g1s, ZPtrs g2s,
                                                       int has1, int has2,
                                                       const double *__restrict__ integ,
                                                       const double *__restrict__ tsums, int W,
                                                       int H, int tw, int th, float *Ms,
                                                       float *Ebs, size_t zstride,
                                                       double *snccs, double *sssds)
{
   __shared__ double red[4];
   const size_t zo = blockIdx.y * zstride;
   const float *C = Cs + zo;
   float *M = Ms + zo, *Eb = Ebs + zo;
   const char *G1 = has1 ? g1s.p[blockIdx.y] : nullptr, *G2 = has2 ? g2s.p[blockIdx.y] : nullptr;
   const int g1step = g1s.step[blockIdx.y], g2step = g2s.step[blockIdx.y];
   double *sncc = snccs + (size_t)blockIdx.y * gridDim.x, *sssd = sssds + (size_t)blockIdx.y * gridDim.x;
   const double tt = tsums[0];
   const int cx = tw / 2, cy = th / 2;
   double a1 = 0, a2 = 0;
   for (long long k = (long long)blockIdx.x * kThreads + threadIdx.x; k < (long long)W * H;
        k += (long long)gridDim.x * kThreads) {
      const int y = (int)(k / W), x = (int)(k - (long long)y * W);
      const double ss = G1 ? rect_sum(integ, W, H, x - cx, x - cx + tw, y - cy, y - cy + th) : 0.0;
      const double g1 = G1 ? reinterpret_cast<const float *>(G1 + (size_t)y * g1step)[x] : 0.0;
      const double g2 = G2 ? reinterpret_cast<const float *>(G2 + (size_t)y * g2step)[x] : 0.0;
      grad_pixel(G1 != nullptr, G2 != nullptr, C[k], ss, tt, g1, g2, M[k], Eb[k], a1, a2);
   }
   if (G1) {
      a1 = block_sum_d(a1, red);
      if (threadIdx.x == 0) sncc[blockIdx.x] = a1;
   }
   if (G2) {
      a2 = block_sum_d(a2, red);
      if (threadIdx.x == 0) sssd[blockIdx.x] = a2;
   }
}

// partial[b][v * tw + u] = sum over the tile's output pixels of M(x, y) * P(x+u-cx, y+v-cy)
template <int TILE>
__global__ __launch_bounds__(kThreads) void weight_corr(const char *M, int mstep, const float *S,
                                                         int sstep, int W, int H, int tw, int th,
                                                         float *partial, int ldsw)
{
   extern __shared__ float lds[];
   float *mt = lds;
   float *st = lds + TILE * TILE;
   const int x0 = blockIdx.x * TILE, y0 = blockIdx.y * TILE;
   const int cx = tw / 2, cy = th / 2;
   for (int i = threadIdx.x; i < TILE * TILE; i += kThreads) {
      const int py = i / TILE, px = i - py * TILE;
      const int y = y0 + py, x = x0 + px;
      mt[i] = (y < H && x < W) ? reinterpret_cast<const float *>(M + (size_t)y * mstep)[x] : 0.f;
   }
   const int srows = TILE + th - 1, scols = TILE + tw - 1;
   for (int i = threadIdx.x; i < srows * ldsw; i += kThreads) {
      const int r = i / ldsw, c = i - r * ldsw;
      const int y = y0 - cy + r, x = x0 - cx + c;
      st[i] = (c < scols && y >= 0 && y < H && x >= 0 && x < W)
                  ? reinterpret_cast<const float *>(reinterpret_cast<const char *>(S) +
                                                    (size_t)y * sstep)[x]
                  : 0.f;
   }
   __syncthreads();
   // each thread takes four adjacent taps (u0 .. u0 + 3, v) of one template row, so every
   // window value read feeds four FMAs; each tap still sums over (py, px) in order. The window
   // row is read as 16-byte LDS loads (ldsw and u0 are multiples of 4, ldsw = 4 mod 32): they
   // reach column u0 + TILE + 3 <= tw + TILE + 2 < ldsw, zeros past the staged columns, and
   // taps past tw are not stored.
   const int qw = (tw + 3) / 4, ngroups = qw * th;
   float *out = partial + (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * (tw * th);
   for (int gi = threadIdx.x; gi < ngroups; gi += kThreads) {
      const int v = gi / qw, u0 = 4 * (gi - v * qw);
      float acc[4] = {0.f, 0.f, 0.f, 0.f};
      for (int py = 0; py < TILE; ++py) {
         const float4 *m4 = reinterpret_cast<const float4 *>(mt + py * TILE);
         const float4 *s4 = reinterpret_cast<const float4 *>(st + (size_t)(py + v) * ldsw + u0);
#pragma unroll
         for (int c = 0; c < TILE / 16; ++c) {
            float m[16], w[20];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
               const float4 a = m4[4 * c + k];
               m[4 * k] = a.x; m[4 * k + 1] = a.y; m[4 * k + 2] = a.z; m[4 * k + 3] = a.w;
            }
#pragma unroll
            for (int k = 0; k < 5; ++k) {
               const float4 b = s4[4 * c + k];
               w[4 * k] = b.x; w[4 * k + 1] = b.y; w[4 * k + 2] = b.z; w[4 * k + 3] = b.w;
            }
#pragma unroll
            for (int px = 0; px < 16; ++px)
#pragma unroll
               for (int j = 0; j < 4; ++j) acc[j] = __builtin_fmaf(m[px], w[px + j], acc[j]);
         }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
         if (u0 + j < tw) out[v * tw + u0 + j] = acc[j];
   }
}

// ---- the template-gradient contraction on MFMA (templates up to 32 x 32) -------------------
// partial[(task, wave)][v * tw + u] = sum over the task's pixels (y, x') of
//    A[v][(y, x')] * B[(y, x')][u],   A[v][(y, x')] = P(x', y + v - cy),  B[(y, x')][u] = M(x' - u + cx, y)
// i.e. sum over x, y of M(x, y) P(x + u - cx, y + v - cy) with x' = x + u - cx: a GEMM with
// M = 32 rows (template rows v), N = 32 columns (template columns u) and K = the image's pixels,
// on v_mfma_f32_32x32x16_bf16 with split BF16 operands (x = hi + lo, three products hi hi,
// hi lo, lo hi: relative error below 2^-14 per product) and FP32 accumulation.  Lane l holds
// A[m][k] and B[k][n] for m, n = l % 32 and k = 8 (l / 32) + i (i < 8), and C[m][n] for
// m = 8 (i / 4) + 4 (l / 32) + i % 4 (the layouts the forward MFMA kernels use).
// A task is kTgRows rows y by kTgCols columns x' of the image: the source rows it reads
// (y - cy .. y - cy + 31) and the map rows (columns x' - 31 + cx .. x' + cx) are staged in LDS
// as BF16 hi and lo planes, the map twice (the second copy one element to the left), so that
// every B read of 8 consecutive elements starts on a 4-byte boundary in one of the copies.
// Wave w takes the task's rows y0 + w, y0 + w + 4, ...; the four waves' 32 x 32 sums are
// added in wave order and the task writes one partial per tap, which grad_template reduces in
// FP64.
constexpr int kTgRows = 32, kTgCols = 128;
constexpr int kTgPRows = kTgRows + 31;       // staged source rows
constexpr int kTgMCols = kTgCols + 32;       // staged map columns (31 to the left, 1 spare)

typedef __bf16 tg_bf16x8 __attribute__((ext_vector_type(8)));
typedef float tg_f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short tg_u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short tg_hi(float f)
{
   return __builtin_bit_cast(unsigned short, (__bf16)f);
}
__device__ __forceinline__ unsigned short tg_lo(float f, unsigned short hi)
{
   return __builtin_bit_cast(unsigned short,
                             (__bf16)(f - (float)__builtin_bit_cast(__bf16, hi)));
}
__device__ __forceinline__ tg_f32x16 tg_mma(tg_u16x8 a, tg_u16x8 b, tg_f32x16 c)
{
   return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(tg_bf16x8, a),
                                                  __builtin_bit_cast(tg_bf16x8, b), c, 0, 0, 0);
}

__global__ __launch_bounds__(256) void tgrad_mfma(const char *Ms, int mstep, size_t mzstride,
                                                   ZPtrs srcs, int W, int H, int tw, int th,
                                                   int tasks_x, float *partials, size_t pzstride)
{
   // image blockIdx.y: its map at Ms + blockIdx.y * mzstride bytes, its source from the table,
   // its partials at partials + blockIdx.y * pzstride
   const char *M = Ms + blockIdx.y * mzstride;
   const float *S = reinterpret_cast<const float *>(srcs.p[blockIdx.y]);
   const int sstep = srcs.step[blockIdx.y];
   float *partial = partials + blockIdx.y * pzstride;
   // planes: source hi, source lo (kTgPRows x kTgCols); map hi, lo, each in two copies
   // (kTgRows x kTgMCols)
   __shared__ unsigned short ps[2][kTgPRows * kTgCols];
   __shared__ unsigned short ms[2][2][kTgRows * kTgMCols];
   const int cx = tw / 2, cy = th / 2;
   const int tx = blockIdx.x % tasks_x, ty = blockIdx.x / tasks_x;
   const int xb = tx * kTgCols, y0 = ty * kTgRows;
   // source rows y0 - cy .. y0 - cy + kTgPRows - 1, columns xb .. xb + kTgCols - 1 (0 outside),
   // and map rows y0 .. y0 + kTgRows - 1 (copy 0 element j = map column xb + cx - 31 + j,
   // copy 1 element j = copy 0 element j + 1): every load of the thread issued before any
   // store (the loop counts are constants)
   constexpr int kPIt = (kTgPRows * kTgCols + 255) / 256, kMIt = (kTgRows * kTgMCols + 255) / 256;
   float pf[kPIt], mf0[kMIt], mf1[kMIt];
#pragma unroll
   for (int it = 0; it < kPIt; ++it) {
      const int k = threadIdx.x + 256 * it;
      const int r = k / kTgCols, c = k - r * kTgCols;
      const int y = y0 - cy + r, x = xb + c;
      pf[it] = (k < kTgPRows * kTgCols && y >= 0 && y < H && x < W)
                   ? reinterpret_cast<const float *>(reinterpret_cast<const char *>(S) +
                                                     (size_t)y * sstep)[x]
                   : 0.f;
   }
#pragma unroll
   for (int it = 0; it < kMIt; ++it) {
      const int k = threadIdx.x + 256 * it;
      const int r = k / kTgMCols, j = k - r * kTgMCols;
      const int y = y0 + r;
      const bool rowin = k < kTgRows * kTgMCols && y < H;
      const float *mrow = reinterpret_cast<const float *>(M + (size_t)(rowin ? y : 0) * mstep);
      const int x0c = xb + cx - 31 + j;
      mf0[it] = (rowin && x0c >= 0 && x0c < W) ? mrow[x0c] : 0.f;
      mf1[it] = (rowin && x0c + 1 >= 0 && x0c + 1 < W) ? mrow[x0c + 1] : 0.f;
   }
#pragma unroll
   for (int it = 0; it < kPIt; ++it) {
      const int k = threadIdx.x + 256 * it;
      if (k >= kTgPRows * kTgCols) break;
      const unsigned short h = tg_hi(pf[it]);
      ps[0][k] = h;
      ps[1][k] = tg_lo(pf[it], h);
   }
#pragma unroll
   for (int it = 0; it < kMIt; ++it) {
      const int k = threadIdx.x + 256 * it;
      if (k >= kTgRows * kTgMCols) break;
      const unsigned short h0 = tg_hi(mf0[it]), h1 = tg_hi(mf1[it]);
      ms[0][0][k] = h0;
      ms[1][0][k] = tg_lo(mf0[it], h0);
      ms[0][1][k] = h1;
      ms[1][1][k] = tg_lo(mf1[it], h1);
   }
   __syncthreads();
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int mn = lane & 31, g = lane >> 5;
   tg_f32x16 acc;
#pragma unroll
   for (int i = 0; i < 16; ++i) acc[i] = 0.f;
   for (int r = wave; r < kTgRows; r += 4) {
      if (y0 + r >= H) break;
      // A: source row (task row r + m), columns 16 kb + 8 g .. + 7
      const unsigned short *ah = &ps[0][(r + mn) * kTgCols + 8 * g];
      const unsigned short *al = &ps[1][(r + mn) * kTgCols + 8 * g];
#pragma unroll 2
      for (int kb = 0; kb < kTgCols / 16; ++kb) {
         if (xb + 16 * kb >= W) break;
         const tg_u16x8 a_hi = *reinterpret_cast<const tg_u16x8 *>(ah + 16 * kb);
         const tg_u16x8 a_lo = *reinterpret_cast<const tg_u16x8 *>(al + 16 * kb);
         // B: map elements at copy-0 index 16 kb + 8 g - n + 31 .. + 7 (n = mn)
         const int s0 = 16 * kb + 8 * g - mn + 31;
         const int cp = s0 & 1, base = r * kTgMCols + (s0 - cp);
         tg_u16x8 b_hi, b_lo;
#pragma unroll
         for (int q = 0; q < 4; ++q) {
            const unsigned wh = *reinterpret_cast<const unsigned *>(&ms[0][cp][base + 2 * q]);
            const unsigned wl = *reinterpret_cast<const unsigned *>(&ms[1][cp][base + 2 * q]);
            b_hi[2 * q] = (unsigned short)(wh & 0xffffu);
            b_hi[2 * q + 1] = (unsigned short)(wh >> 16);
            b_lo[2 * q] = (unsigned short)(wl & 0xffffu);
            b_lo[2 * q + 1] = (unsigned short)(wl >> 16);
         }
         acc = tg_mma(a_hi, b_hi, acc);
         acc = tg_mma(a_hi, b_lo, acc);
         acc = tg_mma(a_lo, b_hi, acc);
      }
   }
   // the four waves' C[m][n] summed in wave order through LDS (the source planes are no longer
   // read), then partial[task][m * tw + n] for m < th, n < tw
   __syncthreads();
   float *red = reinterpret_cast<float *>(&ps[0][0]);   // 4 x 32 x 32 floats
#pragma unroll
   for (int i = 0; i < 16; ++i) {
      const int m = 8 * (i / 4) + 4 * g + i % 4;
      red[(wave * 32 + m) * 32 + mn] = acc[i];
   }
   __syncthreads();
   float *out = partial + (size_t)blockIdx.x * (size_t)(tw * th);
   for (int k = threadIdx.x; k < 32 * 32; k += 256) {
      const int m = k >> 5, n = k & 31;
      if (m < th && n < tw)
         out[m * tw + n] = ((red[k] + red[1024 + k]) + red[2048 + k]) + red[3072 + k];
   }
}

// The template-gradient partials are reduced in two stages.  First stage:
// red[g][tap] = sum of partial[b][tap] over the kRedBlocks consecutive blocks b of group g, in
// FP64; 64 taps per workgroup (lane = tap), the four waves take blocks b = wave (mod 4) eight
// at a time (loads in flight) in order, then lane's four sums are added in wave order.
constexpr int kRedBlocks = 128;

__global__ __launch_bounds__(kThreads) void grad_partials(const float *partial, int nblocks,
                                                           int ntaps, double *red)
{
   __shared__ double part[4][64];
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int tap = blockIdx.x * 64 + lane;
   const int b0 = blockIdx.y * kRedBlocks, b1 = min(nblocks, b0 + kRedBlocks);
   double r = 0;
   if (tap < ntaps) {
      int k = b0 + wave;
      for (; k + 28 < b1; k += 32) {
         float q[8];
#pragma unroll
         for (int j = 0; j < 8; ++j) q[j] = partial[(size_t)(k + 4 * j) * ntaps + tap];
#pragma unroll
         for (int j = 0; j < 8; ++j) r += q[j];
      }
      for (; k < b1; k += 4) r += partial[(size_t)k * ntaps + tap];
   }
   part[wave][lane] = r;
   __syncthreads();
   if (wave == 0 && tap < ntaps)
      red[(size_t)blockIdx.y * ntaps + tap] =
          ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
}

// The scalar sums of the per-block partials (n of them each, either n may be 0): out[0] =
// sum of sncc, out[1] = sum of sssd, in FP64; one workgroup, each thread's share eight loads
// at a time, then the workgroup sums in wave order.
__global__ __launch_bounds__(kThreads) void grad_scalars(const double *sncc, int nncc,
                                                          const double *sssd, int nssd,
                                                          double *out)
{
   __shared__ double red[4];
   double s[2] = {0.0, 0.0};
#pragma unroll
   for (int which = 0; which < 2; ++which) {
      const double *p = which ? sssd : sncc;
      const int n = which ? nssd : nncc;
      int k = threadIdx.x;
      for (; k + 7 * kThreads < n; k += 8 * kThreads) {
         double q[8];
#pragma unroll
         for (int j = 0; j < 8; ++j) q[j] = p[k + j * kThreads];
#pragma unroll
         for (int j = 0; j < 8; ++j) s[which] += q[j];
      }
      for (; k < n; k += kThreads) s[which] += p[k];
   }
   const double a = block_sum_d(s[0], red);
   const double b = block_sum_d(s[1], red);
   if (threadIdx.x == 0) {
      out[0] = a;
      out[1] = b;
   }
}

// Second stage: dT(u, v) = sum_g red[g][tap] + c t(u, v) (groups in order), c = 2 (sum G2) -
// (sum G1 N) / TT from the scalar sums sc[0] = sum G1 N, sc[1] = sum G2.  64 taps per
// workgroup (lane = tap); the four waves take groups g = wave (mod 4) in order, then lane's
// four sums are added in wave order.
__global__ __launch_bounds__(kThreads) void grad_template(const double *red, int ngroups,
                                                           const double *sc, const float *tstage,
                                                           int tpitch, const double *tsums,
                                                           int tw, int th, char *dT, int dtstep)
{
   __shared__ double part[4][64];
   const double cterm = 2.0 * sc[1] - sc[0] / tsums[0];
   const int ntaps = tw * th, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int tap = blockIdx.x * 64 + lane;
   double r = 0;
   if (tap < ntaps)
      for (int g = wave; g < ngroups; g += 4) r += red[(size_t)g * ntaps + tap];
   part[wave][lane] = r;
   __syncthreads();
   if (wave == 0 && tap < ntaps) {
      r = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
      const int v = tap / tw, u = tap - v * tw;
      const double t = tstage[(size_t)v * tpitch + u];
      reinterpret_cast<float *>(dT + (size_t)v * dtstep)[u] = (float)(r + cterm * t);
   }
}

// Flipped staged template with a leading zero column when tw is even and a leading zero row
// when th is even (twf = tw + ox, thf = th + oy): tf(u, v) = t(tw-1-(u-ox), th-1-(v-oy)).
// Correlating a map with tf at anchor (twf/2, thf/2) is correlating it with the flipped
// template at anchor (tw-1-tw/2, th-1-th/2).
__global__ void flip_template(const float *tstage, int tpitch, int tw, int th, int twf, int thf,
                              float *tflip)
{
   const int ox = twf - tw, oy = thf - th;
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < twf * thf; i += gridDim.x * blockDim.x) {
      const int v = i / twf - oy, u = i % twf - ox;
      tflip[i] = (u >= 0 && v >= 0) ? tstage[(size_t)(th - 1 - v) * tpitch + (tw - 1 - u)] : 0.f;
   }
}

// dS(j, i) = F(j, i) + s(j, i) * (flipped window sum of Eb from its exclusive FP64 integral
// I), in FP64, rounded once (templates whose box tile does not fit LDS; box_f64<kBoxSource>
// otherwise).  Image z = blockIdx.y: F at z * zstride elements from its base, source and dS
// from the tables (a null dS: nothing to write).
__global__ __launch_bounds__(kThreads) void grad_source(const float *Fs, ZPtrs srcs,
                                                         const double *__restrict__ I,
                                                         size_t zstride, int W, int H, int tw,
                                                         int th, ZPtrs dss)
{
   char *dS = const_cast<char *>(dss.p[blockIdx.y]);
   if (!dS) return;
   const int dsstep = dss.step[blockIdx.y];
   const float *F = Fs + blockIdx.y * zstride;
   const float *S = reinterpret_cast<const float *>(srcs.p[blockIdx.y]);
   const int sstep = srcs.step[blockIdx.y];
   const int cx = tw / 2, cy = th / 2;
   for (long long k = (long long)blockIdx.x * kThreads + threadIdx.x; k < (long long)W * H;
        k += (long long)gridDim.x * kThreads) {
      const int i = (int)(k / W), j = (int)(k - (long long)i * W);
      const double box = rect_sum(I, W, H, j + cx - tw + 1, j + cx + 1, i + cy - th + 1, i + cy + 1);
      const double s = reinterpret_cast<const float *>(reinterpret_cast<const char *>(S) +
                                                       (size_t)i * sstep)[j];
      reinterpret_cast<float *>(dS + (size_t)i * dsstep)[j] = (float)((double)F[k] + s * box);
   }
}

int grid_for(long long n) { return (int)std::min<long long>(1024, (n + kThreads - 1) / kThreads); }

// Images of one chunk: their CrossCorr maps (then their dS correlations), M and Eb maps are
// held together so that each correlation runs on the whole chunk in one launch.
constexpr int kChunkPixels = 16 << 20;

struct BwLayout {
   size_t F, M, E, I, partial, red, scal, tflip, fragT, fragF, sums, total;
   int nb;        // template-gradient partial blocks per image
   int zc;        // images per chunk
   int twf, thf;  // flipped template, zero-padded to odd sizes
};

size_t align_up(size_t n) { return (n + 255) / 256 * 256; }

// the template-gradient contraction: tgrad_mfma for templates up to 32 x 32, else weight_corr
// on 64- (or, when its LDS does not fit, 32-) pixel tiles
bool tgrad_on_mfma(int tw, int th) { return tw <= 32 && th <= 32; }
int wc_tile(int tw, int th)
{
   return (64 + th - 1) * (64 + tw - 1 + 4) * 4 + 64 * 64 * 4 <= 64 * 1024 ? 64 : 32;
}
int tgrad_blocks(int W, int H, int tw, int th)
{
   if (tgrad_on_mfma(tw, th)) return ((W + kTgCols - 1) / kTgCols) * ((H + kTgRows - 1) / kTgRows);
   const int t = wc_tile(tw, th);
   return ((W + t - 1) / t) * ((H + t - 1) / t);
}

BwLayout bw_layout(int n, int W, int H, int tw, int th)
{
   const size_t px = (size_t)W * H;
   BwLayout L;
   L.nb = tgrad_blocks(W, H, tw, th);
   L.zc = (int)std::max<size_t>(1, std::min<size_t>({(size_t)std::max(n, 1), (size_t)kSplitImages,
                                                     kChunkPixels / std::max<size_t>(px, 1)}));
   L.twf = tw + (tw % 2 == 0);
   L.thf = th + (th % 2 == 0);
   const size_t maps = (size_t)L.zc * px * sizeof(float);
   size_t o = 0;
   L.F = o;       o = align_up(o + maps);
   L.M = o;       o = align_up(o + maps);
   L.E = o;       o = align_up(o + maps);
   L.I = o;       o = align_up(o + excl_bytes(W, H));   // templates too large for box_f64
   L.partial = o; o = align_up(o + (size_t)n * L.nb * tw * th * sizeof(float));
   L.red = o;     o = align_up(o + (size_t)((n * L.nb + kRedBlocks - 1) / kRedBlocks) * tw * th *
                                   sizeof(double));
   // per image, one partial sum of G1 N and of G2 per block of grad_maps (at most 1024) or
   // per tile of box_f64
   L.scal = o;    o = align_up(o + (size_t)2 * n * std::max(1024, box_tiles(W, H)) * sizeof(double));
   L.tflip = o;   o = align_up(o + (size_t)L.twf * L.thf * sizeof(float));
   L.fragT = o;   o = align_up(o + crosscorr_split_frag_bytes(tw, th));
   L.fragF = o;   o = align_up(o + crosscorr_split_frag_bytes(L.twf, L.thf));
   L.sums = o;    o = align_up(o + 8 * sizeof(double));
   L.total = o;
   return L;
}

}  // namespace

size_t backward_work_bytes(int n, int W, int H, int tw, int th, int tpitch)
{
   (void)tpitch;
   return bw_layout(n, W, H, tw, th).total;
}

// Gradients of L = sum_i [sum G1_i NCC(S_i, t) + sum G2_i SSD(S_i, t)] (G1 / G2 null: that op
// absent): dS_i for the non-null dS[i], dT (if not null) summed over the batch.  Images go in
// chunks: the forward CrossCorr maps of a chunk (split-BF16 MFMA), then per image the window
// sums of s^2, the maps M and Eb and the template-gradient contraction, then the chunk's
// correlations of M with the flipped template (split-BF16 MFMA), then per image dS.
void launch_backward_batch(hipStream_t s, int n, const View *S, const View *G1, const View *G2,
                           const float *tstage, int tpitch, int tw, int th, const double *tsums,
                           void *work, void *aux, char *const *dS, const int *dsstep, char *dT,
                           int dtstep)
{
   const int W = S[0].width, H = S[0].height;
   const size_t px = (size_t)W * H;
   const BwLayout L = bw_layout(n, W, H, tw, th);
   char *w = static_cast<char *>(work);
   float *F0 = reinterpret_cast<float *>(w + L.F);
   float *M0 = reinterpret_cast<float *>(w + L.M);
   float *E0 = reinterpret_cast<float *>(w + L.E);
   double *I = reinterpret_cast<double *>(w + L.I);
   float *partial = reinterpret_cast<float *>(w + L.partial);
   double *sncc = reinterpret_cast<double *>(w + L.scal);
   double *sssd = sncc + (size_t)n * std::max(1024, box_tiles(W, H));
   float *tflip = reinterpret_cast<float *>(w + L.tflip);
   double *sums = reinterpret_cast<double *>(w + L.sums);
   // window sums straight from the maps (box_f64, its epilogues forming the gradient maps
   // and dS) when its LDS tile fits, else from integrals
   const bool use_box = box_f64_lds(tw, th) <= (size_t)(160 * 1024 - 256);
   // partial sums per image: one per tile of box_f64<kBoxMaps>, else per block of grad_maps
   const int nscal = use_box && G1 ? box_tiles(W, H) : grid_for((long long)px);
   const int taps = tw * th;
   KernelTimer kt(s, "backward");

   bool any_ds = false;
   for (int i = 0; i < n; ++i) any_ds = any_ds || dS[i];
   if (G1) {
      const View tv{tstage, tw, th, tpitch * (int)sizeof(float), 1, SImage::F32};
      prep_crosscorr_split(s, tv, w + L.fragT, sums);
   }
   if (any_ds) {
      flip_template<<<grid_for((long long)L.twf * L.thf), kThreads, 0, s>>>(tstage, tpitch, tw, th,
                                                                           L.twf, L.thf, tflip);
      launch_check("flip_template");
      const View fv{tflip, L.twf, L.thf, L.twf * (int)sizeof(float), 1, SImage::F32};
      prep_crosscorr_split(s, fv, w + L.fragF, sums + 4);
   }
   const float *src[kSplitImages];
   float *dst[kSplitImages];
   int step[kSplitImages];
   auto one = [](const void *p, int st) {
      ZPtrs z{};
      z.p[0] = static_cast<const char *>(p);
      z.step[0] = st;
      return z;
   };
   for (int c0 = 0; c0 < n; c0 += L.zc) {
      const int m = std::min(L.zc, n - c0);
      // the chunk's tables: sources, gradient maps, dS outputs, and the internal maps E
      ZPtrs zs{}, zg1{}, zg2{}, zds{}, ze{};
      bool chunk_ds = false;
      for (int k = 0; k < m; ++k) {
         const int i = c0 + k;
         zs.p[k] = static_cast<const char *>(S[i].ptr);
         zs.step[k] = S[i].step;
         if (G1) { zg1.p[k] = static_cast<const char *>(G1[i].ptr); zg1.step[k] = G1[i].step; }
         if (G2) { zg2.p[k] = static_cast<const char *>(G2[i].ptr); zg2.step[k] = G2[i].step; }
         zds.p[k] = dS[i];
         zds.step[k] = dS[i] ? dsstep[i] : 0;
         ze.p[k] = reinterpret_cast<const char *>(E0 + (size_t)k * px);
         ze.step[k] = W * 4;
         chunk_ds = chunk_ds || dS[i];
      }
      if (G1) {
         // the chunk's forward CrossCorr maps
         for (int k = 0; k < m; ++k) {
            src[k] = static_cast<const float *>(S[c0 + k].ptr);
            step[k] = S[c0 + k].step;
            dst[k] = F0 + (size_t)k * px;
         }
         launch_crosscorr_split(s, m, src, step, dst, W * 4, W, H, tw, th, w + L.fragT, sums);
      }
      // the maps M and Eb (and partial sums): from the windows' sums of s^2 in the epilogue
      // of box_f64, or with those sums from integrals, or without them (no CrossCorr_Norm term)
      if (use_box && G1) {
         BoxEpi ep{};
         ep.C = F0;
         ep.M = M0;
         ep.Eb = E0;
         ep.zstride = px;
         ep.g1 = zg1;
         ep.g2 = zg2;
         ep.has1 = 1;
         ep.has2 = G2 ? 1 : 0;
         ep.tsums = tsums;
         ep.sncc = sncc + (size_t)c0 * nscal;
         ep.sssd = sssd + (size_t)c0 * nscal;
         launch_box_f64<kBoxMaps>(s, zs, m, W, H, tw, th, tw / 2, th / 2, true, ep);
      } else if (!G1) {
         grad_maps<<<dim3(nscal, m), kThreads, 0, s>>>(
             F0, zg1, zg2, 0, 1, nullptr, tsums, W, H, tw, th, M0, E0, px,
             sncc + (size_t)c0 * nscal, sssd + (size_t)c0 * nscal);
         launch_check("grad_maps");
      } else {
         for (int k = 0; k < m; ++k) {
            const int i = c0 + k;
            launch_integral(s, S[i], 0, true, I, 0, -1, aux);
            grad_maps<<<dim3(nscal, 1), kThreads, 0, s>>>(
                F0 + (size_t)k * px, one(G1[i].ptr, G1[i].step),
                G2 ? one(G2[i].ptr, G2[i].step) : ZPtrs{}, 1, G2 ? 1 : 0, I, tsums, W, H, tw,
                th, M0 + (size_t)k * px, E0 + (size_t)k * px, 0, sncc + (size_t)i * nscal,
                sssd + (size_t)i * nscal);
            launch_check("grad_maps");
         }
      }
      if (dT) {
         float *pc = partial + (size_t)c0 * L.nb * taps;
         if (tgrad_on_mfma(tw, th)) {
            // the contraction on MFMA (split BF16); the FP64 reduction follows the batch
            const int tasks_x = (W + kTgCols - 1) / kTgCols;
            tgrad_mfma<<<dim3(L.nb, m), kThreads, 0, s>>>(
                reinterpret_cast<const char *>(M0), W * 4, px * sizeof(float), zs, W, H, tw, th,
                tasks_x, pc, (size_t)L.nb * taps);
            launch_check("tgrad_mfma");
         } else {
            const int tile = wc_tile(tw, th);
            // a multiple of 4 floats = 4 mod 32 (16-byte window loads, rows on different
            // banks), at least tile + tw + 3
            int ldsw = (tile + tw - 1 + 4 + 3) / 4 * 4;
            while (ldsw % 32 != 4) ldsw += 4;
            const size_t lds = (size_t)tile * tile * 4 + (size_t)(tile + th - 1) * ldsw * 4;
            if (lds > 160 * 1024) fail(CL_INVALID_VALUE, "template too large for the gradient");
            dim3 grid((W + tile - 1) / tile, (H + tile - 1) / tile);
            for (int k = 0; k < m; ++k) {
               const char *Mc = reinterpret_cast<const char *>(M0 + (size_t)k * px);
               const float *Sp = static_cast<const float *>(S[c0 + k].ptr);
               float *pi = pc + (size_t)k * L.nb * taps;
               if (tile == 64) {
                  hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(weight_corr<64>),
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), "attr");
                  weight_corr<64><<<grid, kThreads, lds, s>>>(Mc, W * 4, Sp, S[c0 + k].step, W, H, tw, th, pi, ldsw);
               } else {
                  hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(weight_corr<32>),
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), "attr");
                  weight_corr<32><<<grid, kThreads, lds, s>>>(Mc, W * 4, Sp, S[c0 + k].step, W, H, tw, th, pi, ldsw);
               }
               launch_check("weight_corr");
            }
         }
      }
      if (!chunk_ds) continue;
      // the chunk's correlations of M with the flipped template, into F
      for (int k = 0; k < m; ++k) {
         src[k] = M0 + (size_t)k * px;
         step[k] = W * 4;
         dst[k] = F0 + (size_t)k * px;
      }
      launch_crosscorr_split(s, m, src, step, dst, W * 4, W, H, L.twf, L.thf, w + L.fragF, sums + 4);
      // dS from F, s and the flipped window sums of Eb (window anchored at tw - 1 - cx,
      // th - 1 - cy)
      if (use_box) {
         BoxEpi ep{};
         ep.C = F0;
         ep.zstride = px;
         ep.src = zs;
         ep.ds = zds;
         launch_box_f64<kBoxSource>(s, ze, m, W, H, tw, th, tw - 1 - tw / 2, th - 1 - th / 2,
                                    false, ep);
      } else {
         for (int k = 0; k < m; ++k) {
            const int i = c0 + k;
            if (!dS[i]) continue;
            const View ev{E0 + (size_t)k * px, W, H, W * 4, 1, SImage::F32};
            launch_integral(s, ev, 0, false, I, 0, -1, aux);
            grad_source<<<dim3(grid_for((long long)px), 1), kThreads, 0, s>>>(
                F0 + (size_t)k * px, one(S[i].ptr, S[i].step), I, 0, W, H, tw, th,
                one(dS[i], dsstep[i]));
            launch_check("grad_source");
         }
      }
   }
   if (dT) {
      const int nblocks = n * L.nb, ngroups = (nblocks + kRedBlocks - 1) / kRedBlocks;
      double *red = reinterpret_cast<double *>(w + L.red);
      grad_partials<<<dim3((taps + 63) / 64, ngroups), kThreads, 0, s>>>(partial, nblocks, taps, red);
      launch_check("grad_partials");
      double *sc = sums + 6;
      grad_scalars<<<1, kThreads, 0, s>>>(sncc, G1 ? n * nscal : 0, sssd, G2 ? n * nscal : 0, sc);
      launch_check("grad_scalars");
      grad_template<<<(taps + 63) / 64, kThreads, 0, s>>>(red, ngroups, sc, tstage, tpitch,
                                                         tsums, tw, th, dT, dtstep);
      launch_check("grad_template");
   }
}

}  // namespace ocip

// proximity.hip -- direct (VALU) template matching for every type and channel count.
//
// Semantics (reference cl-files/ImageProximity.cl:44-466):
//   D(x,y) = sum_{v<th} sum_{u<tw} g(s(x+u-tw/2, y+v-th/2), t(u,v)),  s = 0 outside the image,
//   g = (s-t)^2 (SqrDistance), |s-t| (AbsDistance), s*t (CrossCorr);
//   the Norm ops divide by sqrt(sum s^2 * sum t^2) over the same zero-padded window.
//
// Kernels:
//   prox_tile_f32  REAL = float sources (F32, U32, S32).  Each workgroup owns a 64-row x
//                  (4*R)-column output tile; the source tile plus halo is staged in LDS
//                  (zero padded, converted to float), one band of template rows at a time.
//                  A thread owns R horizontally adjacent outputs of one row (lanes = rows,
//                  waves = column groups), keeps a sliding window of R+U-1 source values in
//                  registers and takes template values as wave-uniform scalar loads, so every
//                  LDS float feeds up to U FMAs and every template float R FMAs.  Each output
//                  accumulates in the reference order (template row outer, column inner) with
//                  one fused multiply-add per tap, which is the reference's order under
//                  OpenCL's default FP contraction.
//   prox_tile_int  8- and 16-bit integer sources: the same tiling with exact integer sums
//                  (int32 where the bound allows, int64 otherwise), rounded once to float.
//   prox_direct    one output per thread, reference order, any type (F64 sources and
//                  templates too wide for the LDS tile).
// Norm denominators come from the FP64 sum-of-squares integral of the source (exclusive,
// (W+1) x (H+1) layout) by four corners and the FP64 sum of squares of the template.
#include "device_common.h"

namespace ocip {

namespace {

constexpr int kTileRows = 64;      // output rows per workgroup (one per lane)
constexpr int kWaves = 4;          // waves per workgroup, side by side in x
constexpr int kLdsBudget = 64 * 1024;

// ------------------------------------------------------------------------- template staging
template <int TYPE, int KIND>
__global__ __launch_bounds__(256) void template_prep(const char *tpl, int tw, int th, int tstep,
                                                      int nch, int ch, int pitch, void *out,
                                                      double *sums)
{
   __shared__ double red[2][4];
   double s2 = 0, s1 = 0;
   const int n = pitch * th;
   for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const int v = i / pitch, u = i - v * pitch;
      double val = 0;
      if (u < tw) val = (double)ld<TYPE>(tpl + (size_t)v * tstep, u * nch + ch);
      if (KIND == 0) static_cast<float *>(out)[i] = (float)val;
      else if (KIND == 1) static_cast<double *>(out)[i] = val;
      else static_cast<int32_t *>(out)[i] = (int32_t)val;
      s2 += val * val;
      s1 += val;
   }
   s2 = wave_sum(s2);
   s1 = wave_sum(s1);
   const int wave = threadIdx.x >> 6;
   if ((threadIdx.x & 63) == 0) { red[0][wave] = s2; red[1][wave] = s1; }
   __syncthreads();
   if (threadIdx.x == 0) {
      sums[0] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
      sums[1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
   }
}


// ------------------------------------------------------------------------- per-tap maths
template <int OP> __device__ __forceinline__ float tap(float acc, float s, float t)
{
   if (OP == kSSD || OP == kSSDNorm) { float d = s - t; return __builtin_fmaf(d, d, acc); }
   if (OP == kSAD) return acc + __builtin_fabsf(s - t);
   return __builtin_fmaf(s, t, acc);
}

// U template columns starting at LDS column `col` (row pointer `lrow`) for R outputs.
// ALIGNED: col - SHIFT is a multiple of 4 floats, so the window is read with 16-byte LDS
// loads from there (SHIFT values before col read and not used).
template <int OP, int R, int U, bool ALIGNED, int SHIFT = 0>
__device__ __forceinline__ void chunk_f32(float (&acc)[R], const float *lrow, int col,
                                          const float *__restrict__ trow)
{
   constexpr int N = R + U - 1;
   float s[(N + SHIFT + 3) / 4 * 4];
   if (ALIGNED) {
#pragma unroll
      for (int i = 0; i < (N + SHIFT + 3) / 4; ++i) {
         float4 q = *reinterpret_cast<const float4 *>(lrow + col - SHIFT + 4 * i);
         s[4 * i] = q.x; s[4 * i + 1] = q.y; s[4 * i + 2] = q.z; s[4 * i + 3] = q.w;
      }
   } else {
#pragma unroll
      for (int i = 0; i < N; ++i) s[i + SHIFT] = lrow[col + i];
   }
   float t[U];
#pragma unroll
   for (int k = 0; k < U; ++k) t[k] = trow[k];
#pragma unroll
   for (int k = 0; k < U; ++k)
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] = tap<OP>(acc[r], s[r + k + SHIFT], t[k]);
}

template <int TYPE>
__device__ __forceinline__ void stage_tile_f32(float *lds, int ldsw, int nrows, int ncols,
                                               const char *src, int W, int H, int sstep,
                                               int nch, int ch, int sx0, int sy0)
{
   // eight elements per thread in flight; element i = r ldsw + c, (r, c) stepped along with i
   // (256 elements a step) instead of divided out for each element
   const int n = nrows * ldsw;
   const int dq = 256 / ldsw, dr = 256 - dq * ldsw;
   int r = threadIdx.x / ldsw, c = threadIdx.x - r * ldsw;
   for (int i0 = threadIdx.x; i0 < n; i0 += 256 * 8) {
      float v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
         const int i = i0 + 256 * k;
         const int sx = sx0 + c, sy = sy0 + r;
         v[k] = (i < n && c < ncols && sx >= 0 && sx < W && sy >= 0 && sy < H)
                    ? (float)ld<TYPE>(src + (size_t)sy * sstep, sx * nch + ch) : 0.f;
         c += dr;
         r += dq;
         if (c >= ldsw) {
            c -= ldsw;
            ++r;
         }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k)
         if (i0 + 256 * k < n) lds[i0 + 256 * k] = v[k];
   }
}

// stage_tile_f32 for 1-channel F32 rows whose staged columns start 16-byte aligned (sx0 a
// multiple of 4, 16-byte aligned rows): groups of four columns, one 16-byte load each where
// the group lies wholly inside the image, eight groups per thread in flight; the same values
// in LDS.
__device__ __forceinline__ void stage_tile_f32_v4(float *lds, int ldsw, int nrows, int ncols,
                                                  const char *src, int W, int H, int sstep,
                                                  int sx0, int sy0)
{
   const int gpr = ldsw / 4, n4 = nrows * gpr;
   // group i = r gpr + g, (r, g) stepped along with i (256 groups a step) instead of divided out
   const int dq = 256 / gpr, dr = 256 - dq * gpr;
   int r = threadIdx.x / gpr, g = threadIdx.x - r * gpr;
   for (int i0 = threadIdx.x; i0 < n4; i0 += 256 * 8) {
      float4 v[8];
      int at[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
         const int i = i0 + 256 * k;
         const int c = 4 * g;
         at[k] = r * ldsw + c;
         const int sx = sx0 + c, sy = sy0 + r;
         const bool rowin = i < n4 && sy >= 0 && sy < H;
         const float *srow = reinterpret_cast<const float *>(src + (size_t)(rowin ? sy : 0) * sstep);
         if (rowin && c + 4 <= ncols && sx >= 0 && sx + 4 <= W) {
            v[k] = *reinterpret_cast<const float4 *>(srow + sx);
         } else {
            float e[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
               e[j] = (rowin && c + j < ncols && sx + j >= 0 && sx + j < W) ? srow[sx + j] : 0.f;
            v[k] = make_float4(e[0], e[1], e[2], e[3]);
         }
         g += dr;
         r += dq;
         if (g >= gpr) {
            g -= gpr;
            ++r;
         }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
         const int i = i0 + 256 * k;
         if (i >= n4) break;
         *reinterpret_cast<float4 *>(lds + at[k]) = v[k];
      }
   }
}

// Window sums of s^2 of a tile's outputs from the staged tile (F32 values of rows y0 - cy ..
// y0 - cy + kTileRows + th - 2, columns x0 - cx .. x0 - cx + kWaves R + tw - 2, zero outside
// the image), in FP64, for the Norm ops when the whole template fits one band.  Two passes
// through LDS past the staged tile (tile_sqsum_lds bytes):
//   h(rho, q)  = sum over u < tw of s(q + u, rho)^2    (rho < kTileRows + th - 1, q < kWaves R)
//   SS(y, q)   = sum over v < th of h(y + v, q)         (y < kTileRows)
// each as the difference of two running sums (lead - lag) of exact FP64 squares along its
// row or column, started at the tile (a window of zeros gives exactly 0).  Returns SS of the
// lane's row y = lane and columns colbase .. colbase + R - 1.  Every thread of the workgroup
// takes part (two barriers).
__host__ __device__ constexpr size_t tile_sqsum_offset(int rows, int ldsw)
{
   return ((size_t)rows * ldsw * sizeof(float) + 15) / 16 * 16;
}
size_t tile_sqsum_lds(int R, int th)
{
   return (size_t)(kTileRows + th - 1 + kTileRows) * kWaves * R * sizeof(double);
}

template <int R>
__device__ __forceinline__ void tile_sqsums(float *lds, int ldsw, int tw, int th, int lane,
                                            int colbase, double (&ss)[R])
{
   constexpr int C4 = kWaves * R;
   const int nrows = kTileRows + th - 1;
   double *hb = reinterpret_cast<double *>(reinterpret_cast<char *>(lds) +
                                           tile_sqsum_offset(nrows, ldsw));
   double *sb = hb + (size_t)nrows * C4;
   // h: one task per staged row and R-column group
   for (int t = threadIdx.x; t < nrows * kWaves; t += 256) {
      const int rho = t / kWaves, q0 = (t - rho * kWaves) * R;
      const float *row = lds + (size_t)rho * ldsw + q0;
      double lead = 0.0, lag = 0.0;
      for (int u = 0; u < tw; ++u) lead = fma((double)row[u], (double)row[u], lead);
      double *hrow = hb + (size_t)rho * C4 + q0;
      hrow[0] = lead;
#pragma unroll
      for (int r = 1; r < R; ++r) {
         lead = fma((double)row[tw + r - 1], (double)row[tw + r - 1], lead);
         lag = fma((double)row[r - 1], (double)row[r - 1], lag);
         hrow[r] = lead - lag;
      }
   }
   __syncthreads();
   // SS: 256 threads = C4 columns x (256 / C4) groups of consecutive output rows
   {
      constexpr int kGroups = 256 / C4, kRows = kTileRows / kGroups;
      static_assert(256 % C4 == 0 && kTileRows % kGroups == 0, "row groups");
      const int q = threadIdx.x % C4, y0 = (threadIdx.x / C4) * kRows;
      double lead = 0.0, lag = 0.0;
      for (int v = 0; v < th; ++v) lead += hb[(size_t)(y0 + v) * C4 + q];
      sb[(size_t)y0 * C4 + q] = lead;
      for (int y = y0 + 1; y < y0 + kRows; ++y) {
         lead += hb[(size_t)(y + th - 1) * C4 + q];
         lag += hb[(size_t)(y - 1) * C4 + q];
         sb[(size_t)y * C4 + q] = lead - lag;
      }
   }
   __syncthreads();
#pragma unroll
   for (int r = 0; r < R; ++r) ss[r] = sb[(size_t)lane * C4 + colbase + r];
}

template <int OP, int TYPE, int R>
__global__ __launch_bounds__(256) void prox_tile_f32(const char *src, int W, int H, int sstep,
                                                      int nch, int ch,
                                                      const float *__restrict__ tp, int tpitch,
                                                      int tw, int th,
                                                      const double *__restrict__ tsums,
                                                      const double *__restrict__ integ,
                                                      char *dst, int dstep, int vb, int ldsw,
                                                      int cx, int cy)
{
   extern __shared__ float lds[];
   constexpr int U = 8;
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   // one-dimensional grid in XCD-aware tile order
   const int tiles_x = (W + kWaves * R - 1) / (kWaves * R);
   const int tile = xcd_tile(blockIdx.x, gridDim.x);
   const int x0 = (tile % tiles_x) * (kWaves * R), y0 = (tile / tiles_x) * kTileRows;
   const int ncols = kWaves * R + tw - 1;
   const int nfull = tw / U, rem = tw - nfull * U;
   const int colbase = wave * R;

   float acc[R];
#pragma unroll
   for (int r = 0; r < R; ++r) acc[r] = 0.f;
   // 16-byte staging loads: 1-channel F32 rows, staged columns from x0 - cx with x0 a
   // multiple of 4 R and cx a multiple of 4, 16-byte aligned rows; for the 16-output tiles of
   // large images (2.24 -> 2.09-2.12 ms at 4096^2 / 64x64)
   const bool v4 = R == 16 && TYPE == SImage::F32 && nch == 1 && (cx & 3) == 0 &&
                   (sstep & 15) == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0;

   for (int v0 = 0; v0 < th; v0 += vb) {
      const int nv = min(vb, th - v0);
      __syncthreads();
      if (v4)
         stage_tile_f32_v4(lds, ldsw, kTileRows + nv - 1, ncols, src, W, H, sstep, x0 - cx,
                           y0 - cy + v0);
      else
         stage_tile_f32<TYPE>(lds, ldsw, kTileRows + nv - 1, ncols, src, W, H, sstep, nch, ch,
                              x0 - cx, y0 - cy + v0);
      __syncthreads();
      for (int v = 0; v < nv; ++v) {
         const float *lrow = lds + (size_t)(lane + v) * ldsw;
         const float *trow = tp + (size_t)(v0 + v) * tpitch;
         int u0 = 0;
         for (int c = 0; c < nfull; ++c, u0 += U)
            chunk_f32<OP, R, U, true>(acc, lrow, colbase + u0, trow + u0);
         if (rem & 4) { chunk_f32<OP, R, 4, true>(acc, lrow, colbase + u0, trow + u0); u0 += 4; }
         // the 2- and 1-column chunks from 16-byte reads of a 16-byte aligned column (4-byte
         // LDS reads of consecutive rows would meet in the same banks), the 1-column chunk after
         // a 2-column one from two columns before its own: the reads end at most at column
         // ncols + 2, inside the row pitch
         if (rem & 2) { chunk_f32<OP, R, 2, true>(acc, lrow, colbase + u0, trow + u0); u0 += 2; }
         if (rem & 1) {
            if (rem & 2) chunk_f32<OP, R, 1, true, 2>(acc, lrow, colbase + u0, trow + u0);
            else chunk_f32<OP, R, 1, true>(acc, lrow, colbase + u0, trow + u0);
         }
      }
   }

   const int oy = y0 + lane;
   if (OP == kSSDNorm || OP == kCCNorm) {
      const double tt = tsums[0];
      double ss[R];
      if (integ) {
         // all integral corners of the R outputs first, then the divisions
#pragma unroll
         for (int r = 0; r < R; ++r) ss[r] = window_sqsum(integ, W, H, x0 + colbase + r, oy, tw, th);
      } else {
         // the whole template was staged in one band (the host passes no integral only
         // then): the window sums of s^2 from the staged tile, in FP64
         tile_sqsums<R>(lds, ldsw, tw, th, lane, colbase, ss);
      }
      if (oy >= H) return;
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] = (float)((double)acc[r] / sqrt(ss[r] * tt));
   }
   if (oy >= H) return;
   float *drow = reinterpret_cast<float *>(dst + (size_t)oy * dstep);
#pragma unroll
   for (int r = 0; r < R; ++r) {
      const int ox = x0 + colbase + r;
      if (ox < W) drow[ox * nch + ch] = acc[r];
   }
}

// ------------------------------------------------------------------------- exact integers
template <int TYPE>
__device__ __forceinline__ void stage_tile_int(int32_t *lds, int ldsw, int nrows, int ncols,
                                               const char *src, int W, int H, int sstep,
                                               int nch, int ch, int sx0, int sy0)
{
   const int n = nrows * ldsw;
   // element i = r ldsw + c, (r, c) stepped along with i (256 elements a step) instead of
   // divided out for each element
   const int dq = 256 / ldsw, dr = 256 - dq * ldsw;
   int r = threadIdx.x / ldsw, c = threadIdx.x - r * ldsw;
   for (int i0 = threadIdx.x; i0 < n; i0 += 256 * 8) {
      int32_t v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
         const int i = i0 + 256 * k;
         const int sx = sx0 + c, sy = sy0 + r;
         v[k] = (i < n && c < ncols && sx >= 0 && sx < W && sy >= 0 && sy < H)
                    ? (int32_t)ld<TYPE>(src + (size_t)sy * sstep, sx * nch + ch) : 0;
         c += dr;
         r += dq;
         if (c >= ldsw) {
            c -= ldsw;
            ++r;
         }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k)
         if (i0 + 256 * k < n) lds[i0 + 256 * k] = v[k];
   }
}

// One tap.  int32 sums come from 8-bit sources, whose values and differences fit the 24-bit
// multiplies (v_mad_i32_i24), and for SAD from 16-bit sources too; SAD takes v_sad_u16 (the
// absolute differences of both 16-bit halves plus c; its operands are kept in 0..65535, see
// sad_key, so the high halves are zero).
template <int OP, class ACC> __device__ __forceinline__ ACC itap(ACC acc, int32_t s, int32_t t)
{
   if constexpr (sizeof(ACC) == 4) {
      if (OP == kSSD || OP == kSSDNorm) { int32_t d = s - t; return acc + __mul24(d, d); }
      if (OP == kSAD) return (ACC)__builtin_amdgcn_sad_u16((uint32_t)s, (uint32_t)t, (uint32_t)acc);
      return acc + __mul24(s, t);
   } else {
      if (OP == kSSD || OP == kSSDNorm) { int32_t d = s - t; return acc + (ACC)d * (ACC)d; }
      if (OP == kSAD) { int32_t d = s - t; return acc + (ACC)(d < 0 ? -d : d); }
      return acc + (ACC)s * (ACC)t;
   }
}

// Operand of itap: for SAD with int32 sums on a signed source, the value + 128 (8-bit) or
// + 32768 (16-bit), which maps it in order onto 0..255 or 0..65535 (a sign-extended
// negative value would put 0xffff in v_sad_u16's high half); any other value unchanged.
template <int OP, int TYPE, class ACC> __device__ __forceinline__ int32_t sad_key(int32_t v)
{
   if (OP != kSAD || sizeof(ACC) != 4) return v;
   return TYPE == SImage::S8 ? v + 128 : TYPE == SImage::S16 ? v + 32768 : v;
}

constexpr int kIntU = 8;   // template columns per unrolled chunk of prox_tile_int

// Four 8-bit values (low bytes of a..d) in one dword, a in the low byte.
__device__ __forceinline__ uint32_t pack4(int32_t a, int32_t b, int32_t c, int32_t d)
{
   const uint32_t lo = __builtin_amdgcn_perm((uint32_t)b, (uint32_t)a, 0x0C0C0400u);
   const uint32_t hi = __builtin_amdgcn_perm((uint32_t)d, (uint32_t)c, 0x0C0C0400u);
   return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

// Four taps at once on packed bytes: v_dot4_u32_u8 / v_dot4_i32_i8 for the cross term,
// v_sad_u8 (bytes kept in 0..255 by sad_key) for AbsDistance.
template <int OP, int TYPE> __device__ __forceinline__ int32_t itap4(int32_t acc, uint32_t s, uint32_t t)
{
   if (OP == kSAD) return (int32_t)__builtin_amdgcn_sad_u8(s, t, (uint32_t)acc);
   if (TYPE == SImage::S8) return __builtin_amdgcn_sdot4((int32_t)s, (int32_t)t, acc, false);
   return (int32_t)__builtin_amdgcn_udot4(s, t, (uint32_t)acc, false);
}

// AbsDistance on 8-bit sources: the staged tile's keyed values (sad_key) repacked in place as
// P(c) = bytes (v[c], v[c + 4], v[c + 8], v[c + 12]) for the columns c < ncols (the tile was
// staged ncols + 12 columns wide, within the pitch), so that the 16 taps u0 .. u0 + 15 of the
// output at window column x are four v_sad_u8 of P(x + k) against the template bytes
// (t[u0 + k], t[u0 + k + 4], t[u0 + k + 8], t[u0 + k + 12]), k < 4, with no byte
// realignment.  Rounds of 8 elements per thread in increasing order: every read of a round
// (elements i .. i + 12) precedes its writes (elements i) at a barrier, and later rounds read
// only elements no earlier round wrote.
template <int TYPE>
__device__ __forceinline__ void pack_sad8(int32_t *lds, int ldsw, int nrows, int ncols)
{
   const int n = nrows * ldsw;
   // column c = i % ldsw stepped along with i (256 elements a step)
   const int dr = 256 % ldsw;
   int c = threadIdx.x % ldsw;
   for (int i0 = 0; i0 < n; i0 += 256 * 8) {
      uint32_t p[8];
      bool in[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
         const int i = i0 + threadIdx.x + 256 * k;
         in[k] = i < n && c < ncols;
         p[k] = 0;
         if (in[k])
            p[k] = pack4(sad_key<kSAD, TYPE, int32_t>(lds[i]), sad_key<kSAD, TYPE, int32_t>(lds[i + 4]),
                         sad_key<kSAD, TYPE, int32_t>(lds[i + 8]),
                         sad_key<kSAD, TYPE, int32_t>(lds[i + 12]));
         c += dr;
         if (c >= ldsw) c -= ldsw;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 8; ++k) {
         const int i = i0 + threadIdx.x + 256 * k;
         if (in[k]) lds[i] = (int32_t)p[k];
      }
      __syncthreads();
   }
}

// AbsDistance on 8-bit sources: the staged template (int32 per column, row pitch tpitch) gets,
// in each block of 16 columns, the four dwords of the block's keyed bytes that the repacked
// tile meets, in place at the block's first four columns: dword k = bytes of columns 16 b + k,
// + 4, + 8, + 12, a zero byte for a column past the template (a last partial block).  One
// thread per (row, block).
__host__ __device__ constexpr int sad8_blocks(int tw) { return (tw + 15) / 16; }

template <int TYPE>
__global__ __launch_bounds__(256) void sad8_template_pack(int32_t *tp, int tpitch, int tw, int th)
{
   const int nb = sad8_blocks(tw);
   const int i = blockIdx.x * 256 + threadIdx.x;
   if (i >= nb * th) return;
   const int v = i / nb, b = i - v * nb;
   int32_t *blk = tp + (size_t)v * tpitch + 16 * b;
   int32_t t[16];
#pragma unroll
   for (int k = 0; k < 16; ++k) t[k] = 16 * b + k < tw ? sad_key<kSAD, TYPE, int32_t>(blk[k]) : 0;
#pragma unroll
   for (int k = 0; k < 4; ++k) blk[k] = (int32_t)pack4(t[k], t[k + 4], t[k + 8], t[k + 12]);
}

// With int32 sums (8-bit sources) SqrDistance and SqrDistance_Norm take the cross term only,
// one multiply-add per tap, and finish as sum s^2 - 2 sum s*t + sum t^2: the window's sum of
// squares from the FP64 integral (an exact integer below 2^53), sum t^2 from tsums[0].
template <int OP, class ACC> constexpr bool ssd_via_cc()
{
   return (OP == kSSD || OP == kSSDNorm) && sizeof(ACC) == 4;
}

template <int OP, int TYPE, int R, class ACC>
__global__ __launch_bounds__(256) void prox_tile_int(const char *src, int W, int H, int sstep,
                                                      int nch, int ch,
                                                      const int32_t *__restrict__ tp, int tpitch,
                                                      int tw, int th,
                                                      const double *__restrict__ tsums,
                                                      const double *__restrict__ integ,
                                                      char *dst, int dstep, int vb, int ldsw)
{
   extern __shared__ int32_t ldsi[];
   constexpr bool kViaCC = ssd_via_cc<OP, ACC>();
   constexpr int TOP = kViaCC ? kCC : OP;   // the op of each tap
   // 8-bit sources with int32 sums: four taps per instruction on packed bytes
   constexpr bool kPacked = sizeof(ACC) == 4 && (TYPE == SImage::U8 || TYPE == SImage::S8) &&
                            (TOP == kCC || TOP == kCCNorm || TOP == kSAD);
   // AbsDistance on 8-bit sources: the tile repacked by pack_sad8 (the host sized the pitch
   // for ncols + 12 staged columns), the template's full 16-column blocks by
   // sad8_template_pack
   constexpr bool kSad8 = kPacked && TOP == kSAD;
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int x0 = blockIdx.x * (kWaves * R), y0 = blockIdx.y * kTileRows;
   const int cx = tw / 2, cy = th / 2;
   const int ncols = kWaves * R + tw - 1;
   const int colbase = wave * R;

   ACC acc[R];
#pragma unroll
   for (int r = 0; r < R; ++r) acc[r] = 0;

   for (int v0 = 0; v0 < th; v0 += vb) {
      const int nv = min(vb, th - v0);
      __syncthreads();
      stage_tile_int<TYPE>(ldsi, ldsw, kTileRows + nv - 1, kSad8 ? ncols + 12 : ncols, src, W,
                           H, sstep, nch, ch, x0 - cx, y0 - cy + v0);
      __syncthreads();
      if constexpr (kSad8) pack_sad8<TYPE>(ldsi, ldsw, kTileRows + nv - 1, ncols);
      for (int v = 0; v < nv; ++v) {
         const int32_t *lrow = ldsi + (size_t)(lane + v) * ldsw + colbase;
         const int32_t *trow = tp + (size_t)(v0 + v) * tpitch;
         if constexpr (kSad8) {
            // 16 taps per step: the window's P values of columns u0 .. u0 + R + 2 (16-byte
            // reads from a 16-byte aligned column: ldsw, colbase and u0 are multiples of 4),
            // four v_sad_u8 per output; the last tw % 16 taps in one masked step of the same
            // kind
            int u0 = 0;
            for (; u0 + 16 <= tw; u0 += 16) {
               uint32_t pw[R + 4];
               // (an empty asm with a memory clobber: each step reads its window afresh as 16-byte
               // loads, rather than the compiler carrying one dword over from the step before
               // and reading the rest as conflicting 4-byte loads)
               asm volatile("" ::: "memory");
               const int4 *w4 = reinterpret_cast<const int4 *>(lrow + u0);
#pragma unroll
               for (int m = 0; m < (R + 4) / 4; ++m) {
                  const int4 q4 = w4[m];
                  pw[4 * m] = (uint32_t)q4.x; pw[4 * m + 1] = (uint32_t)q4.y;
                  pw[4 * m + 2] = (uint32_t)q4.z; pw[4 * m + 3] = (uint32_t)q4.w;
               }
               // the block's four packed template dwords (sad8_template_pack)
               uint32_t tq[4];
#pragma unroll
               for (int k = 0; k < 4; ++k) tq[k] = (uint32_t)trow[u0 + k];
#pragma unroll
               for (int r = 0; r < R; ++r)
#pragma unroll
                  for (int k = 0; k < 4; ++k)
                     acc[r] = (ACC)__builtin_amdgcn_sad_u8(pw[r + k], tq[k], (uint32_t)acc[r]);
            }
            if (u0 < tw) {
               // a last block of 1 to 15 taps: the same step, each P masked to the bytes of
               // taps inside the template (its packed template dword holds zeros past it), so
               // that every other byte adds |0 - 0|, and the dwords k >= rem, which hold no
               // tap, skipped.  The window reads stay inside the row pitch: columns up to
               // colbase + u0 + 19 <= ncols + 3.
               const int rem = tw - u0;
               asm volatile("" ::: "memory");
               uint32_t pw[R + 4];
               const int4 *w4 = reinterpret_cast<const int4 *>(lrow + u0);
#pragma unroll
               for (int m = 0; m < (R + 4) / 4; ++m) {
                  const int4 q4 = w4[m];
                  pw[4 * m] = (uint32_t)q4.x; pw[4 * m + 1] = (uint32_t)q4.y;
                  pw[4 * m + 2] = (uint32_t)q4.z; pw[4 * m + 3] = (uint32_t)q4.w;
               }
               uint32_t tq[4], mk[4];
#pragma unroll
               for (int k = 0; k < 4; ++k) {
                  tq[k] = (uint32_t)trow[u0 + k];
                  // bytes j of tap u0 + k + 4 j < tw: (rem - k + 3) / 4 of them
                  const int nbytes = (rem - k + 3) >> 2;
                  mk[k] = nbytes >= 4 ? 0xffffffffu : (1u << (8 * nbytes)) - 1u;
               }
#pragma unroll
               for (int k = 0; k < 4; ++k) {
                  if (k >= rem) break;
#pragma unroll
                  for (int r = 0; r < R; ++r)
                     acc[r] = (ACC)__builtin_amdgcn_sad_u8(pw[r + k] & mk[k], tq[k], (uint32_t)acc[r]);
               }
            }
            continue;
         }
         // the template row in chunks of kIntU columns: kIntU template values and the
         // R + kIntU - 1 window values they meet, then R x kIntU taps (exact integer sums,
         // so the order is free); LDS columns used: colbase .. colbase + R + tw - 2 (the
         // 16-byte reads below take up to one more, inside the row pitch)
         int u0 = 0;
         if constexpr (kPacked) {
            // the window as bytes: aligned dwords q[m] = sw[4m .. 4m+3], the dword starting
            // at byte j = 4m + a by v_alignbyte_b32 of q[m+1]:q[m].  sw is read as 16-byte
            // LDS loads (lrow + u0 is 16-byte aligned: ldsw, colbase and u0 are multiples of
            // 4); the last of them takes one column past the R + kIntU - 1 used, at most
            // column ncols, inside the row pitch (ldsw >= ncols + 4), and its value is
            // never used.
            constexpr int NQ = (R + kIntU - 1 + 3) / 4;
            for (; u0 + kIntU <= tw; u0 += kIntU) {
               int32_t sw[4 * NQ];
               const int4 *w4 = reinterpret_cast<const int4 *>(lrow + u0);
#pragma unroll
               for (int m = 0; m < NQ; ++m) {
                  const int4 v4 = w4[m];
                  sw[4 * m] = v4.x; sw[4 * m + 1] = v4.y; sw[4 * m + 2] = v4.z; sw[4 * m + 3] = v4.w;
               }
#pragma unroll
               for (int j = 0; j < 4 * NQ; ++j) sw[j] = sad_key<TOP, TYPE, ACC>(sw[j]);
               uint32_t q[NQ], tq[kIntU / 4];
#pragma unroll
               for (int m = 0; m < NQ; ++m) q[m] = pack4(sw[4 * m], sw[4 * m + 1], sw[4 * m + 2], sw[4 * m + 3]);
#pragma unroll
               for (int g = 0; g < kIntU / 4; ++g)
                  tq[g] = pack4(sad_key<TOP, TYPE, ACC>(trow[u0 + 4 * g]),
                                sad_key<TOP, TYPE, ACC>(trow[u0 + 4 * g + 1]),
                                sad_key<TOP, TYPE, ACC>(trow[u0 + 4 * g + 2]),
                                sad_key<TOP, TYPE, ACC>(trow[u0 + 4 * g + 3]));
#pragma unroll
               for (int g = 0; g < kIntU / 4; ++g)
#pragma unroll
                  for (int r = 0; r < R; ++r) {
                     const int j = r + 4 * g;
                     const uint32_t pj = (j & 3) == 0 ? q[j >> 2]
                                         : __builtin_amdgcn_alignbyte(q[(j >> 2) + 1], q[j >> 2], j & 3);
                     acc[r] = itap4<TOP, TYPE>(acc[r], pj, tq[g]);
                  }
            }
         }
         for (; u0 + kIntU <= tw; u0 += kIntU) {
            // the same 16-byte window reads as above
            constexpr int NQ = (R + kIntU - 1 + 3) / 4;
            int32_t t[kIntU], sw[4 * NQ];
#pragma unroll
            for (int k = 0; k < kIntU; ++k) t[k] = sad_key<TOP, TYPE, ACC>(trow[u0 + k]);
            const int4 *w4 = reinterpret_cast<const int4 *>(lrow + u0);
#pragma unroll
            for (int m = 0; m < NQ; ++m) {
               const int4 v4 = w4[m];
               sw[4 * m] = v4.x; sw[4 * m + 1] = v4.y; sw[4 * m + 2] = v4.z; sw[4 * m + 3] = v4.w;
            }
#pragma unroll
            for (int j = 0; j < R + kIntU - 1; ++j) sw[j] = sad_key<TOP, TYPE, ACC>(sw[j]);
#pragma unroll
            for (int k = 0; k < kIntU; ++k)
#pragma unroll
               for (int r = 0; r < R; ++r) acc[r] = itap<TOP, ACC>(acc[r], sw[r + k], t[k]);
         }
         if (u0 < tw) {
            // the last tw % kIntU taps (1 to kIntU - 1) from 16-byte reads of the aligned column
            // u0 too (4-byte LDS reads of consecutive rows would meet in the same banks); the
            // reads end at most at column colbase + R + tw + 1 <= ncols + 2, inside the row
            // pitch
            const int rem = tw - u0;
            constexpr int NQ = (R + kIntU - 1 + 3) / 4;
            int32_t t[kIntU - 1], sw[4 * NQ];
            const int4 *w4 = reinterpret_cast<const int4 *>(lrow + u0);
#pragma unroll
            for (int m = 0; m < NQ; ++m) {
               int4 v4 = make_int4(0, 0, 0, 0);
               if (4 * m < R + rem - 1) v4 = w4[m];
               sw[4 * m] = v4.x; sw[4 * m + 1] = v4.y; sw[4 * m + 2] = v4.z; sw[4 * m + 3] = v4.w;
            }
#pragma unroll
            for (int j = 0; j < R + kIntU - 2; ++j) sw[j] = sad_key<TOP, TYPE, ACC>(sw[j]);
#pragma unroll
            for (int k = 0; k < kIntU - 1; ++k) t[k] = k < rem ? sad_key<TOP, TYPE, ACC>(trow[u0 + k]) : 0;
#pragma unroll
            for (int k = 0; k < kIntU - 1; ++k) {
               if (k >= rem) break;
#pragma unroll
               for (int r = 0; r < R; ++r) acc[r] = itap<TOP, ACC>(acc[r], sw[r + k], t[k]);
            }
         }
      }
   }

   const int oy = y0 + lane;
   if (oy >= H) return;
   float *drow = reinterpret_cast<float *>(dst + (size_t)oy * dstep);
   float val[R];
   if constexpr (kViaCC || OP == kSSDNorm || OP == kCCNorm) {
      // the sums are integers, exact in FP64 below 2^53 (always, for 8-bit sources).
      // Four outputs at a time, each group
      // stored before the next group's integral corners are loaded (the scheduling
      // barriers keep them apart): all sixteen outputs' corners at once would hold about
      // 100 more VGPRs and cost a wave per SIMD
      const double tt = tsums[0];
#pragma unroll
      for (int g = 0; g < R; g += 4) {
         double ss[4];
#pragma unroll
         for (int r = 0; r < 4; ++r)
            ss[r] = window_sqsum(integ, W, H, x0 + colbase + g + r, oy, tw, th);
#pragma unroll
         for (int r = 0; r < 4; ++r) {
            const double c = (double)acc[g + r];
            const double d = kViaCC ? ss[r] - 2.0 * c + tt : c;
            const float v = OP == kSSD ? (float)d : (float)(d / sqrt(ss[r] * tt));
            const int ox = x0 + colbase + g + r;
            if (ox < W) drow[ox * nch + ch] = v;
         }
         __builtin_amdgcn_sched_barrier(0);
      }
      return;
   } else {
#pragma unroll
      for (int r = 0; r < R; ++r) val[r] = (float)acc[r];
   }
#pragma unroll
   for (int r = 0; r < R; ++r) {
      const int ox = x0 + colbase + r;
      if (ox < W) drow[ox * nch + ch] = val[r];
   }
}

// ------------------------------------------------------------------------- direct
template <int OP, int TYPE, class REALT, class STAGED>
__global__ __launch_bounds__(256) void prox_direct(const char *src, int W, int H, int sstep,
                                                    int nch, int ch,
                                                    const STAGED *__restrict__ tp, int tpitch,
                                                    int tw, int th,
                                                    const double *__restrict__ tsums,
                                                    const double *__restrict__ integ,
                                                    char *dst, int dstep)
{
   const int x = blockIdx.x * 16 + (threadIdx.x & 15);
   const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
   if (x >= W || y >= H) return;
   const int left = -(tw / 2), top = -(th / 2);
   REALT sum = 0;
   for (int h = 0; h < th; ++h) {
      const int py = y + h + top;
      const bool rowin = py >= 0 && py < H;
      const char *srow = src + (size_t)(rowin ? py : 0) * sstep;
      for (int w = 0; w < tw; ++w) {
         const int px = x + w + left;
         const REALT t = (REALT)tp[(size_t)h * tpitch + w];
         REALT s = 0;
         if (rowin && px >= 0 && px < W) s = (REALT)ld<TYPE>(srow, px * nch + ch);
         if (OP == kSSD || OP == kSSDNorm) { REALT d = s - t; sum = fma(d, d, sum); }
         else if (OP == kSAD) sum += fabs(s - t);
         else sum = fma(s, t, sum);
      }
   }
   REALT val = sum;
   if (OP == kSSDNorm || OP == kCCNorm)
      val = (REALT)((double)sum / sqrt(window_sqsum(integ, W, H, x, y, tw, th) * tsums[0]));
   reinterpret_cast<REALT *>(dst + (size_t)y * dstep)[x * nch + ch] = val;
}

int lds_pitch(int ncols)
{
   // multiple of 4 floats (16-byte loads) and = 4 mod 32 so that lanes reading the same
   // column of consecutive rows start on different banks
   int p = (ncols + 4 + 3) / 4 * 4;
   while (p % 32 != 4) p += 4;
   return p;
}

}  // namespace

void launch_template_prep(hipStream_t s, const View &t, int channel, int pitch, int kind,
                          void *out, double *sums)
{
   KernelTimer kt(s, "template_prep");
   OCIP_TYPE_SWITCH(t.type, TT, {
      const char *p = static_cast<const char *>(t.ptr);
      if (kind == 0)
         template_prep<TT, 0><<<1, 256, 0, s>>>(p, t.width, t.height, t.step, t.channels,
                                                channel, pitch, out, sums);
      else if (kind == 1)
         template_prep<TT, 1><<<1, 256, 0, s>>>(p, t.width, t.height, t.step, t.channels,
                                                channel, pitch, out, sums);
      else
         template_prep<TT, 2><<<1, 256, 0, s>>>(p, t.width, t.height, t.step, t.channels,
                                                channel, pitch, out, sums);
   });
   launch_check("template_prep");
}

// Output columns per lane of the tiled kernels
static int tile_r(int W, int tw, int th) { return (W >= 2048 || tw * th >= 1024) ? 16 : 8; }

bool prox_norm_from_tile(int op, const View &S, const View &T)
{
   if ((op != kSSDNorm && op != kCCNorm) || S.type != SImage::F32 || T.width <= 0 || T.height <= 0)
      return false;
   const int R = tile_r(S.width, T.width, T.height);
   const int ldsw = lds_pitch(kWaves * R + T.width - 1);
   const int rows_fit = kLdsBudget / (ldsw * 4) - (kTileRows - 1);
   if (rows_fit < T.height) return false;   // one band
   return tile_sqsum_offset(kTileRows + T.height - 1, ldsw) + tile_sqsum_lds(R, T.height) <=
          (size_t)kLdsBudget;
}

template <int OP>
static void launch_op(hipStream_t s, const View &S, const View &T, void *dst, int dstep,
                      const void *tstage, int tpitch, const double *tsums, const double *integ,
                      int ch)
{
   const int W = S.width, H = S.height, tw = T.width, th = T.height;
   const char *src = static_cast<const char *>(S.ptr);
   const bool small_int = S.type == SImage::U8 || S.type == SImage::S8 ||
                          S.type == SImage::U16 || S.type == SImage::S16;
   const int R = tile_r(W, tw, th);
   const int ncols = kWaves * R + tw - 1;
   // int32 accumulation is exact when taps * max|term| < 2^31: 8-bit sources up to 33025
   // taps, and AbsDistance on 16-bit sources up to 32768 taps; the other 16-bit cases
   // accumulate in int64.
   const bool bits8 = S.type == SImage::U8 || S.type == SImage::S8;
   const long long taps = (long long)tw * th;
   const bool fits32 = bits8 ? taps * 65025LL < (1LL << 31)
                             : OP == kSAD && taps * 65535LL < (1LL << 31);
   // AbsDistance on 8-bit sources with int32 sums stages 12 more columns (pack_sad8)
   const int ldsw = lds_pitch(ncols + (OP == kSAD && bits8 && fits32 ? 12 : 0));
   const int rows_fit = kLdsBudget / (ldsw * 4) - (kTileRows - 1);
   const bool tiled = S.type != SImage::F64 && rows_fit >= 1 && th > 0 && tw > 0;

   if (!tiled) {
      KernelTimer kt(s, "prox_direct");
      dim3 grid((W + 15) / 16, (H + 15) / 16);
      if (S.type == SImage::F64) {
         prox_direct<OP, SImage::F64, double, double><<<grid, 256, 0, s>>>(
             src, W, H, S.step, S.channels, ch, static_cast<const double *>(tstage), tpitch, tw,
             th, tsums, integ, static_cast<char *>(dst), dstep);
      } else if (small_int) {
         OCIP_TYPE_SWITCH(S.type, TT, {
            prox_direct<OP, TT, float, int32_t><<<grid, 256, 0, s>>>(
                src, W, H, S.step, S.channels, ch, static_cast<const int32_t *>(tstage), tpitch,
                tw, th, tsums, integ, static_cast<char *>(dst), dstep);
         });
      } else {
         OCIP_TYPE_SWITCH(S.type, TT, {
            prox_direct<OP, TT, float, float><<<grid, 256, 0, s>>>(
                src, W, H, S.step, S.channels, ch, static_cast<const float *>(tstage), tpitch,
                tw, th, tsums, integ, static_cast<char *>(dst), dstep);
         });
      }
      launch_check("prox_direct");
      return;
   }

   const int vb = min(th, rows_fit);
   size_t lds = (size_t)(kTileRows + vb - 1) * ldsw * 4;
   dim3 grid((W + kWaves * R - 1) / (kWaves * R), (H + kTileRows - 1) / kTileRows);
   // Norm ops on F32 sources with no integral: window sums from the staged tile
   if ((OP == kSSDNorm || OP == kCCNorm) && !integ) {
      if (!prox_norm_from_tile(OP, S, T)) fail(CL_INVALID_VALUE, "integral required");
      lds = tile_sqsum_offset(kTileRows + vb - 1, ldsw) + tile_sqsum_lds(R, th);
   }

#define OCIP_TILE_F32(TT_, R_)                                                                 \
   prox_tile_f32<OP, TT_, R_><<<grid.x * grid.y, 256, lds, s>>>(                                         \
       src, W, H, S.step, S.channels, ch, static_cast<const float *>(tstage), tpitch, tw, th, \
       tsums, integ, static_cast<char *>(dst), dstep, vb, ldsw, tw / 2, th / 2)
#define OCIP_TILE_INT(TT_, R_, ACC_)                                                             \
   prox_tile_int<OP, TT_, R_, ACC_><<<grid, 256, lds, s>>>(                                     \
       src, W, H, S.step, S.channels, ch, static_cast<const int32_t *>(tstage), tpitch, tw, th, \
       tsums, integ, static_cast<char *>(dst), dstep, vb, ldsw)

   if (small_int) {
      if (OP == kSAD && bits8 && fits32 && sad8_blocks(tw) > 0) {
         // the packed template blocks of the tile's AbsDistance loop (in the staged template,
         // which the next call stages again)
         KernelTimer kt(s, "template_prep");
         const int items = sad8_blocks(tw) * th;
         int32_t *tp = static_cast<int32_t *>(const_cast<void *>(tstage));
         OCIP_TYPE_SWITCH(S.type, TT, {
            if constexpr (TT == SImage::U8 || TT == SImage::S8)
               sad8_template_pack<TT><<<(items + 255) / 256, 256, 0, s>>>(tp, tpitch, tw, th);
         });
         launch_check("sad8_template_pack");
      }
      KernelTimer kt(s, "prox_tile_int");
      OCIP_TYPE_SWITCH(S.type, TT, {
         if (fits32) {
            if (R == 16) OCIP_TILE_INT(TT, 16, int32_t); else OCIP_TILE_INT(TT, 8, int32_t);
         } else {
            if (R == 16) OCIP_TILE_INT(TT, 16, long long); else OCIP_TILE_INT(TT, 8, long long);
         }
      });
      launch_check("prox_tile_int");
   } else {
      KernelTimer kt(s, "prox_tile_f32");
      switch (S.type) {
      case SImage::F32: if (R == 16) OCIP_TILE_F32(SImage::F32, 16); else OCIP_TILE_F32(SImage::F32, 8); break;
      case SImage::U32: if (R == 16) OCIP_TILE_F32(SImage::U32, 16); else OCIP_TILE_F32(SImage::U32, 8); break;
      case SImage::S32: if (R == 16) OCIP_TILE_F32(SImage::S32, 16); else OCIP_TILE_F32(SImage::S32, 8); break;
      default: fail(CL_IMAGE_FORMAT_NOT_SUPPORTED, "type");
      }
      launch_check("prox_tile_f32");
   }
#undef OCIP_TILE_F32
#undef OCIP_TILE_INT
}

void launch_proximity(hipStream_t s, int op, const View &S, const View &T, void *dst, int dstep,
                      const void *tstage, int tpitch, const double *tsums, const double *integ,
                      int ch)
{
   switch (op) {
   case kSSD: launch_op<kSSD>(s, S, T, dst, dstep, tstage, tpitch, tsums, integ, ch); break;
   case kSSDNorm: launch_op<kSSDNorm>(s, S, T, dst, dstep, tstage, tpitch, tsums, integ, ch); break;
   case kSAD: launch_op<kSAD>(s, S, T, dst, dstep, tstage, tpitch, tsums, integ, ch); break;
   case kCC: launch_op<kCC>(s, S, T, dst, dstep, tstage, tpitch, tsums, integ, ch); break;
   case kCCNorm: launch_op<kCCNorm>(s, S, T, dst, dstep, tstage, tpitch, tsums, integ, ch); break;
   default: fail(CL_INVALID_VALUE, "op");
   }
}

}  // namespace ocip

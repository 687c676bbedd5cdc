// integral.hip -- inclusive 2-D prefix sum / sum of squares in FP64.
//
// Semantics: reference Integral.cl / Integral.cpp (I(x,y) = sum over j<=y, i<=x of f(S[j][i]),
// same size as the source, f = identity or square; include/OpenCLIPP.h:639-645).  The
// reference sums in the Dest precision through a 5-launch, 64x8 block scan with serial
// junction loops (Integral.cl:39-460).  Here every sum is formed in FP64 (exact for 8/16-bit
// integer sources, whose partial sums stay below 2^53) and rounded once to the Dest type.
//
// Design: the image is cut into tiles of 32 rows x 256 columns, one wavefront per tile
// (4 columns per lane), so every scan is a wavefront shuffle scan with no barrier.  Three
// launches, with the source read twice and the output written once:
//   1. aggregates: per tile, its column sums, its row sums and its total;
//   2. carries   : exclusive scans of the column sums over tile rows, of the row sums over
//                  tile columns, and the 2-D exclusive scan of the tile totals;
//   3. final     : per tile, the in-tile 2-D scan (a running column sum per lane and one
//                  wavefront scan per row) plus the three carries.
// HBM traffic: 2 x source + 1 x output (+ small aggregates).
#include "device_common.h"

namespace ocip {

namespace {

constexpr int kThreads = 64;            // one wavefront per tile: no barriers
constexpr int kVec = 4;                 // columns per lane
constexpr int kCW = kThreads * kVec;    // tile width (256)
constexpr int kRH = 32;                 // tile height
constexpr int kRowUnroll = 8;           // rows whose loads are issued together

struct IntegralAux {
   // layout inside the aux buffer (doubles):
   //   colsum  [RB][W]     column sums per tile row, then exclusive-scanned over tile rows
   //   rowsum  [H][CB]     row sums per tile column, then exclusive-scanned over tile columns
   //   tiletot [RB][CB]    tile totals, then 2-D exclusive-scanned (corner carries)
   double *colsum, *rowsum, *tiletot;
   int CB, RB;
};

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }

IntegralAux aux_layout(void *base, int W, int H)
{
   IntegralAux a;
   a.CB = cdiv(W, kCW);
   a.RB = cdiv(H, kRH);
   double *p = static_cast<double *>(base);
   a.colsum = p;
   a.rowsum = a.colsum + (size_t)a.RB * W;
   a.tiletot = a.rowsum + (size_t)H * a.CB;
   return a;
}

template <class T> struct alignas(4 * sizeof(T)) Vec4 {
   T v[4];
};

// f(s) for the lane's 4 columns of row y (0 outside [0, y1) x [0, W)).  VEC: 1-channel rows
// whose 4-element groups are aligned, read with one vector load.
template <int TYPE, bool SQR, bool VEC>
__device__ __forceinline__ void load_row(const char *src, int step, int y, int y1, int x0, int W,
                                         int nch, int ch, double (&v)[kVec])
{
   typedef typename ElemT<TYPE>::type T;
   const char *row = src + (size_t)(y < y1 ? y : 0) * step;
   if (VEC && y < y1 && x0 + kVec <= W) {
      const Vec4<T> q = *reinterpret_cast<const Vec4<T> *>(row + (size_t)x0 * sizeof(T));
#pragma unroll
      for (int j = 0; j < kVec; ++j) {
         const double f = (double)q.v[j];
         v[j] = SQR ? f * f : f;
      }
      return;
   }
#pragma unroll
   for (int j = 0; j < kVec; ++j) {
      const int x = x0 + j;
      const double f = (y < y1 && x < W) ? (double)ld<TYPE>(row, x * nch + ch) : 0.0;
      v[j] = SQR ? f * f : f;
   }
}

// 1. per-tile aggregates: column sums, row sums, tile total
template <int TYPE, bool SQR, bool VEC>
__global__ __launch_bounds__(kThreads) void integral_aggregate(const char *src, int W, int H,
                                                                int step, int nch, int ch,
                                                                IntegralAux aux)
{
   // each lane's partial row sums, rows x lanes (odd pitch: lane r's walk along row r
   // spreads over the banks); lane r then sums row r in lane order
   __shared__ double part[kRH][kThreads + 1];
   static_assert(kRH % kRowUnroll == 0 && kRH <= kThreads, "rows of the tile");
   const int cb = blockIdx.x, rb = blockIdx.y;
   const int x0 = cb * kCW + threadIdx.x * kVec;
   const int y0 = rb * kRH, y1 = min(y0 + kRH, H);
   double col[kVec] = {0, 0, 0, 0};
   for (int y = y0; y < y1; y += kRowUnroll) {
      double v[kRowUnroll][kVec];
#pragma unroll
      for (int r = 0; r < kRowUnroll; ++r)
         load_row<TYPE, SQR, VEC>(src, step, y + r, y1, x0, W, nch, ch, v[r]);
#pragma unroll
      for (int r = 0; r < kRowUnroll; ++r) {
         double rs = 0;
#pragma unroll
         for (int j = 0; j < kVec; ++j) {
            col[j] += v[r][j];
            rs += v[r][j];
         }
         part[y - y0 + r][threadIdx.x] = rs;   // rows past y1 hold 0 and are not summed
      }
   }
#pragma unroll
   for (int j = 0; j < kVec; ++j)
      if (x0 + j < W) aux.colsum[(size_t)rb * W + x0 + j] = col[j];
   __syncthreads();
   const int r = threadIdx.x;
   double rs = 0;
   if (r < y1 - y0) {
      for (int l = 0; l < kThreads; ++l) rs += part[r][l];
      aux.rowsum[(size_t)(y0 + r) * aux.CB + cb] = rs;
   }
   const double tot = wave_sum(rs);
   if (threadIdx.x == 0) aux.tiletot[(size_t)rb * aux.CB + cb] = tot;
}

// exclusive scan in place of n values at p[0], p[stride], ...
__device__ __forceinline__ void exclusive_scan_strided(double *p, int n, size_t stride)
{
   // 32 loads in flight: the scans are chains of round trips (128 tile rows at 4096 rows)
   constexpr int B = 32;
   double acc = 0;
   for (int k = 0; k < n; k += B) {
      double v[B];
#pragma unroll
      for (int i = 0; i < B; ++i) v[i] = k + i < n ? p[(size_t)(k + i) * stride] : 0.0;
#pragma unroll
      for (int i = 0; i < B; ++i) {
         if (k + i < n) p[(size_t)(k + i) * stride] = acc;
         acc += v[i];
      }
   }
}

// 2. carries.  blockIdx.y = 0: column sums over tile rows, one thread per column;
//              blockIdx.y = 1: row sums over tile columns, one thread per row;
//              blockIdx.y = 2: 2-D exclusive scan of the tile totals (block 0 only).
__global__ __launch_bounds__(256) void integral_carries(int W, int H, IntegralAux aux)
{
   const int i = blockIdx.x * 256 + threadIdx.x;
   if (blockIdx.y == 0) {
      if (i < W) exclusive_scan_strided(aux.colsum + i, aux.RB, (size_t)W);
   } else if (blockIdx.y == 1) {
      if (i < H) exclusive_scan_strided(aux.rowsum + (size_t)i * aux.CB, aux.CB, 1);
   } else if (blockIdx.x == 0) {
      // along each tile row, then down each tile column:
      // corner[rb][cb] = sum over rb' < rb, cb' < cb of the tile totals
      for (int rb = threadIdx.x; rb < aux.RB; rb += 256)
         exclusive_scan_strided(aux.tiletot + (size_t)rb * aux.CB, aux.CB, 1);
      __syncthreads();
      for (int cb = threadIdx.x; cb < aux.CB; cb += 256)
         exclusive_scan_strided(aux.tiletot + cb, aux.RB, (size_t)aux.CB);
   }
}

// 4 output values of a row at column x0 (j = 0..3), vector store when OVEC and in range
template <class T, bool OVEC>
__device__ __forceinline__ void store4(T *o, int x0, int W, const double (&val)[kVec])
{
   if (OVEC && x0 + kVec <= W) {
      Vec4<T> q;
#pragma unroll
      for (int j = 0; j < kVec; ++j) q.v[j] = (T)val[j];
      *reinterpret_cast<Vec4<T> *>(o + x0) = q;
      return;
   }
#pragma unroll
   for (int j = 0; j < kVec; ++j)
      if (x0 + j < W) o[x0 + j] = (T)val[j];
}

// 3. final pass.  OUT: 0 = F32 inclusive, 1 = F64 inclusive (both with the caller's pitch),
// 2 = the internal exclusive layout.  OVEC: output rows allow aligned 4-element stores.
template <int TYPE, bool SQR, bool VEC, int OUT, bool OVEC>
__global__ __launch_bounds__(kThreads) void integral_final(const char *src, int W, int H,
                                                            int step, int nch, int ch,
                                                            IntegralAux aux, char *out,
                                                            int ostep)
{
   const int cb = blockIdx.x, rb = blockIdx.y;
   const int x0 = cb * kCW + threadIdx.x * kVec;
   const int y0 = rb * kRH, y1 = min(y0 + kRH, H);

   // carry from the tiles above: prefix over this tile's columns of the column carries
   double px[kVec];
   {
      double loc = 0;
#pragma unroll
      for (int j = 0; j < kVec; ++j) {
         const int x = x0 + j;
         loc += x < W ? aux.colsum[(size_t)rb * W + x] : 0.0;
         px[j] = loc;
      }
      const double excl = wave_inclusive_scan(loc) - loc;
#pragma unroll
      for (int j = 0; j < kVec; ++j) px[j] += excl;
   }
   const double corner = aux.tiletot[(size_t)rb * aux.CB + cb];
   const size_t P = excl_pitch(W);

   if (OUT == 2) {
      // row 0 and column 0 of the exclusive layout are 0
      double *o = reinterpret_cast<double *>(out) + kExclOffset;
      if (rb == 0) {
#pragma unroll
         for (int j = 0; j < kVec; ++j)
            if (x0 + j < W) o[x0 + j + 1] = 0.0;
         if (cb == 0 && threadIdx.x == 0) o[0] = 0.0;
      }
      if (cb == 0)
         for (int y = y0 + threadIdx.x; y < y1; y += kThreads) o[(size_t)(y + 1) * P] = 0.0;
   }

   double colrun[kVec] = {0, 0, 0, 0};
   double py = 0;
   for (int y = y0; y < y1; y += kRowUnroll) {
      double v[kRowUnroll][kVec];
      double rc[kRowUnroll];
#pragma unroll
      for (int r = 0; r < kRowUnroll; ++r) {
         load_row<TYPE, SQR, VEC>(src, step, y + r, y1, x0, W, nch, ch, v[r]);
         rc[r] = y + r < y1 ? aux.rowsum[(size_t)(y + r) * aux.CB + cb] : 0.0;
      }
      // the lane's prefixes of the batch's rows, then their wavefront scans together
      double p[kRowUnroll][kVec], inc[kRowUnroll];
#pragma unroll
      for (int r = 0; r < kRowUnroll; ++r) {
         double loc = 0;
#pragma unroll
         for (int j = 0; j < kVec; ++j) {
            colrun[j] += v[r][j];
            loc += colrun[j];
            p[r][j] = loc;
         }
         inc[r] = loc;
      }
      wave_incl_scan_f64<kRowUnroll>(inc);
#pragma unroll
      for (int r = 0; r < kRowUnroll; ++r) {
         py += rc[r];
         const double base = corner + py + (inc[r] - p[r][kVec - 1]);
         const int yy = y + r;
         if (yy >= y1) continue;
         double val[kVec];
#pragma unroll
         for (int j = 0; j < kVec; ++j) val[j] = base + p[r][j] + px[j];
         if (OUT == 0)
            store4<float, OVEC>(reinterpret_cast<float *>(out + (size_t)yy * ostep), x0, W, val);
         else if (OUT == 1)
            store4<double, OVEC>(reinterpret_cast<double *>(out + (size_t)yy * ostep), x0, W, val);
         else
            store4<double, true>(reinterpret_cast<double *>(out) + (size_t)(yy + 1) * P +
                                     kExclOffset + 1, x0, W, val);
      }
   }
}

}  // namespace

size_t integral_aux_bytes(int W, int H)
{
   size_t CB = cdiv(W, kCW), RB = cdiv(H, kRH);
   return (RB * W + (size_t)H * CB + RB * CB) * sizeof(double);
}

template <int TT, bool VEC>
static void launch_integral_t(hipStream_t s, const char *src, const View &v, int channel,
                              bool sqr, void *out, int ostep, int mode, bool ovec, IntegralAux aux)
{
   const int W = v.width, H = v.height;
   dim3 grid(aux.CB, aux.RB);
   KernelTimer kt(s, "integral");
   if (sqr)
      integral_aggregate<TT, true, VEC><<<grid, kThreads, 0, s>>>(src, W, H, v.step, v.channels,
                                                                 channel, aux);
   else
      integral_aggregate<TT, false, VEC><<<grid, kThreads, 0, s>>>(src, W, H, v.step, v.channels,
                                                                  channel, aux);
   launch_check("integral_aggregate");
   integral_carries<<<dim3(cdiv(W > H ? W : H, 256), 3), 256, 0, s>>>(W, H, aux);
   launch_check("integral_carries");
   char *o = static_cast<char *>(out);
#define OCIP_FINAL(SQ, M, OV) \
   integral_final<TT, SQ, VEC, M, OV><<<grid, kThreads, 0, s>>>(src, W, H, v.step, v.channels, channel, aux, o, ostep)
#define OCIP_FINAL_M(SQ)                                                    \
   if (mode == 2) OCIP_FINAL(SQ, 2, true);                                  \
   else if (mode == 1) { if (ovec) OCIP_FINAL(SQ, 1, true); else OCIP_FINAL(SQ, 1, false); } \
   else { if (ovec) OCIP_FINAL(SQ, 0, true); else OCIP_FINAL(SQ, 0, false); }
   if (sqr) { OCIP_FINAL_M(true) } else { OCIP_FINAL_M(false) }
#undef OCIP_FINAL_M
#undef OCIP_FINAL
   launch_check("integral_final");
}

void launch_integral(hipStream_t s, const View &v, int channel, bool sqr, void *out, int ostep,
                     int out_type, void *auxp)
{
   IntegralAux aux = aux_layout(auxp, v.width, v.height);
   const char *src = static_cast<const char *>(v.ptr);
   const int mode = out_type == SImage::F32 ? 0 : (out_type == SImage::F64 ? 1 : 2);
   const int esz = elem_size(v.type);
   // 4-element vector loads: one channel, rows and base aligned to 4 elements
   const bool vec = v.channels == 1 && (v.step % (4 * esz)) == 0 &&
                    (reinterpret_cast<uintptr_t>(v.ptr) % (4 * esz)) == 0;
   const int osz = mode == 0 ? 4 : 8;
   const bool ovec = mode != 2 && (ostep % (4 * osz)) == 0 &&
                     (reinterpret_cast<uintptr_t>(out) % (4 * osz)) == 0;
   OCIP_TYPE_SWITCH(v.type, TT, {
      if (vec) launch_integral_t<TT, true>(s, src, v, channel, sqr, out, ostep, mode, ovec, aux);
      else launch_integral_t<TT, false>(s, src, v, channel, sqr, out, ostep, mode, ovec, aux);
   });
}

}  // namespace ocip

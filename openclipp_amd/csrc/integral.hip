// integral.hip -- inclusive 2-D prefix sum / sum of squares in FP64.
//
// Semantics: reference Integral.cl / Integral.cpp (I(x,y) = sum over j<=y, i<=x of f(S[j][i]),
// same size as the source, f = identity or square; include/OpenCLIPP.h:639-645).  The
// reference sums in the Dest precision through a 5-launch, 64x8 block scan with serial
// junction loops (Integral.cl:39-460).  Here every sum is formed in FP64 (exact for 8/16-bit
// integer sources, whose partial sums stay below 2^53) and rounded once to the Dest type.
//
// Design: the image is cut into tiles of RH rows x CW columns (CW = 4 columns per thread x
// 256 threads).  Three launches, with the source read twice and the output written once:
//   1. aggregates: per tile, its column sums, its row sums and its total;
//   2. carries   : exclusive scans of the column sums over tile rows, of the row sums over
//                  tile columns, and the 2-D exclusive scan of the tile totals;
//   3. final     : per tile, the in-tile 2-D scan (a running column sum per thread and one
//                  block-wide scan per row) plus the three carries.
// HBM traffic: 2 x source + 1 x output (+ small aggregates).
#include "device_common.h"

namespace ocip {

namespace {

constexpr int kThreads = 256;
constexpr int kVec = 4;                 // columns per thread
constexpr int kCW = kThreads * kVec;    // tile width
constexpr int kRH = 64;                 // tile height

struct IntegralAux {
   // layout inside the aux buffer (doubles):
   //   colsum  [RB][W]     column sums per tile row, then exclusive-scanned in place
   //   rowsum  [H][CB]     row sums per tile column, then exclusive-scanned in place
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

template <int TYPE, bool SQR>
__device__ __forceinline__ double fval(const char *row, int x, int nch, int ch)
{
   double v = (double)ld<TYPE>(row, x * nch + ch);
   return SQR ? v * v : v;
}

// Block-wide inclusive scan of one value per thread (256 threads = 4 waves).
__device__ __forceinline__ double block_inclusive_scan(double v, double *lds4)
{
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   v = wave_inclusive_scan(v);
   if (lane == 63) lds4[wave] = v;
   __syncthreads();
   double off = 0;
   for (int w = 0; w < wave; ++w) off += lds4[w];
   __syncthreads();
   return v + off;
}

__device__ __forceinline__ double block_sum(double v, double *lds4)
{
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   v = wave_sum(v);
   if (lane == 0) lds4[wave] = v;
   __syncthreads();
   double t = lds4[0] + lds4[1] + lds4[2] + lds4[3];
   __syncthreads();
   return t;
}

// 1. per-tile aggregates
template <int TYPE, bool SQR>
__global__ __launch_bounds__(kThreads) void integral_aggregate(const char *src, int W, int H,
                                                                int step, int nch, int ch,
                                                                IntegralAux aux)
{
   __shared__ double lds4[4];
   const int cb = blockIdx.x, rb = blockIdx.y;
   const int x0 = cb * kCW + threadIdx.x * kVec;
   const int y0 = rb * kRH, y1 = min(y0 + kRH, H);
   double col[kVec] = {0, 0, 0, 0};
   double tot = 0;
   for (int y = y0; y < y1; ++y) {
      const char *row = src + (size_t)y * step;
      double r = 0;
#pragma unroll
      for (int j = 0; j < kVec; ++j) {
         int x = x0 + j;
         double v = x < W ? fval<TYPE, SQR>(row, x, nch, ch) : 0.0;
         col[j] += v;
         r += v;
      }
      r = block_sum(r, lds4);
      if (threadIdx.x == 0) aux.rowsum[(size_t)y * aux.CB + cb] = r;
      tot += r;
   }
#pragma unroll
   for (int j = 0; j < kVec; ++j)
      if (x0 + j < W) aux.colsum[(size_t)rb * W + x0 + j] = col[j];
   if (threadIdx.x == 0) aux.tiletot[(size_t)rb * aux.CB + cb] = tot;
}

// 2. carries: exclusive scans of the aggregates
__global__ void integral_carries(int W, int H, IntegralAux aux)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < W) {
      double acc = 0;
      for (int rb = 0; rb < aux.RB; ++rb) {
         double *p = aux.colsum + (size_t)rb * W + i;
         double v = *p;
         *p = acc;
         acc += v;
      }
   } else if (i < W + H) {
      const int y = i - W;
      double acc = 0;
      for (int cb = 0; cb < aux.CB; ++cb) {
         double *p = aux.rowsum + (size_t)y * aux.CB + cb;
         double v = *p;
         *p = acc;
         acc += v;
      }
   } else if (i == W + H) {
      // 2-D exclusive prefix of the tile totals, in place: first along each tile row ...
      for (int rb = 0; rb < aux.RB; ++rb) {
         double rowacc = 0;
         for (int cb = 0; cb < aux.CB; ++cb) {
            double *p = aux.tiletot + (size_t)rb * aux.CB + cb;
            double v = *p;
            *p = rowacc;
            rowacc += v;
         }
      }
      // ... then down each tile column: corner[rb][cb] = sum_{rb'<rb, cb'<cb} tot[rb'][cb']
      for (int cb = 0; cb < aux.CB; ++cb) {
         double acc = 0;
         for (int rb = 0; rb < aux.RB; ++rb) {
            double *p = aux.tiletot + (size_t)rb * aux.CB + cb;
            double v = *p;
            *p = acc;
            acc += v;
         }
      }
   }
}

// 3. final pass.  OUT: 0 = F32 inclusive, 1 = F64 inclusive, 2 = F64 exclusive padded.
template <int TYPE, bool SQR, int OUT>
__global__ __launch_bounds__(kThreads) void integral_final(const char *src, int W, int H,
                                                            int step, int nch, int ch,
                                                            IntegralAux aux, char *out,
                                                            int ostep)
{
   __shared__ double lds4[4];
   const int cb = blockIdx.x, rb = blockIdx.y;
   const int x0 = cb * kCW + threadIdx.x * kVec;
   const int y0 = rb * kRH, y1 = min(y0 + kRH, H);

   // carry from the tiles above: prefix over the tile's columns of colsum (exclusive over rb)
   double px[kVec];
   {
      double loc = 0;
#pragma unroll
      for (int j = 0; j < kVec; ++j) {
         int x = x0 + j;
         loc += x < W ? aux.colsum[(size_t)rb * W + x] : 0.0;
         px[j] = loc;
      }
      double incl = block_inclusive_scan(loc, lds4);
      double excl = incl - loc;
#pragma unroll
      for (int j = 0; j < kVec; ++j) px[j] += excl;
   }
   const double corner = aux.tiletot[(size_t)rb * aux.CB + cb];

   if (OUT == 2) {
      // zero row 0 and column 0 of the padded layout
      if (rb == 0) {
         double *o = reinterpret_cast<double *>(out);
#pragma unroll
         for (int j = 0; j < kVec; ++j)
            if (x0 + j < W) o[x0 + j + 1] = 0.0;
         if (cb == 0 && threadIdx.x == 0) o[0] = 0.0;
      }
      if (cb == 0 && threadIdx.x == 0)
         for (int y = y0; y < y1; ++y)
            reinterpret_cast<double *>(out + (size_t)(y + 1) * ostep)[0] = 0.0;
   }

   double colrun[kVec] = {0, 0, 0, 0};
   double py = 0;
   for (int y = y0; y < y1; ++y) {
      const char *row = src + (size_t)y * step;
      py += aux.rowsum[(size_t)y * aux.CB + cb];
      double p[kVec];
      double loc = 0;
#pragma unroll
      for (int j = 0; j < kVec; ++j) {
         int x = x0 + j;
         colrun[j] += x < W ? fval<TYPE, SQR>(row, x, nch, ch) : 0.0;
         loc += colrun[j];
         p[j] = loc;
      }
      double incl = block_inclusive_scan(loc, lds4);
      double base = corner + py + (incl - loc);
      if (OUT == 0) {
         float *o = reinterpret_cast<float *>(out + (size_t)y * ostep);
#pragma unroll
         for (int j = 0; j < kVec; ++j)
            if (x0 + j < W) o[x0 + j] = (float)(base + p[j] + px[j]);
      } else if (OUT == 1) {
         double *o = reinterpret_cast<double *>(out + (size_t)y * ostep);
#pragma unroll
         for (int j = 0; j < kVec; ++j)
            if (x0 + j < W) o[x0 + j] = base + p[j] + px[j];
      } else {
         double *o = reinterpret_cast<double *>(out + (size_t)(y + 1) * ostep);
#pragma unroll
         for (int j = 0; j < kVec; ++j)
            if (x0 + j < W) o[x0 + j + 1] = base + p[j] + px[j];
      }
   }
}

}  // namespace

size_t integral_aux_bytes(int W, int H)
{
   size_t CB = cdiv(W, kCW), RB = cdiv(H, kRH);
   return (RB * W + (size_t)H * CB + RB * CB) * sizeof(double);
}

template <int TT>
static void launch_integral_t(hipStream_t s, const char *src, const View &v, int channel,
                              bool sqr, void *out, int ostep, int mode, IntegralAux aux)
{
   const int W = v.width, H = v.height;
   dim3 grid(aux.CB, aux.RB);
   KernelTimer kt(s, "integral");
   if (sqr)
      integral_aggregate<TT, true><<<grid, kThreads, 0, s>>>(src, W, H, v.step, v.channels,
                                                            channel, aux);
   else
      integral_aggregate<TT, false><<<grid, kThreads, 0, s>>>(src, W, H, v.step, v.channels,
                                                             channel, aux);
   launch_check("integral_aggregate");
   integral_carries<<<cdiv(W + H + 1, 256), 256, 0, s>>>(W, H, aux);
   launch_check("integral_carries");
   char *o = static_cast<char *>(out);
#define OCIP_FINAL(SQ, M) \
   integral_final<TT, SQ, M><<<grid, kThreads, 0, s>>>(src, W, H, v.step, v.channels, channel, aux, o, ostep)
   if (sqr) {
      if (mode == 0) OCIP_FINAL(true, 0); else if (mode == 1) OCIP_FINAL(true, 1); else OCIP_FINAL(true, 2);
   } else {
      if (mode == 0) OCIP_FINAL(false, 0); else if (mode == 1) OCIP_FINAL(false, 1); else OCIP_FINAL(false, 2);
   }
#undef OCIP_FINAL
   launch_check("integral_final");
}

void launch_integral(hipStream_t s, const View &v, int channel, bool sqr, void *out, int ostep,
                     int out_type, void *auxp)
{
   IntegralAux aux = aux_layout(auxp, v.width, v.height);
   const char *src = static_cast<const char *>(v.ptr);
   const int mode = out_type == SImage::F32 ? 0 : (out_type == SImage::F64 ? 1 : 2);
   OCIP_TYPE_SWITCH(v.type, TT,
                    launch_integral_t<TT>(s, src, v, channel, sqr, out, ostep, mode, aux));
}

}  // namespace ocip

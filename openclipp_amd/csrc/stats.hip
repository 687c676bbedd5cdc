// stats.hip -- minimum / maximum of a 1-channel image and the pixel holding it; per-channel
// sum and sum of squares.
//
// Reference: Statistics::Min / Max (C++/programs/Statistics.cpp:244-270) with the
// min_coord / max_coord kernels (cl-files/Statistics.cl:342-517) and the host final reduce
// (C++/programs/StatisticsHelpers.cpp:126-145).  The reference keeps partial results as
// float and breaks ties by the order of its reduction tree; here values are compared in
// FP64 (exact for every element type) and ties go to the first pixel in row-major order.
// NaN pixels are skipped.  The call blocks until the result is on the host, as the
// reference's does.
//
// Sum / SumSqr: reference Statistics::Sum / SumSqr (C++/programs/Statistics.cpp:350-370) with
// reduce_sum / reduce_sum_sqr (cl-files/Statistics.cl:47-121, 279-288, 310-311) and the host
// final reduce in double (C++/programs/StatisticsHelpers.cpp:66-76).  The reference sums each
// work-group's 4096 pixels in REAL (float) and adds the float partials in double; here 8- and
// 16-bit sources are summed exactly in int64 and other types in FP64, in a fixed order (the
// result is deterministic), rounded once to double.
#include "device_common.h"

namespace ocip {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;

struct Cand {
   double v;
   long long i;   // linear index y*W + x; -1 = none
};

template <bool MAX> __device__ __forceinline__ bool better(const Cand &a, const Cand &b)
{
   if (a.i < 0) return false;
   if (b.i < 0) return true;
   if (MAX ? a.v > b.v : a.v < b.v) return true;
   if (a.v == b.v) return a.i < b.i;
   return false;
}

template <bool MAX> __device__ Cand block_best(Cand c)
{
   __shared__ double sv[kThreads / 64];
   __shared__ long long si[kThreads / 64];
#pragma unroll
   for (int d = 32; d >= 1; d >>= 1) {
      Cand o{__shfl_xor(c.v, d, 64), __shfl_xor(c.i, d, 64)};
      if (better<MAX>(o, c)) c = o;
   }
   const int wave = threadIdx.x >> 6;
   if ((threadIdx.x & 63) == 0) { sv[wave] = c.v; si[wave] = c.i; }
   __syncthreads();
   if (threadIdx.x == 0) {
      for (int w = 1; w < kThreads / 64; ++w) {
         Cand o{sv[w], si[w]};
         if (better<MAX>(o, c)) c = o;
      }
   }
   return c;
}

template <int TYPE, bool MAX>
__global__ __launch_bounds__(kThreads) void minmax_partial(const char *src, int W, int H,
                                                            int step, double *pv,
                                                            long long *pi)
{
   Cand best{0.0, -1};
   const long long n = (long long)W * H;
   for (long long k = (long long)blockIdx.x * kThreads + threadIdx.x; k < n;
        k += (long long)gridDim.x * kThreads) {
      const int y = (int)(k / W), x = (int)(k - (long long)y * W);
      double v = (double)ld<TYPE>(src + (size_t)y * step, x);
      if (v != v) continue;
      Cand c{v, k};
      if (better<MAX>(c, best)) best = c;
   }
   best = block_best<MAX>(best);
   if (threadIdx.x == 0) { pv[blockIdx.x] = best.v; pi[blockIdx.x] = best.i; }
}

template <bool MAX>
__global__ __launch_bounds__(kThreads) void minmax_final(int nparts, double *pv, long long *pi)
{
   Cand best{0.0, -1};
   for (int k = threadIdx.x; k < nparts; k += kThreads) {
      Cand c{pv[k], pi[k]};
      if (better<MAX>(c, best)) best = c;
   }
   best = block_best<MAX>(best);
   if (threadIdx.x == 0) { pv[nparts] = best.v; pi[nparts] = best.i; }
}

int nblocks(int W, int H)
{
   long long n = (long long)W * H;
   long long b = (n + kThreads * 8 - 1) / (kThreads * 8);
   return (int)(b < 1 ? 1 : (b > kMaxBlocks ? kMaxBlocks : b));
}

// ---- sum / sum of squares --------------------------------------------------------------------

template <int TYPE> struct SumAcc { typedef double type; };
template <> struct SumAcc<SImage::U8>  { typedef long long type; };
template <> struct SumAcc<SImage::S8>  { typedef long long type; };
template <> struct SumAcc<SImage::U16> { typedef long long type; };
template <> struct SumAcc<SImage::S16> { typedef long long type; };

// One partial per block and channel: grid-stride over pixels, then wave and block sums.
template <int TYPE, bool SQR, int NCH>
__global__ __launch_bounds__(kThreads) void sum_partial(const char *src, int W, int H, int step,
                                                         typename SumAcc<TYPE>::type *part)
{
   typedef typename SumAcc<TYPE>::type A;
   __shared__ A red[kThreads / 64][NCH];
   A acc[NCH];
#pragma unroll
   for (int c = 0; c < NCH; ++c) acc[c] = 0;
   const long long n = (long long)W * H;
   for (long long k = (long long)blockIdx.x * kThreads + threadIdx.x; k < n;
        k += (long long)gridDim.x * kThreads) {
      const int y = (int)(k / W), x = (int)(k - (long long)y * W);
      const char *row = src + (size_t)y * step;
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
         const A v = (A)ld<TYPE>(row, x * NCH + c);
         acc[c] += SQR ? v * v : v;
      }
   }
   const int wave = threadIdx.x >> 6;
#pragma unroll
   for (int c = 0; c < NCH; ++c) {
      const A w = wave_sum(acc[c]);
      if ((threadIdx.x & 63) == 0) red[wave][c] = w;
   }
   __syncthreads();
   if (threadIdx.x < NCH) {
      A t = 0;
      for (int w = 0; w < kThreads / 64; ++w) t += red[w][threadIdx.x];
      part[(size_t)blockIdx.x * NCH + threadIdx.x] = t;
   }
}

// Sum of the nparts partials of each channel into out[c] (as double), fixed order.
template <class A, int NCH>
__global__ __launch_bounds__(kThreads) void sum_final(const A *part, int nparts, double *out)
{
   __shared__ A red[kThreads / 64][NCH];
   A acc[NCH];
#pragma unroll
   for (int c = 0; c < NCH; ++c) acc[c] = 0;
   for (int k = threadIdx.x; k < nparts; k += kThreads)
#pragma unroll
      for (int c = 0; c < NCH; ++c) acc[c] += part[(size_t)k * NCH + c];
   const int wave = threadIdx.x >> 6;
#pragma unroll
   for (int c = 0; c < NCH; ++c) {
      const A w = wave_sum(acc[c]);
      if ((threadIdx.x & 63) == 0) red[wave][c] = w;
   }
   __syncthreads();
   if (threadIdx.x < NCH) {
      A t = 0;
      for (int w = 0; w < kThreads / 64; ++w) t += red[w][threadIdx.x];
      out[threadIdx.x] = (double)t;
   }
}

template <int TYPE, bool SQR, int NCH>
void launch_sum_t(hipStream_t s, const View &v, void *partials, int nb)
{
   typedef typename SumAcc<TYPE>::type A;
   A *part = static_cast<A *>(partials);
   double *out = reinterpret_cast<double *>(part + (size_t)nb * 4);
   sum_partial<TYPE, SQR, NCH><<<nb, kThreads, 0, s>>>(static_cast<const char *>(v.ptr), v.width,
                                                         v.height, v.step, part);
   launch_check("sum_partial");
   sum_final<A, NCH><<<1, kThreads, 0, s>>>(part, nb, out);
   launch_check("sum_final");
}

}  // namespace

size_t sum_partials_bytes(int W, int H) { return ((size_t)nblocks(W, H) * 4 + 4) * 8; }

void launch_sum(hipStream_t s, bool sqr, const View &v, void *partials, double *result)
{
   const int nb = nblocks(v.width, v.height);
   KernelTimer kt(s, "sum");
   const int nch = v.channels;
   if (nch != 1 && nch != 2 && nch != 4) fail(CL_INVALID_VALUE, "channels");
   OCIP_TYPE_SWITCH(v.type, TT, {
      if (sqr) {
         if (nch == 1) launch_sum_t<TT, true, 1>(s, v, partials, nb);
         else if (nch == 2) launch_sum_t<TT, true, 2>(s, v, partials, nb);
         else launch_sum_t<TT, true, 4>(s, v, partials, nb);
      } else {
         if (nch == 1) launch_sum_t<TT, false, 1>(s, v, partials, nb);
         else if (nch == 2) launch_sum_t<TT, false, 2>(s, v, partials, nb);
         else launch_sum_t<TT, false, 4>(s, v, partials, nb);
      }
   });
   const double *out = reinterpret_cast<const double *>(static_cast<char *>(partials) + (size_t)nb * 4 * 8);
   hip_check(hipMemcpyAsync(result, out, sizeof(double) * nch, hipMemcpyDeviceToHost, s), "read");
   hip_check(hipStreamSynchronize(s), "sync");
}

size_t minmax_partials_bytes(int W, int H)
{
   return (size_t)(nblocks(W, H) + 1) * (sizeof(double) + sizeof(long long));
}

void launch_minmax(hipStream_t s, bool want_max, const View &v, void *partials, double *result,
                   int *ix, int *iy)
{
   const int nb = nblocks(v.width, v.height);
   double *pv = static_cast<double *>(partials);
   long long *pi = reinterpret_cast<long long *>(pv + nb + 1);
   const char *src = static_cast<const char *>(v.ptr);
   KernelTimer kt(s, "minmax");
   OCIP_TYPE_SWITCH(v.type, TT, {
      if (want_max) minmax_partial<TT, true><<<nb, kThreads, 0, s>>>(src, v.width, v.height, v.step, pv, pi);
      else minmax_partial<TT, false><<<nb, kThreads, 0, s>>>(src, v.width, v.height, v.step, pv, pi);
   });
   launch_check("minmax_partial");
   if (want_max) minmax_final<true><<<1, kThreads, 0, s>>>(nb, pv, pi);
   else minmax_final<false><<<1, kThreads, 0, s>>>(nb, pv, pi);
   launch_check("minmax_final");
   double hv = 0;
   long long hi = -1;
   hip_check(hipMemcpyAsync(&hv, pv + nb, sizeof(double), hipMemcpyDeviceToHost, s), "read");
   hip_check(hipMemcpyAsync(&hi, pi + nb, sizeof(long long), hipMemcpyDeviceToHost, s), "read");
   hip_check(hipStreamSynchronize(s), "sync");
   if (hi < 0) {
      // every pixel is NaN
      *result = __builtin_nan("");
      *ix = 0;
      *iy = 0;
      return;
   }
   *result = hv;
   *iy = (int)(hi / v.width);
   *ix = (int)(hi - (long long)(*iy) * v.width);
}

}  // namespace ocip

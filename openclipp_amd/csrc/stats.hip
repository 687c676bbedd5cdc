// stats.hip -- minimum / maximum of a 1-channel image and the pixel holding it; per-channel
// sum and sum of squares.
//
// Reference: Statistics::Min / Max (C++/programs/Statistics.cpp:244-270) with the
// min_coord / max_coord kernels (cl-files/Statistics.cl:332-425, _flush variant :427-507) and
// the host final reduce (C++/programs/StatisticsHelpers.cpp:105-145).  Same pixels per
// work-item, same comparisons, same pairings and the same float partials, so ties and NaNs
// resolve to the reference's pixel:
//   * work-group (256 threads) = 256 columns x 16 rows; thread lid = 16 ly + lx scans row
//     16 gyg + ly at columns 256 gxg + lx + 16 i, i = 0..15, keeping the first of equal values;
//   * halving tree over lid (k = 128 .. 1): position lid takes lid + k only when both hold
//     pixels and lid + k is strictly better (k = 128, 64 through LDS, k <= 32 in wave 0 by
//     shuffles: positions < k only ever read positions < 2k, written at the level before);
//   * each group's extreme is kept as float with its pixel; the final pass takes, as the
//     reference's host loop does, the lowest group id among the best floats (group 0 if its
//     float is NaN, since the host loop never replaces a NaN it starts from).
// Values are compared as doubles (every element type converts exactly, so the comparisons are
// those of the element type) and the result is the float partial, as the reference returns.
// The call blocks until the result is on the host, as the reference's does.
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

// One group's (or the final) result: the extreme as float and its pixel.
struct MinMaxPart {
   float v;
   int x, y, pad;
};

template <bool MAX> __device__ __forceinline__ bool strictly_better(double b, double a)
{
   return MAX ? b > a : b < a;
}

template <int TYPE, bool MAX>
__global__ __launch_bounds__(256) void minmax_groups(const char *src, int W, int H, int step,
                                                      int ngx, MinMaxPart *part)
{
   __shared__ double sv[256];
   __shared__ int sx[256], sy[256], sok[256];
   const int lid = threadIdx.x, lx = lid & 15, ly = lid >> 4;
   const int gxg = blockIdx.x % ngx, gyg = blockIdx.x / ngx;
   const int gx = 256 * gxg + lx, gy = 16 * gyg + ly;
   int ok = gx < W && gy < H;
   double v = 0.0;
   int cx = gx, cy = gy;
   if (ok) {
      const char *row = src + (size_t)gy * step;
      v = (double)ld<TYPE>(row, gx);
#pragma unroll
      for (int i = 16; i < 256; i += 16)
         if (gx + i < W) {
            const double px = (double)ld<TYPE>(row, gx + i);
            if (strictly_better<MAX>(px, v)) { v = px; cx = gx + i; }
         }
   }
   sv[lid] = v; sx[lid] = cx; sy[lid] = cy; sok[lid] = ok;
   __syncthreads();
#pragma unroll
   for (int k = 128; k >= 64; k >>= 1) {
      if (lid < k && sok[lid] && sok[lid + k] && strictly_better<MAX>(sv[lid + k], sv[lid])) {
         sv[lid] = sv[lid + k]; sx[lid] = sx[lid + k]; sy[lid] = sy[lid + k];
      }
      __syncthreads();
   }
   if (lid < 64) {
      v = sv[lid]; cx = sx[lid]; cy = sy[lid]; ok = sok[lid];
#pragma unroll
      for (int k = 32; k >= 1; k >>= 1) {
         const double ov = __shfl_down(v, k, 64);
         const int ox = __shfl_down(cx, k, 64), oy = __shfl_down(cy, k, 64);
         const int ook = __shfl_down(ok, k, 64);
         if (lid < k && ok && ook && strictly_better<MAX>(ov, v)) { v = ov; cx = ox; cy = oy; }
      }
      if (lid == 0) part[blockIdx.x] = MinMaxPart{(float)v, cx, cy, 0};
   }
}

// First group holding the best float (NaN never taken), or group 0 when its float is NaN:
// the reference's serial host loop.  Each thread scans a strided subset keeping its first
// best; subsets merge by (value, lower group id), which is order-independent.
template <bool MAX>
__global__ __launch_bounds__(1024) void minmax_final(const MinMaxPart *part, int n,
                                                     MinMaxPart *out)
{
   __shared__ float rv[1024 / 64];
   __shared__ int rg[1024 / 64];
   float v = 0.0f;
   int g = 0x7fffffff;  // no candidate
   for (int k = threadIdx.x; k < n; k += blockDim.x) {
      const float pv = part[k].v;
      if (pv != pv) continue;
      if (g == 0x7fffffff || (MAX ? pv > v : pv < v)) { v = pv; g = k; }
   }
#pragma unroll
   for (int d = 32; d >= 1; d >>= 1) {
      const float ov = __shfl_xor(v, d, 64);
      const int og = __shfl_xor(g, d, 64);
      if (og != 0x7fffffff &&
          (g == 0x7fffffff || (MAX ? ov > v : ov < v) || (ov == v && og < g))) { v = ov; g = og; }
   }
   const int wave = threadIdx.x >> 6;
   if ((threadIdx.x & 63) == 0) { rv[wave] = v; rg[wave] = g; }
   __syncthreads();
   if (threadIdx.x == 0) {
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
         const float ov = rv[w];
         const int og = rg[w];
         if (og != 0x7fffffff &&
             (g == 0x7fffffff || (MAX ? ov > v : ov < v) || (ov == v && og < g))) { v = ov; g = og; }
      }
      const float p0 = part[0].v;
      if (p0 != p0 || g == 0x7fffffff) g = 0;
      *out = part[g];
   }
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

static int minmax_groups_count(int W, int H) { return ((W + 255) / 256) * ((H + 15) / 16); }

size_t minmax_partials_bytes(int W, int H)
{
   return (size_t)(minmax_groups_count(W, H) + 1) * sizeof(MinMaxPart);
}

void launch_minmax(hipStream_t s, bool want_max, const View &v, void *partials, double *result,
                   int *ix, int *iy)
{
   const int ngx = (v.width + 255) / 256;
   const int n = minmax_groups_count(v.width, v.height);
   MinMaxPart *part = static_cast<MinMaxPart *>(partials);
   const char *src = static_cast<const char *>(v.ptr);
   KernelTimer kt(s, "minmax");
   OCIP_TYPE_SWITCH(v.type, TT, {
      if (want_max) minmax_groups<TT, true><<<n, 256, 0, s>>>(src, v.width, v.height, v.step, ngx, part);
      else minmax_groups<TT, false><<<n, 256, 0, s>>>(src, v.width, v.height, v.step, ngx, part);
   });
   launch_check("minmax_groups");
   if (want_max) minmax_final<true><<<1, 1024, 0, s>>>(part, n, part + n);
   else minmax_final<false><<<1, 1024, 0, s>>>(part, n, part + n);
   launch_check("minmax_final");
   MinMaxPart h{};
   hip_check(hipMemcpyAsync(&h, part + n, sizeof h, hipMemcpyDeviceToHost, s), "read");
   hip_check(hipStreamSynchronize(s), "sync");
   *result = (double)h.v;
   *ix = h.x;
   *iy = h.y;
}

}  // namespace ocip

// backward.hip -- gradients of CrossCorr_Norm and SqrDistance (F32, 1 channel).
//
// The reference has no gradients; the forward maps are cl-files/ImageProximity.cl:44-73
// (SqrDistance) and :375-412 (CrossCorr_Norm).  With the template origin at (-cx, -cy),
// P the zero-padded source, G the gradient of the loss with respect to the output map,
// C = sum P t, SS = window sum of P^2 (from the FP64 integral), TT = sum t^2,
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
// Building blocks: weight_corr (sum over all output pixels of a map times the shifted
// source, one value per template tap, FP32 partials per tile reduced in FP64), the FP32
// correlation of a map with the flipped template (the plain tiled kernel with anchor
// (tw-1-cx, th-1-cy)), and flipped window sums from an FP64 integral of a map.
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

// Per-pixel maps of the CrossCorr_Norm gradient: Gd = G / D, E = G N / SS (0 where D = 0),
// and per-block partial sums of G N.
__global__ __launch_bounds__(kThreads) void ncc_maps(const float *C, int cstep, const char *G,
                                                      int gstep, const double *__restrict__ integ,
                                                      const double *__restrict__ tsums, int W,
                                                      int H, int tw, int th, float *Gd, float *E,
                                                      double *partial)
{
   __shared__ double red[4];
   const double tt = tsums[0];
   const int cx = tw / 2, cy = th / 2;
   double acc = 0;
   for (long long k = (long long)blockIdx.x * kThreads + threadIdx.x; k < (long long)W * H;
        k += (long long)gridDim.x * kThreads) {
      const int y = (int)(k / W), x = (int)(k - (long long)y * W);
      const double ss = rect_sum(integ, W, H, x - cx, x - cx + tw, y - cy, y - cy + th);
      const double d = sqrt(ss * tt);
      const double g = reinterpret_cast<const float *>(G + (size_t)y * gstep)[x];
      const double c = reinterpret_cast<const float *>(reinterpret_cast<const char *>(C) +
                                                       (size_t)y * cstep)[x];
      float gd = 0.f, e = 0.f;
      if (d > 0) {
         const double n = c / d;
         gd = (float)(g / d);
         e = (float)(g * n / ss);
         acc += g * n;
      }
      Gd[k] = gd;
      E[k] = e;
   }
   acc = block_sum_d(acc, red);
   if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// Per-block partial sums of a pitched F32 map.
__global__ __launch_bounds__(kThreads) void map_sum(const char *G, int gstep, int W, int H,
                                                     double *partial)
{
   __shared__ double red[4];
   double acc = 0;
   for (long long k = (long long)blockIdx.x * kThreads + threadIdx.x; k < (long long)W * H;
        k += (long long)gridDim.x * kThreads) {
      const int y = (int)(k / W), x = (int)(k - (long long)y * W);
      acc += reinterpret_cast<const float *>(G + (size_t)y * gstep)[x];
   }
   acc = block_sum_d(acc, red);
   if (threadIdx.x == 0) partial[blockIdx.x] = acc;
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

// dT(u, v) = a * sum_b partial[b][tap] + b * t(u, v) with (a, b) from the scalar sums:
//   mode 0 (CrossCorr_Norm): a = 1,  b = -(sum G N) / TT
//   mode 1 (SqrDistance):    a = -2, b = 2 sum G
__global__ __launch_bounds__(kThreads) void grad_template(const float *partial, int nblocks,
                                                           const double *scal_partial,
                                                           int nscal, const float *tstage,
                                                           int tpitch, const double *tsums,
                                                           int tw, int th, int mode, char *dT,
                                                           int dtstep)
{
   __shared__ double red[4];
   double sc = 0;
   for (int k = threadIdx.x; k < nscal; k += kThreads) sc += scal_partial[k];
   sc = block_sum_d(sc, red);
   const double a = mode == 0 ? 1.0 : -2.0;
   const double b = mode == 0 ? -sc / tsums[0] : 2.0 * sc;
   // 64 taps per workgroup (lane = tap); the four waves sum the blocks k = wave (mod 4),
   // then lane's four sums are added in wave order
   __shared__ double part[4][64];
   const int ntaps = tw * th, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   for (int t0 = blockIdx.x * 64; t0 < ntaps; t0 += gridDim.x * 64) {
      const int tap = t0 + lane;
      double r = 0;
      if (tap < ntaps)
         for (int k = wave; k < nblocks; k += 4) r += partial[(size_t)k * ntaps + tap];
      part[wave][lane] = r;
      __syncthreads();
      if (wave == 0 && tap < ntaps) {
         r = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
         const int v = tap / tw, u = tap - v * tw;
         const double t = tstage[(size_t)v * tpitch + u];
         reinterpret_cast<float *>(dT + (size_t)v * dtstep)[u] = (float)(a * r + b * t);
      }
      __syncthreads();
   }
}

// flipped staged template: tf(u, v) = t(tw-1-u, th-1-v)
__global__ void flip_template(const float *tstage, int tpitch, int tw, int th, float *tflip)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < tpitch * th; i += gridDim.x * blockDim.x) {
      const int v = i / tpitch, u = i - v * tpitch;
      tflip[i] = u < tw ? tstage[(size_t)(th - 1 - v) * tpitch + (tw - 1 - u)] : 0.f;
   }
}

// dS(j, i) = alpha * F(j, i) + beta * s(j, i) * (flipped window sum of the map integrated in I)
__global__ __launch_bounds__(kThreads) void grad_source(const float *F, const float *S, int sstep,
                                                         const double *__restrict__ I, int W,
                                                         int H, int tw, int th, float alpha,
                                                         float beta, char *dS, int dsstep)
{
   const int cx = tw / 2, cy = th / 2;
   for (long long k = (long long)blockIdx.x * kThreads + threadIdx.x; k < (long long)W * H;
        k += (long long)gridDim.x * kThreads) {
      const int i = (int)(k / W), j = (int)(k - (long long)i * W);
      const double box = rect_sum(I, W, H, j + cx - tw + 1, j + cx + 1, i + cy - th + 1, i + cy + 1);
      const double s = reinterpret_cast<const float *>(reinterpret_cast<const char *>(S) +
                                                       (size_t)i * sstep)[j];
      reinterpret_cast<float *>(dS + (size_t)i * dsstep)[j] =
          (float)(alpha * (double)F[k] + beta * s * box);
   }
}

int grid_for(long long n) { return (int)std::min<long long>(1024, (n + kThreads - 1) / kThreads); }

struct BwLayout {
   size_t F, Gd, E, I, partial, scal, tflip, total;
};

size_t align_up(size_t n) { return (n + 255) / 256 * 256; }

BwLayout bw_layout(int W, int H, int tw, int th, int tpitch)
{
   const size_t px = (size_t)W * H;
   // weight_corr partials: room for 32-pixel tiles (at most 4x the 64-pixel tile count)
   const size_t nb = (size_t)((W + 31) / 32) * ((H + 31) / 32);
   BwLayout L;
   size_t o = 0;
   L.F = o;       o = align_up(o + px * sizeof(float));
   L.Gd = o;      o = align_up(o + px * sizeof(float));
   L.E = o;       o = align_up(o + px * sizeof(float));
   L.I = o;       o = align_up(o + excl_bytes(W, H));
   L.partial = o; o = align_up(o + nb * tw * th * sizeof(float));
   L.scal = o;    o = align_up(o + 1024 * sizeof(double));
   L.tflip = o;   o = align_up(o + (size_t)tpitch * th * sizeof(float));
   L.total = o;
   return L;
}

}  // namespace

size_t backward_work_bytes(int W, int H, int tw, int th, int tpitch)
{
   return bw_layout(W, H, tw, th, tpitch).total;
}

// mode 0: CrossCorr_Norm, 1: SqrDistance.  integ_ss: exclusive FP64 integral of s^2 (mode 0).
void launch_backward(hipStream_t s, int mode, const View &S, const View &G, const float *tstage,
                     int tpitch, int tw, int th, const double *tsums, const double *integ_ss,
                     void *work, void *aux, char *dS, int dsstep, char *dT, int dtstep)
{
   const int W = S.width, H = S.height;
   const size_t px = (size_t)W * H;
   const BwLayout L = bw_layout(W, H, tw, th, tpitch);
   char *w = static_cast<char *>(work);
   float *F = reinterpret_cast<float *>(w + L.F);
   float *Gd = reinterpret_cast<float *>(w + L.Gd);
   float *E = reinterpret_cast<float *>(w + L.E);
   double *I = reinterpret_cast<double *>(w + L.I);
   float *partial = reinterpret_cast<float *>(w + L.partial);
   double *scal = reinterpret_cast<double *>(w + L.scal);
   float *tflip = reinterpret_cast<float *>(w + L.tflip);
   const int tile = (64 + th - 1) * (64 + tw - 1 + 4) * 4 + 64 * 64 * 4 <= 64 * 1024 ? 64 : 32;
   const int nbx = (W + tile - 1) / tile, nby = (H + tile - 1) / tile;
   const int nscal = grid_for((long long)px);
   const float *Sp = static_cast<const float *>(S.ptr);
   KernelTimer kt(s, "backward");

   // the map M that multiplies the shifted source in dT, and that is correlated with the
   // flipped template in dS
   const char *M;
   int mstep;
   if (mode == 0) {
      // forward CC with the same order as ocipCrossCorr, then the per-pixel maps
      launch_crosscorr_f32_anchored(s, S, tstage, tpitch, tw, th, tw / 2, th / 2, F, W * 4);
      ncc_maps<<<nscal, kThreads, 0, s>>>(F, W * 4, static_cast<const char *>(G.ptr), G.step,
                                           integ_ss, tsums, W, H, tw, th, Gd, E, scal);
      launch_check("ncc_maps");
      M = reinterpret_cast<const char *>(Gd);
      mstep = W * 4;
   } else {
      map_sum<<<nscal, kThreads, 0, s>>>(static_cast<const char *>(G.ptr), G.step, W, H, scal);
      launch_check("map_sum");
      M = static_cast<const char *>(G.ptr);
      mstep = G.step;
   }

   if (dT) {
      // a multiple of 4 floats = 4 mod 32 (16-byte window loads, rows on different banks),
      // at least tile + tw + 3
      int ldsw = (tile + tw - 1 + 4 + 3) / 4 * 4;
      while (ldsw % 32 != 4) ldsw += 4;
      const size_t lds = (size_t)tile * tile * 4 + (size_t)(tile + th - 1) * ldsw * 4;
      dim3 grid(nbx, nby);
      if (tile == 64) {
         hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(weight_corr<64>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), "attr");
         weight_corr<64><<<grid, kThreads, lds, s>>>(M, mstep, Sp, S.step, W, H, tw, th, partial, ldsw);
      } else {
         if (lds > 160 * 1024) fail(CL_INVALID_VALUE, "template too large for the gradient");
         hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(weight_corr<32>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), "attr");
         weight_corr<32><<<grid, kThreads, lds, s>>>(M, mstep, Sp, S.step, W, H, tw, th, partial, ldsw);
      }
      launch_check("weight_corr");
      grad_template<<<(tw * th + 63) / 64, kThreads, 0, s>>>(
          partial, nbx * nby, scal, nscal, tstage, tpitch, tsums, tw, th, mode, dT, dtstep);
      launch_check("grad_template");
   }

   if (dS) {
      flip_template<<<grid_for((long long)tpitch * th), kThreads, 0, s>>>(tstage, tpitch, tw, th, tflip);
      launch_check("flip_template");
      View mv{M, W, H, mstep, 1, SImage::F32};
      launch_crosscorr_f32_anchored(s, mv, tflip, tpitch, tw, th, tw - 1 - tw / 2, th - 1 - th / 2,
                                    F, W * 4);
      // integral of the map whose flipped window sums multiply s
      View bv = mode == 0 ? View{E, W, H, W * 4, 1, SImage::F32} : G;
      launch_integral(s, bv, 0, false, I, 0, -1, aux);
      const float alpha = mode == 0 ? 1.f : -2.f, beta = mode == 0 ? -1.f : 2.f;
      grad_source<<<grid_for((long long)px), kThreads, 0, s>>>(F, Sp, S.step, I, W, H, tw, th,
                                                               alpha, beta, dS, dsstep);
      launch_check("grad_source");
   }
}

}  // namespace ocip

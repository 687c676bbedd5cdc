// fft.hip -- template matching through 2-D FFTs (the reference's large-template path).
//
// Reference: ImageProximityFFT::{CrossCorr, CrossCorr_Norm, SqrDistance, SqrDistance_Norm}
// (C++/programs/ImageProximityFFT.cpp:115-267) and its kernels
// (cl-files/ImageProximityFFT.cl:55-212): the source is placed at (tw/2, th/2) in a zero image
// of Wb x Hb (W + tw/2, H + th/2 rounded up to a length whose only prime factors are 2, 3
// and 5, ImageProximityFFT.cpp:48-56, FFT.cpp:85-103), the template at (0, 0) in another; the
// correlation is the inverse FFT of FFT(source) * conj(FFT(template)) / (Wb Hb)
// (mulAndScaleSpectrums), cropped to W x H; the Norm ops divide by sqrt(SS TT) with SS the
// window sum of s^2 from the square integral (AreaSqSum) and TT = SumSqr(template).  The
// circular correlation equals the zero-padded one because every wrapped index lands in the
// zero border.
//
// This build computes in FP64 complex arithmetic (the reference uses clFFT in single
// precision and an F32 integral) with its own mixed-radix (2, 3, 4, 5, 8, 9) Stockham FFT in LDS:
//   pass 1 (fft_rows_pack): one workgroup per row y of the packed image z = s + i t (source
//          and template share one complex transform), FFT along x, written transposed;
//   pass 2 (fft_cols_pair): one workgroup per pair of spectrum columns (kx, -kx): FFT along y
//          of both, the spectra of s and t separated by Hermitian symmetry
//          (S = (Z(k) + conj Z(-k)) / 2, T = (Z(k) - conj Z(-k)) / 2i), the product
//          conj(S conj T) / (Wb Hb), FFT along y again (an inverse transform of conjugated
//          data), written transposed;
//   pass 3 (fft_rows_epi): one workgroup per output row: FFT along x, real part (the
//          conjugations cancel in it), the epilogue of the op, stored as float.
#include <cmath>
#include <vector>

#include "device_common.h"

namespace ocip {

namespace {

constexpr int kFftThreads = 1024;
constexpr int kFftMaxStages = 24;
// one row of complex doubles in the 160 KiB of LDS (less room for static LDS)
constexpr int kFftMaxLen = (160 * 1024 - 256) / 16;

struct FftPlan {
   int n, nstages, twsize;
   int radix[kFftMaxStages];
   int span[kFftMaxStages];   // product of the radices of the earlier stages
   int twoff[kFftMaxStages];  // offset of the stage's twiddles exp(-2 pi i k / (span radix))
};

// q = x / d and r = x mod d for 0 <= x < 2^24, 0 < d, from a float reciprocal corrected by
// one step (the float quotient is within one of the true one)
__device__ __forceinline__ int div_small(int x, int d, float rcp, int &r)
{
   int q = (int)((float)x * rcp);
   r = x - q * d;
   if (r < 0) {
      --q;
      r += d;
   } else if (r >= d) {
      ++q;
      r -= d;
   }
   return q;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
   return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// -i z and +i z
__device__ __forceinline__ double2 mul_mi(double2 z) { return make_double2(z.y, -z.x); }
__device__ __forceinline__ double2 mul_pi(double2 z) { return make_double2(-z.y, z.x); }
__device__ __forceinline__ double2 cscale(double2 z, double s) { return make_double2(z.x * s, z.y * s); }

// Forward DFTs of length R (sign -1) in place.
template <int R> __device__ __forceinline__ void dft(double2 (&u)[R]);

template <> __device__ __forceinline__ void dft<2>(double2 (&u)[2])
{
   const double2 a = u[0], b = u[1];
   u[0] = cadd(a, b);
   u[1] = csub(a, b);
}

template <> __device__ __forceinline__ void dft<3>(double2 (&u)[3])
{
   const double s = 0.86602540378443864676;  // sin(2 pi / 3)
   const double2 a = u[0], t1 = cadd(u[1], u[2]), d = csub(u[1], u[2]);
   const double2 t2 = make_double2(a.x - 0.5 * t1.x, a.y - 0.5 * t1.y);
   const double2 t3 = cscale(mul_mi(d), s);
   u[0] = cadd(a, t1);
   u[1] = cadd(t2, t3);
   u[2] = csub(t2, t3);
}

template <> __device__ __forceinline__ void dft<4>(double2 (&u)[4])
{
   const double2 t0 = cadd(u[0], u[2]), t1 = csub(u[0], u[2]);
   const double2 t2 = cadd(u[1], u[3]), t3 = mul_mi(csub(u[1], u[3]));
   u[0] = cadd(t0, t2);
   u[2] = csub(t0, t2);
   u[1] = cadd(t1, t3);
   u[3] = csub(t1, t3);
}

template <> __device__ __forceinline__ void dft<5>(double2 (&u)[5])
{
   const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;  // cos(2pi/5), cos(4pi/5)
   const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;   // sin(2pi/5), sin(4pi/5)
   const double2 a = u[0];
   const double2 t1 = cadd(u[1], u[4]), t2 = cadd(u[2], u[3]);
   const double2 t3 = csub(u[1], u[4]), t4 = csub(u[2], u[3]);
   const double2 m1 = make_double2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
   const double2 m2 = make_double2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
   const double2 n1 = make_double2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
   const double2 n2 = make_double2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
   u[0] = make_double2(a.x + t1.x + t2.x, a.y + t1.y + t2.y);
   u[1] = cadd(m1, mul_mi(n1));
   u[4] = cadd(m1, mul_pi(n1));
   u[2] = cadd(m2, mul_mi(n2));
   u[3] = cadd(m2, mul_pi(n2));
}

// 8 points: DFT4 of the even and of the odd inputs, the odd ones times w8^k = exp(-2 pi i k/8)
template <> __device__ __forceinline__ void dft<8>(double2 (&u)[8])
{
   const double r = 0.70710678118654752440;  // 1 / sqrt(2)
   double2 e[4] = {u[0], u[2], u[4], u[6]}, o[4] = {u[1], u[3], u[5], u[7]};
   dft<4>(e);
   dft<4>(o);
   const double2 t[4] = {o[0], make_double2(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x)),
                         mul_mi(o[2]), make_double2(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y))};
#pragma unroll
   for (int k = 0; k < 4; ++k) {
      u[k] = cadd(e[k], t[k]);
      u[k + 4] = csub(e[k], t[k]);
   }
}

// 9 points: DFT3 over m of x[3m + p] for p = 0, 1, 2, twiddles w9^(p k1), DFT3 over p:
// X[k1 + 3 k2]
template <> __device__ __forceinline__ void dft<9>(double2 (&u)[9])
{
   // w9^1, w9^2, w9^4 with w9 = exp(-2 pi i / 9)
   const double2 w1 = make_double2(0.76604444311897803520, -0.64278760968653932632);
   const double2 w2 = make_double2(0.17364817766693034885, -0.98480775301220805936);
   const double2 w4 = make_double2(-0.93969262078590838405, -0.34202014332566873304);
   double2 a[3][3];
#pragma unroll
   for (int p = 0; p < 3; ++p) {
      double2 v[3] = {u[p], u[p + 3], u[p + 6]};
      dft<3>(v);
      a[p][0] = v[0];
      a[p][1] = v[1];
      a[p][2] = v[2];
   }
   a[1][1] = cmul(a[1][1], w1);
   a[1][2] = cmul(a[1][2], w2);
   a[2][1] = cmul(a[2][1], w2);
   a[2][2] = cmul(a[2][2], w4);
#pragma unroll
   for (int k1 = 0; k1 < 3; ++k1) {
      double2 v[3] = {a[0][k1], a[1][k1], a[2][k1]};
      dft<3>(v);
      u[k1] = v[0];
      u[k1 + 3] = v[1];
      u[k1 + 6] = v[2];
   }
}

// Workgroup barrier that orders LDS only (a full __syncthreads also waits for the
// workgroup's outstanding global loads and stores, which the FFT stages do not need).
__device__ __forceinline__ void lds_barrier()
{
   __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
   __builtin_amdgcn_s_barrier();
   __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One Stockham stage of radix R over nseq sequences of length n stored back to back in LDS:
// butterfly b (k = b mod span) takes x[b + j n/R], multiplies input j by w^(j k),
// w = exp(-2 pi i / (span R)), and writes DFT output j to x[(b - k) R + k + j span].  All
// inputs are read into registers before any output is written (one buffer).
template <int R>
__device__ __forceinline__ void fft_stage(double2 *x, int n, int nseq, int span,
                                          const double2 *__restrict__ tw)
{
   constexpr int MAXB = (kFftMaxLen / R + kFftThreads - 1) / kFftThreads;
   const int m = n / R, total = m * nseq;
   const float rm = 1.0f / (float)m, rs = 1.0f / (float)span;
   double2 u[MAXB][R];
#pragma unroll
   for (int q = 0; q < MAXB; ++q) {
      const int g = threadIdx.x + q * kFftThreads;
      if (g < total) {
         int b, k;
         const int seq = nseq == 1 ? (b = g, 0) : div_small(g, m, rm, b);
         div_small(b, span, rs, k);
         const double2 *xs = x + seq * n;
#pragma unroll
         for (int j = 0; j < R; ++j) u[q][j] = xs[b + j * m];
         if (k) {
            const double2 w = tw[k];
            double2 wj = w;
#pragma unroll
            for (int j = 1; j < R; ++j) {
               u[q][j] = cmul(u[q][j], wj);
               if (j + 1 < R) wj = cmul(wj, w);
            }
         }
         dft<R>(u[q]);
      }
   }
   lds_barrier();
#pragma unroll
   for (int q = 0; q < MAXB; ++q) {
      const int g = threadIdx.x + q * kFftThreads;
      if (g < total) {
         int b, k;
         const int seq = nseq == 1 ? (b = g, 0) : div_small(g, m, rm, b);
         div_small(b, span, rs, k);
         double2 *xs = x + seq * n + (b - k) * R + k;
#pragma unroll
         for (int j = 0; j < R; ++j) xs[j * span] = u[q][j];
      }
   }
   lds_barrier();
}

// Forward FFT (unnormalised) of nseq sequences of length P.n in LDS, natural order in and out;
// tw: the plan's twiddles, in LDS.
// LARGE: the plan may hold radix-8 and radix-9 stages (row plans only: the column kernels,
// which hold two columns, have no registers to spare for them)
template <bool LARGE = false>
__device__ __forceinline__ void lds_fft(double2 *x, int nseq, const FftPlan *__restrict__ P, const double2 *tw)
{
   const int n = P->n, ns = P->nstages;
   for (int s = 0; s < ns; ++s) {
      const double2 *t = tw + P->twoff[s];
      const int span = P->span[s];
      switch (P->radix[s]) {
      case 2: fft_stage<2>(x, n, nseq, span, t); break;
      case 3: fft_stage<3>(x, n, nseq, span, t); break;
      case 4: fft_stage<4>(x, n, nseq, span, t); break;
      case 8:
         if constexpr (LARGE) fft_stage<8>(x, n, nseq, span, t);
         break;
      case 9:
         if constexpr (LARGE) fft_stage<9>(x, n, nseq, span, t);
         break;
      default: fft_stage<5>(x, n, nseq, span, t); break;
      }
   }
}


// One workgroup per row (passes 1 and 3) or column pair (pass 2).  Passes 1 and 2 take their
// rows in XCD-aware order (xcd_tile): the workgroups of one XCD take consecutive rows, so the
// transposed 16-byte stores of neighbouring rows fill the same lines of that XCD's L2.

// Pass 1: row y of z = s + i t (s at (tw/2, th/2), t at (0, 0), zero elsewhere), FFT along
// x, stored transposed: out[kx * Hb + y].
__global__ __launch_bounds__(kFftThreads) void fft_rows_pack(const char *__restrict__ S, int W,
                                                             int H, int sstep,
                                                             const char *__restrict__ T, int tw,
                                                             int th, int tstep, int Hb,
                                                             const FftPlan *__restrict__ pw,
                                                             const double2 *__restrict__ tww,
                                                             double2 *__restrict__ out)
{
   extern __shared__ double2 fbuf[];
   const int Wb = pw->n, cx = tw / 2;
   const int y = xcd_tile(blockIdx.x, gridDim.x);
   double2 *ltw = fbuf + Wb;
   for (int i = threadIdx.x; i < pw->twsize; i += kFftThreads) ltw[i] = tww[i];
   const int sy = y - th / 2;
   const bool srow = sy >= 0 && sy < H, trow = y < th;
   const float *sr = reinterpret_cast<const float *>(S + (size_t)(srow ? sy : 0) * sstep);
   const float *tr = reinterpret_cast<const float *>(T + (size_t)(trow ? y : 0) * tstep);
   for (int x = threadIdx.x; x < Wb; x += kFftThreads) {
      const int sx = x - cx;
      const double sv = (srow && sx >= 0 && sx < W) ? (double)sr[sx] : 0.0;
      const double tv = (trow && x < tw) ? (double)tr[x] : 0.0;
      fbuf[x] = make_double2(sv, tv);
   }
   lds_barrier();
   lds_fft<true>(fbuf, 1, pw, ltw);
   for (int x = threadIdx.x; x < Wb; x += kFftThreads) out[(size_t)x * Hb + y] = fbuf[x];
}

// Spectrum product at k from Z(k) = z1 and Z(-k) = z2, conjugated for the inverse transform.
__device__ __forceinline__ double2 spectrum_product(double2 z1, double2 z2, double scale)
{
   const double2 s = cscale(cadd(z1, cconj(z2)), 0.5);           // FFT of the source
   const double2 t = cscale(mul_mi(csub(z1, cconj(z2))), 0.5);   // FFT of the template
   return cconj(cscale(cmul(s, cconj(t)), scale));
}

// Pass 2: spectrum columns kx and Wb - kx (rows of the transposed array), kx = 0..Wb/2: FFT
// along y, product, FFT along y again; stored transposed (rows y < H only): out[y * Wb + kx].
__global__ __launch_bounds__(kFftThreads) void fft_cols_pair(const double2 *__restrict__ in, int Wb,
                                                             int H, const FftPlan *__restrict__ ph,
                                                             const double2 *__restrict__ twh,
                                                             double scale,
                                                             double2 *__restrict__ out)
{
   extern __shared__ double2 fbuf[];
   const int Hb = ph->n;
   double2 *ltw = fbuf + 2 * Hb;
   for (int i = threadIdx.x; i < ph->twsize; i += kFftThreads) ltw[i] = twh[i];
   {
      const int kx = xcd_tile(blockIdx.x, gridDim.x);
      const int kx2 = (Wb - kx) % Wb;
      const int nseq = kx2 == kx ? 1 : 2;
      double2 *A = fbuf, *B = nseq == 2 ? fbuf + Hb : fbuf;
      for (int y = threadIdx.x; y < Hb; y += kFftThreads) {
         A[y] = in[(size_t)kx * Hb + y];
         if (nseq == 2) B[y] = in[(size_t)kx2 * Hb + y];
      }
      lds_barrier();
      lds_fft(fbuf, nseq, ph, ltw);
      // the four entries of a pair (ky, -ky) of both columns are read and written by one thread
      for (int ky = threadIdx.x; ky <= Hb / 2; ky += kFftThreads) {
         const int ky2 = ky == 0 ? 0 : Hb - ky;
         const double2 a1 = A[ky], a2 = A[ky2], b1 = B[ky], b2 = B[ky2];
         A[ky] = spectrum_product(a1, b2, scale);
         A[ky2] = spectrum_product(a2, b1, scale);
         if (nseq == 2) {
            B[ky] = spectrum_product(b1, a2, scale);
            B[ky2] = spectrum_product(b2, a1, scale);
         }
      }
      lds_barrier();
      lds_fft(fbuf, nseq, ph, ltw);
      for (int y = threadIdx.x; y < H; y += kFftThreads) {
         out[(size_t)y * Wb + kx] = A[y];
         if (nseq == 2) out[(size_t)y * Wb + kx2] = B[y];
      }
   }
}

// Pass 3: output row y < H: FFT along x, real part = the correlation, the op's epilogue
// (ImageProximityFFT.cl:86-147, in FP64 with SS from the FP64 integral).
__global__ __launch_bounds__(kFftThreads) void fft_rows_epi(const double2 *__restrict__ in, int W,
                                                            int H, const FftPlan *__restrict__ pw,
                                                            const double2 *__restrict__ tww,
                                                            int op, int tw, int th,
                                                            const double *__restrict__ integ,
                                                            const double *__restrict__ tsums,
                                                            char *__restrict__ dst, int dstep)
{
   extern __shared__ double2 fbuf[];
   const int Wb = pw->n;
   double2 *ltw = fbuf + Wb;
   for (int i = threadIdx.x; i < pw->twsize; i += kFftThreads) ltw[i] = tww[i];
   const double tt = op == kCC ? 0.0 : tsums[0];
   {
      const int y = blockIdx.x;
      for (int x = threadIdx.x; x < Wb; x += kFftThreads) fbuf[x] = in[(size_t)y * Wb + x];
      lds_barrier();
      lds_fft<true>(fbuf, 1, pw, ltw);
      float *drow = reinterpret_cast<float *>(dst + (size_t)y * dstep);
      for (int x = threadIdx.x; x < W; x += kFftThreads) {
         const double cc = fbuf[x].x;
         double v = cc;
         if (op != kCC) {
            const double ss = window_sqsum(integ, W, H, x, y, tw, th);
            if (op == kCCNorm) v = cc / sqrt(ss * tt);
            else if (op == kSSD) v = ss - 2.0 * cc + tt;
            else v = (ss - 2.0 * cc + tt) / sqrt(ss * tt);
         }
         drow[x] = (float)v;
      }
   }
}

// ---- the generic 2-D FFT (reference FFT::Forward / Inverse, C++/programs/FFT.cpp:140-283):
// real W x H image <-> Hermitian half spectrum (W/2 + 1) x H, complex values interleaved
// (re, im) in F32 or F64, unnormalised both ways (clfftSetPlanScale(..., 1)).  Computed in FP64;
// work holds (W/2 + 1) x H complex doubles, one row per spectrum column kx.

// Forward, pass 1: row y of the real image, FFT along x, columns kx <= W/2 stored transposed.
template <class T>
__global__ __launch_bounds__(kFftThreads) void fft_r2c_rows(const char *__restrict__ src, int sstep,
                                                            int H, const FftPlan *__restrict__ pw,
                                                            const double2 *__restrict__ tww,
                                                            double2 *__restrict__ work)
{
   extern __shared__ double2 fbuf[];
   const int W = pw->n, y = xcd_tile(blockIdx.x, gridDim.x);
   double2 *ltw = fbuf + W;
   for (int i = threadIdx.x; i < pw->twsize; i += kFftThreads) ltw[i] = tww[i];
   const T *row = reinterpret_cast<const T *>(src + (size_t)y * sstep);
   for (int x = threadIdx.x; x < W; x += kFftThreads) fbuf[x] = make_double2((double)row[x], 0.0);
   lds_barrier();
   lds_fft<true>(fbuf, 1, pw, ltw);
   for (int kx = threadIdx.x; kx <= W / 2; kx += kFftThreads) work[(size_t)kx * H + y] = fbuf[kx];
}

// Forward, pass 2: spectrum column kx, FFT along y, stored as row ky, column kx of dst.
template <class T>
__global__ __launch_bounds__(kFftThreads) void fft_r2c_cols(const double2 *__restrict__ work,
                                                            const FftPlan *__restrict__ ph,
                                                            const double2 *__restrict__ twh,
                                                            char *__restrict__ dst, int dstep)
{
   extern __shared__ double2 fbuf[];
   const int H = ph->n, kx = xcd_tile(blockIdx.x, gridDim.x);
   double2 *ltw = fbuf + H;
   for (int i = threadIdx.x; i < ph->twsize; i += kFftThreads) ltw[i] = twh[i];
   for (int y = threadIdx.x; y < H; y += kFftThreads) fbuf[y] = work[(size_t)kx * H + y];
   lds_barrier();
   lds_fft(fbuf, 1, ph, ltw);
   for (int ky = threadIdx.x; ky < H; ky += kFftThreads) {
      T *d = reinterpret_cast<T *>(dst + (size_t)ky * dstep) + 2 * kx;
      d[0] = (T)fbuf[ky].x;
      d[1] = (T)fbuf[ky].y;
   }
}

// Inverse, pass 1: column kx of the half spectrum, inverse FFT along y as the FFT of the
// conjugate; the conjugated result stored as row kx of work.
template <class T>
__global__ __launch_bounds__(kFftThreads) void fft_c2r_cols(const char *__restrict__ src, int sstep,
                                                            const FftPlan *__restrict__ ph,
                                                            const double2 *__restrict__ twh,
                                                            double2 *__restrict__ work)
{
   extern __shared__ double2 fbuf[];
   const int H = ph->n, kx = xcd_tile(blockIdx.x, gridDim.x);
   double2 *ltw = fbuf + H;
   for (int i = threadIdx.x; i < ph->twsize; i += kFftThreads) ltw[i] = twh[i];
   for (int ky = threadIdx.x; ky < H; ky += kFftThreads) {
      const T *v = reinterpret_cast<const T *>(src + (size_t)ky * sstep) + 2 * kx;
      fbuf[ky] = make_double2((double)v[0], -(double)v[1]);
   }
   lds_barrier();
   lds_fft(fbuf, 1, ph, ltw);
   for (int y = threadIdx.x; y < H; y += kFftThreads) work[(size_t)kx * H + y] = fbuf[y];
}

// Inverse, pass 2: row y: the conjugated row spectrum completed by Hermitian symmetry
// (Z(W - kx) = conj Z(kx)), FFT along x, real part stored.
template <class T>
__global__ __launch_bounds__(kFftThreads) void fft_c2r_rows(const double2 *__restrict__ work, int H,
                                                            const FftPlan *__restrict__ pw,
                                                            const double2 *__restrict__ tww,
                                                            char *__restrict__ dst, int dstep)
{
   extern __shared__ double2 fbuf[];
   const int W = pw->n, y = blockIdx.x;
   double2 *ltw = fbuf + W;
   for (int i = threadIdx.x; i < pw->twsize; i += kFftThreads) ltw[i] = tww[i];
   for (int kx = threadIdx.x; kx < W; kx += kFftThreads)
      fbuf[kx] = kx <= W / 2 ? work[(size_t)kx * H + y] : cconj(work[(size_t)(W - kx) * H + y]);
   lds_barrier();
   lds_fft<true>(fbuf, 1, pw, ltw);
   T *row = reinterpret_cast<T *>(dst + (size_t)y * dstep);
   for (int x = threadIdx.x; x < W; x += kFftThreads) row[x] = (T)fbuf[x].x;
}


// Stages for a 2^a 3^b 5^c length, larger radices last (the twiddle table of a stage has span
// entries, so this keeps the table small): LARGE (row plans), radix 8 and 9 where they fit
// (fewer passes through LDS) and 4 / 2 and 3 for what is left; otherwise radices 2 to 5.
// Twiddles appended to `table`, offsets relative to the plan's first entry.
void make_plan(int n, FftPlan &p, std::vector<double2> &table, bool large)
{
   p.n = n;
   p.nstages = 0;
   const int base = (int)table.size();
   int rest = n, span = 1;
   auto add = [&](int r) {
      if (p.nstages >= kFftMaxStages) fail(CL_INVALID_IMAGE_SIZE, "FFT length");
      p.radix[p.nstages] = r;
      p.span[p.nstages] = span;
      p.twoff[p.nstages] = (int)table.size() - base;
      for (int k = 0; k < span; ++k) {
         const long double a = -2.0L * 3.14159265358979323846264338327950288L * k / ((long double)span * r);
         table.push_back(make_double2((double)cosl(a), (double)sinl(a)));
      }
      span *= r;
      rest /= r;
      ++p.nstages;
   };
   int twos = 0, m = n;
   while (m % 2 == 0) {
      m /= 2;
      ++twos;
   }
   if (!large) {
      if (twos % 2) add(2);
      while (rest % 3 == 0) add(3);
      while (rest % 4 == 0) add(4);
      while (rest % 5 == 0) add(5);
      if (rest != 1) fail(CL_INVALID_IMAGE_SIZE, "FFT length with a prime factor above 5");
      p.twsize = (int)table.size() - base;
      return;
   }
   int threes = 0;
   for (int q = m; q % 3 == 0; q /= 3) ++threes;
   // radix 8 and 9 where they fit (fewer LDS passes), 4 / 2 and 3 for the rest
   if (twos % 3 == 1 && twos >= 4) {
      add(4);
      add(4);
      twos -= 4;
   } else if (twos % 3 == 1) {
      add(2);
      twos -= 1;
   } else if (twos % 3 == 2) {
      add(4);
      twos -= 2;
   }
   const int nines = threes / 2;
   for (int k = threes - 2 * nines; k > 0; --k) add(3);
   while (rest % 5 == 0) add(5);
   for (; twos > 0; twos -= 3) add(8);
   for (int k = nines; k > 0; --k) add(9);
   if (rest != 1) fail(CL_INVALID_IMAGE_SIZE, "FFT length with a prime factor above 5");
   p.twsize = (int)table.size() - base;
}

// LDS of the three passes: data plus the plan's twiddles, in complex doubles
size_t row_lds(const FftPlan &p) { return (size_t)(p.n + p.twsize) * sizeof(double2); }
size_t pair_lds(const FftPlan &p) { return (size_t)(2 * p.n + p.twsize) * sizeof(double2); }

}  // namespace

// FFT.cpp:85-103 IsSupportedLength (multiples of 5 accepted: not an NVIDIA platform)
bool fft_supported_length(int n)
{
   if (n < 1) return false;
   for (int f : {2, 3, 5})
      while (n % f == 0) n /= f;
   return n == 1;
}

// ImageProximityFFT.cpp:48-56
int fft_padded_length(int n)
{
   while (!fft_supported_length(n)) ++n;
   return n;
}

struct FftState {
   int Wb = 0, Hb = 0, hbase = 0;   // hbase: first twiddle of the height plan
   FftPlan pw{}, ph{};
   DevBuf twid, plans, a, b;   // plans: pw and ph, read by the kernels
   ~FftState()
   {
      twid.release();
      plans.release();
      a.release();
      b.release();
   }
};

FftState *fft_state_create() { return new FftState(); }
void fft_state_destroy(FftState *s) { delete s; }

cl_int fft_sizes(int W, int H, int tw, int th, int *Wb, int *Hb)
{
   *Wb = fft_padded_length(W + tw / 2);
   *Hb = fft_padded_length(H + th / 2);
   // The reference stops at 2^24 padded points (a clFFT limit, FFT.cpp:170-171); here the
   // limit is LDS: one row of the padded width, or two columns of the padded height, of
   // complex doubles plus the plan's twiddles per workgroup.
   if (*Wb > kFftMaxLen || 2 * *Hb > kFftMaxLen) return CL_INVALID_IMAGE_SIZE;
   std::vector<double2> table;
   FftPlan pw, ph;
   make_plan(*Wb, pw, table, true);
   make_plan(*Hb, ph, table, false);
   if (row_lds(pw) > (size_t)kFftMaxLen * sizeof(double2) ||
       pair_lds(ph) > (size_t)kFftMaxLen * sizeof(double2))
      return CL_INVALID_IMAGE_SIZE;
   return CL_SUCCESS;
}

// Sizes the state's plans and buffers for a W x H source and tw x th template.
void fft_prepare(hipStream_t s, FftState *st, int W, int H, int tw, int th)
{
   int Wb, Hb;
   if (fft_sizes(W, H, tw, th, &Wb, &Hb) != CL_SUCCESS) fail(CL_INVALID_IMAGE_SIZE, "image too big for the FFT path");
   if (st->Wb == Wb && st->Hb == Hb) return;
   std::vector<double2> table;  // the width plan's twiddles, then the height plan's
   make_plan(Wb, st->pw, table, true);
   st->hbase = (int)table.size();
   make_plan(Hb, st->ph, table, false);
   st->twid.ensure(table.size() * sizeof(double2));
   hip_check(hipMemcpyAsync(st->twid.ptr, table.data(), table.size() * sizeof(double2),
                            hipMemcpyHostToDevice, s), "twiddles");
   const FftPlan both[2] = {st->pw, st->ph};
   st->plans.ensure(sizeof(both));
   hip_check(hipMemcpyAsync(st->plans.ptr, both, sizeof(both), hipMemcpyHostToDevice, s), "plans");
   hip_check(hipStreamSynchronize(s), "twiddles");
   st->a.ensure((size_t)Wb * Hb * sizeof(double2));
   st->b.ensure((size_t)Wb * Hb * sizeof(double2));
   st->Wb = Wb;
   st->Hb = Hb;
   static bool lds_set = false;
   if (!lds_set) {
      for (const void *k : {reinterpret_cast<const void *>(fft_rows_pack),
                            reinterpret_cast<const void *>(fft_cols_pair),
                            reinterpret_cast<const void *>(fft_rows_epi)})
         hip_check(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       kFftMaxLen * (int)sizeof(double2)),
                   "lds attribute");
      lds_set = true;
   }
}

void launch_proximity_fft(hipStream_t s, FftState *st, int op, const View &S, const View &T,
                          const double *integ, const double *tsums, void *dst, int dstep)
{
   fft_prepare(s, st, S.width, S.height, T.width, T.height);
   const int Wb = st->Wb, Hb = st->Hb;
   const double2 *tww = static_cast<const double2 *>(st->twid.ptr), *twh = tww + st->hbase;
   double2 *a = static_cast<double2 *>(st->a.ptr), *b = static_cast<double2 *>(st->b.ptr);
   const FftPlan *plans = static_cast<const FftPlan *>(st->plans.ptr);
   {
      KernelTimer kt(s, "fft_rows_pack");
      fft_rows_pack<<<Hb, kFftThreads, row_lds(st->pw), s>>>(
          static_cast<const char *>(S.ptr), S.width, S.height, S.step,
          static_cast<const char *>(T.ptr), T.width, T.height, T.step, Hb, plans, tww, a);
      launch_check("fft_rows_pack");
   }
   {
      KernelTimer kt(s, "fft_cols_pair");
      fft_cols_pair<<<Wb / 2 + 1, kFftThreads, pair_lds(st->ph), s>>>(
          a, Wb, S.height, plans + 1, twh, 1.0 / ((double)Wb * Hb), b);
      launch_check("fft_cols_pair");
   }
   {
      KernelTimer kt(s, "fft_rows_epi");
      fft_rows_epi<<<S.height, kFftThreads, row_lds(st->pw), s>>>(
          b, S.width, S.height, plans, tww, op, T.width, T.height, integ, tsums,
          static_cast<char *>(dst), dstep);
      launch_check("fft_rows_epi");
   }
}

// ---- generic FFT entry points -------------------------------------------------------------

cl_int fft_plain_sizes(int W, int H)
{
   if (!fft_supported_length(W) || !fft_supported_length(H)) return CL_INVALID_IMAGE_SIZE;
   std::vector<double2> table;
   FftPlan pw, ph;
   make_plan(W, pw, table, true);
   make_plan(H, ph, table, false);
   if (row_lds(pw) > (size_t)kFftMaxLen * sizeof(double2) ||
       row_lds(ph) > (size_t)kFftMaxLen * sizeof(double2))
      return CL_INVALID_IMAGE_SIZE;
   return CL_SUCCESS;
}

// Plans for exactly W x H (no padding) and a (W/2 + 1) x H work buffer.
static void fft_prepare_plain(hipStream_t s, FftState *st, int W, int H)
{
   if (fft_plain_sizes(W, H) != CL_SUCCESS) fail(CL_INVALID_IMAGE_SIZE, "FFT size");
   if (st->Wb == W && st->Hb == H) return;
   std::vector<double2> table;
   make_plan(W, st->pw, table, true);
   st->hbase = (int)table.size();
   make_plan(H, st->ph, table, false);
   st->twid.ensure(table.size() * sizeof(double2));
   hip_check(hipMemcpyAsync(st->twid.ptr, table.data(), table.size() * sizeof(double2),
                            hipMemcpyHostToDevice, s), "twiddles");
   const FftPlan both[2] = {st->pw, st->ph};
   st->plans.ensure(sizeof(both));
   hip_check(hipMemcpyAsync(st->plans.ptr, both, sizeof(both), hipMemcpyHostToDevice, s), "plans");
   hip_check(hipStreamSynchronize(s), "twiddles");
   st->a.ensure((size_t)(W / 2 + 1) * H * sizeof(double2));
   st->Wb = W;
   st->Hb = H;
   static bool lds_set = false;
   if (!lds_set) {
      for (const void *k : {reinterpret_cast<const void *>(fft_r2c_rows<float>),
                            reinterpret_cast<const void *>(fft_r2c_rows<double>),
                            reinterpret_cast<const void *>(fft_r2c_cols<float>),
                            reinterpret_cast<const void *>(fft_r2c_cols<double>),
                            reinterpret_cast<const void *>(fft_c2r_cols<float>),
                            reinterpret_cast<const void *>(fft_c2r_cols<double>),
                            reinterpret_cast<const void *>(fft_c2r_rows<float>),
                            reinterpret_cast<const void *>(fft_c2r_rows<double>)})
         hip_check(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       kFftMaxLen * (int)sizeof(double2)),
                   "lds attribute");
      lds_set = true;
   }
}

void fft_plain_prepare(hipStream_t s, FftState *st, int W, int H) { fft_prepare_plain(s, st, W, H); }

template <class T>
static void fft_forward_t(hipStream_t s, FftState *st, const View &real, const View &cplx)
{
   const int W = real.width, H = real.height;
   const double2 *tww = static_cast<const double2 *>(st->twid.ptr), *twh = tww + st->hbase;
   const FftPlan *plans = static_cast<const FftPlan *>(st->plans.ptr);
   double2 *work = static_cast<double2 *>(st->a.ptr);
   KernelTimer kt(s, "fft_r2c");
   fft_r2c_rows<T><<<H, kFftThreads, row_lds(st->pw), s>>>(static_cast<const char *>(real.ptr),
                                                           real.step, H, plans, tww, work);
   launch_check("fft_r2c_rows");
   fft_r2c_cols<T><<<W / 2 + 1, kFftThreads, row_lds(st->ph), s>>>(
       work, plans + 1, twh, const_cast<char *>(static_cast<const char *>(cplx.ptr)), cplx.step);
   launch_check("fft_r2c_cols");
}

template <class T>
static void fft_inverse_t(hipStream_t s, FftState *st, const View &cplx, const View &real)
{
   const int W = real.width, H = real.height;
   const double2 *tww = static_cast<const double2 *>(st->twid.ptr), *twh = tww + st->hbase;
   const FftPlan *plans = static_cast<const FftPlan *>(st->plans.ptr);
   double2 *work = static_cast<double2 *>(st->a.ptr);
   KernelTimer kt(s, "fft_c2r");
   fft_c2r_cols<T><<<W / 2 + 1, kFftThreads, row_lds(st->ph), s>>>(
       static_cast<const char *>(cplx.ptr), cplx.step, plans + 1, twh, work);
   launch_check("fft_c2r_cols");
   fft_c2r_rows<T><<<H, kFftThreads, row_lds(st->pw), s>>>(
       work, H, plans, tww, const_cast<char *>(static_cast<const char *>(real.ptr)), real.step);
   launch_check("fft_c2r_rows");
}

void launch_fft(hipStream_t s, FftState *st, bool forward, const View &real, const View &cplx)
{
   fft_prepare_plain(s, st, real.width, real.height);
   const bool f64 = real.type == SImage::F64;
   if (forward) {
      if (f64) fft_forward_t<double>(s, st, real, cplx);
      else fft_forward_t<float>(s, st, real, cplx);
   } else {
      if (f64) fft_inverse_t<double>(s, st, cplx, real);
      else fft_inverse_t<float>(s, st, cplx, real);
   }
}

}  // namespace ocip

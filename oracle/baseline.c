/* baseline.c -- CPU baseline of the reference's CrossCorr_Norm, for timing only.
 *
 * TEST INFRASTRUCTURE: the benchmark's cpu_baseline leg times this on the host's cores beside
 * the GPU figures.  It is not the checker (oracle.c is) and nothing in the product library
 * links or calls it.
 *
 * Each output is the reference's sum in the reference's order (cl-files/ImageProximity.cl:
 * 375-412, CrossCorr_Norm: template row outer, column inner; one FMA chain each for sum, SS
 * and TT in REAL = float; result sum / sqrt(SS * TT)), run the way a CPU OpenCL runtime runs
 * the kernel: adjacent work-items (outputs of one row) in the lanes of SIMD registers, rows
 * across threads.  Every lane keeps its own TT chain, as every work-item does.  Zero padding
 * of the source replaces the kernel's per-tap bounds test (a zero pixel adds exactly 0 to
 * every chain, so the sums are those of the kernel).  Compiled for AVX-512 and AVX2 hosts
 * (the loader picks the clone the CPU supports).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define XB 64   /* outputs per SIMD block: 4 AVX-512 or 8 AVX2 registers per chain */

#ifdef BASE_ISA
/* The row kernel, compiled once per instruction set (the Makefile builds it with AVX-512,
 * with AVX2 + FMA and without either, as rows_cc_norm_<ISA>). */
#define BASE_CAT2(a, b) a##_##b
#define BASE_CAT(a, b) BASE_CAT2(a, b)
void BASE_CAT(rows_cc_norm, BASE_ISA)(const float *pad, int Wp, const float *tpl, int tw, int th,
                                      int tstep, float *dst, int dstep, int W, int y0, int y1)
{
   #pragma omp parallel for schedule(dynamic, 1)
   for (int y = y0; y < y1; ++y) {
      float *drow = (float *)((char *)dst + (size_t)y * dstep);
      for (int x0 = 0; x0 < W; x0 += XB) {
         float acc[XB], ss[XB], tt[XB];
         for (int i = 0; i < XB; ++i) acc[i] = ss[i] = tt[i] = 0.0f;
         for (int v = 0; v < th; ++v) {
            /* padded row of source row y + v - th/2, from source column x0 - tw/2 */
            const float *prow = pad + (size_t)(y - y0 + v) * Wp + x0;
            const float *trow = (const float *)((const char *)tpl + (size_t)v * tstep);
            for (int u = 0; u < tw; ++u) {
               const float t = trow[u];
               const float *p = prow + u;
               #pragma omp simd
               for (int i = 0; i < XB; ++i) {
                  const float s = p[i];
                  acc[i] = fmaf(s, t, acc[i]);
                  ss[i] = fmaf(s, s, ss[i]);
                  tt[i] = fmaf(t, t, tt[i]);
               }
            }
         }
         const int n = W - x0 < XB ? W - x0 : XB;
         for (int i = 0; i < n; ++i) drow[x0 + i] = acc[i] / sqrtf(ss[i] * tt[i]);
      }
   }
}

#else

typedef void rows_fn(const float *, int, const float *, int, int, int, float *, int, int, int,
                     int);
rows_fn rows_cc_norm_avx512, rows_cc_norm_avx2, rows_cc_norm_generic;

/* the instruction set the row kernel runs with on this CPU */
const char *base_isa(void)
{
   __builtin_cpu_init();
   if (__builtin_cpu_supports("avx512f")) return "avx512";
   if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return "avx2";
   return "generic";
}

/* CrossCorr_Norm of output rows [y0, y1) of a 1-channel F32 image (dst: W x H floats with
 * row pitch dstep bytes).  Returns 0, or -1 when out of memory. */
int base_cc_norm(const float *src, int W, int H, int sstep, const float *tpl, int tw, int th,
                 int tstep, float *dst, int dstep, int y0, int y1, int nthreads)
{
   if (y0 < 0) y0 = 0;
   if (y1 > H) y1 = H;
   if (y1 <= y0) return 0;
#ifdef _OPENMP
   if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
   const int cx = tw / 2, cy = th / 2;
   /* padded rows y0 - cy .. y1 - cy + th - 2, columns -cx .. (blocks of XB outputs) + tw - 1 */
   const int nb = (W + XB - 1) / XB;
   const int Wp = nb * XB + tw;
   const int rows = y1 - y0 + th - 1;
   float *pad = (float *)aligned_alloc(64, ((size_t)rows * Wp * sizeof(float) + 63) / 64 * 64);
   if (!pad) return -1;
   #pragma omp parallel for schedule(static)
   for (int r = 0; r < rows; ++r) {
      float *prow = pad + (size_t)r * Wp;
      memset(prow, 0, (size_t)Wp * sizeof(float));
      const int sy = y0 - cy + r;
      if (sy < 0 || sy >= H) continue;
      const float *srow = (const float *)((const char *)src + (size_t)sy * sstep);
      memcpy(prow + cx, srow, (size_t)W * sizeof(float));
   }
   const char *isa = base_isa();
   rows_fn *fn = isa[0] == 'a' ? (isa[3] == '5' ? rows_cc_norm_avx512 : rows_cc_norm_avx2)
                               : rows_cc_norm_generic;
   fn(pad, Wp, tpl, tw, th, tstep, dst, dstep, W, y0, y1);
   free(pad);
   return 0;
}

#endif

/* oracle.c -- CPU restatement of the reference's template-matching path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the tests, the smoke
 * check and the CPU-baseline leg of the benchmark call it to produce expected
 * values.  Nothing in the product library links or calls it.
 *
 * Contents (each function cites the reference code it restates):
 *   orc_prox_ref      reference-order proximity in REAL precision
 *                     (cl-files/ImageProximity.cl:44-466, cl-files/Buffers.h:58-84);
 *                     orc_prox_ref_rows: the same on a list of output rows
 *   orc_prox_f64      same sums accumulated in FP64 (ground truth for float sources)
 *   orc_prox_i64      exact int64 sums for integer sources
 *   orc_integral_ref  reference summation tree of the 5-launch integral scan
 *                     (cl-files/Integral.cl:39-460, C++/programs/Integral.cpp:143-241)
 *   orc_integral_exact  exact inclusive 2-D prefix (int128 / long double)
 *   orc_area_sum      4-corner window sum over an inclusive integral
 *                     (cl-files/ImageProximityFFT.cl:55-84)
 *   orc_min_max_indx  min/max value and its pixel in the reference's reduction order
 *                     (cl-files/Statistics.cl:332-466, StatisticsHelpers.cpp:105-145)
 *
 * Pinning (what checks these functions against something other than themselves)
 * lives in tests/test_oracle.py: scipy.signal.correlate2d, numpy sliding windows,
 * numpy cumsum, exact-integer agreement, and closed-form invariants.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* SImage::EDataType values (include/SImage.h:36-46 of the reference). */
enum { T_U8 = 0, T_S8, T_U16, T_S16, T_U32, T_S32, T_F32, T_F64 };
/* proximity operations */
enum { OP_SSD = 0, OP_SSD_NORM = 1, OP_SAD = 2, OP_CC = 3, OP_CC_NORM = 4 };

static int type_size(int t)
{
   switch (t) {
   case T_U8: case T_S8: return 1;
   case T_U16: case T_S16: return 2;
   case T_U32: case T_S32: case T_F32: return 4;
   case T_F64: return 8;
   }
   return 0;
}

/* element (x, c) of row `row` as double (exact for every type) */
static inline double elem_d(const unsigned char *row, int t, int i)
{
   switch (t) {
   case T_U8:  return ((const uint8_t  *)row)[i];
   case T_S8:  return ((const int8_t   *)row)[i];
   case T_U16: return ((const uint16_t *)row)[i];
   case T_S16: return ((const int16_t  *)row)[i];
   case T_U32: return ((const uint32_t *)row)[i];
   case T_S32: return ((const int32_t  *)row)[i];
   case T_F32: return ((const float    *)row)[i];
   case T_F64: return ((const double   *)row)[i];
   }
   return 0.0;
}

static inline int64_t elem_i(const unsigned char *row, int t, int i)
{
   switch (t) {
   case T_U8:  return ((const uint8_t  *)row)[i];
   case T_S8:  return ((const int8_t   *)row)[i];
   case T_U16: return ((const uint16_t *)row)[i];
   case T_S16: return ((const int16_t  *)row)[i];
   case T_U32: return ((const uint32_t *)row)[i];
   case T_S32: return ((const int32_t  *)row)[i];
   }
   return 0;
}

/* convert_float of an element: C conversion to float is round-to-nearest-even,
 * as OpenCL's convert_float (default rounding) is. */
static inline float elem_f(const unsigned char *row, int t, int i)
{
   switch (t) {
   case T_U8:  return (float)((const uint8_t  *)row)[i];
   case T_S8:  return (float)((const int8_t   *)row)[i];
   case T_U16: return (float)((const uint16_t *)row)[i];
   case T_S16: return (float)((const int16_t  *)row)[i];
   case T_U32: return (float)((const uint32_t *)row)[i];
   case T_S32: return (float)((const int32_t  *)row)[i];
   case T_F32: return ((const float *)row)[i];
   case T_F64: return (float)((const double *)row)[i];
   }
   return 0.0f;
}

static int set_threads(int nthreads)
{
#ifdef _OPENMP
   if (nthreads > 0) omp_set_num_threads(nthreads);
   return omp_get_max_threads();
#else
   (void)nthreads;
   return 1;
#endif
}

/* ------------------------------------------------------------------------------------------
 * orc_prox_ref: one output per (x, y, c) in the reference's order.
 *   cl-files/ImageProximity.cl:52-53   left = -tw/2, top = -th/2 (C integer division)
 *   cl-files/ImageProximity.cl:55-70   h outer, w inner, one accumulator per sum
 *   cl-files/ImageProximity.cl:64-67   source outside [0,W)x[0,H) reads as 0
 *   cl-files/ImageProximity.cl:150-155, 404-409   Norm = sum / sqrt(sumSS * sumTT)
 *   cl-files/Buffers.h:58-63           REAL = float, double for F64 sources
 * `contract` = 1 forms every `acc += a*b` as one fused multiply-add (OpenCL's default
 * FP_CONTRACT), 0 rounds the product first.  Rows [y0, y1) are computed (the CPU baseline
 * times a band of rows).  dst is float (double for F64) with pitch dstep bytes, channels
 * interleaved.
 * ------------------------------------------------------------------------------------------ */
#define PROX_BODY(REALT, FMA, SQRT, ELEM)                                                     \
   for (int y = y0; y < y1; ++y) {                                                            \
      REALT *drow = (REALT *)((unsigned char *)dst + (size_t)y * dstep);                     \
      for (int x = 0; x < W; ++x) {                                                           \
         for (int c = 0; c < nch; ++c) {                                                      \
            REALT sum = 0, ss = 0, tt = 0;                                                    \
            for (int h = 0; h < th; ++h) {                                                    \
               const unsigned char *trow = (const unsigned char *)tpl + (size_t)h * tstep;    \
               int py = y + h + top;                                                          \
               const unsigned char *srow = (py >= 0 && py < H)                                \
                  ? (const unsigned char *)src + (size_t)py * sstep : NULL;                   \
               for (int w = 0; w < tw; ++w) {                                                 \
                  int px = x + w + left;                                                      \
                  REALT t = ELEM(trow, type, w * nch + c);                                    \
                  REALT s = 0;                                                                \
                  if (srow && px >= 0 && px < W) s = ELEM(srow, type, px * nch + c);          \
                  switch (op) {                                                               \
                  case OP_SSD: { REALT d = s - t; sum = contract ? FMA(d, d, sum) : sum + d * d; break; } \
                  case OP_SSD_NORM: { REALT d = s - t;                                        \
                     sum = contract ? FMA(d, d, sum) : sum + d * d;                           \
                     ss = contract ? FMA(s, s, ss) : ss + s * s;                              \
                     tt = contract ? FMA(t, t, tt) : tt + t * t; break; }                     \
                  case OP_SAD: sum += (REALT)fabs((double)(s - t)); break;                    \
                  case OP_CC: sum = contract ? FMA(s, t, sum) : sum + s * t; break;           \
                  case OP_CC_NORM:                                                            \
                     sum = contract ? FMA(s, t, sum) : sum + s * t;                           \
                     ss = contract ? FMA(s, s, ss) : ss + s * s;                              \
                     tt = contract ? FMA(t, t, tt) : tt + t * t; break;                       \
                  }                                                                           \
               }                                                                              \
            }                                                                                 \
            if (op == OP_SSD_NORM || op == OP_CC_NORM) {                                      \
               REALT den = SQRT(ss * tt);                                                     \
               sum = sum / den;                                                               \
            }                                                                                 \
            drow[x * nch + c] = sum;                                                          \
         }                                                                                    \
      }                                                                                       \
   }

static inline double elem_dd(const unsigned char *row, int t, int i) { return elem_d(row, t, i); }

int orc_prox_ref(int op, int type, int nch,
                 const void *src, int W, int H, int sstep,
                 const void *tpl, int tw, int th, int tstep,
                 void *dst, int dstep, int contract, int y0, int y1, int nthreads)
{
   if (op < 0 || op > 4 || type < 0 || type > T_F64 || nch < 1 || nch > 4) return -1;
   if (y0 < 0) y0 = 0;
   if (y1 > H) y1 = H;
   const int left = -(tw / 2), top = -(th / 2);
   set_threads(nthreads);
   if (type == T_F64) {
      #pragma omp parallel for schedule(dynamic, 1)
      for (int yy = y0; yy < y1; ++yy) {
         int y0_ = yy, y1_ = yy + 1;
         { int y0 = y0_, y1 = y1_; PROX_BODY(double, fma, sqrt, elem_dd) }
      }
   } else {
      #pragma omp parallel for schedule(dynamic, 1)
      for (int yy = y0; yy < y1; ++yy) {
         int y0_ = yy, y1_ = yy + 1;
         { int y0 = y0_, y1 = y1_; PROX_BODY(float, fmaf, sqrtf, elem_f) }
      }
   }
   return 0;
}

/* orc_prox_ref on the listed output rows only (the other rows of dst are left as they are):
 * spread samples of large maps in one parallel call. */
int orc_prox_ref_rows(int op, int type, int nch,
                      const void *src, int W, int H, int sstep,
                      const void *tpl, int tw, int th, int tstep,
                      void *dst, int dstep, int contract, const int *rows, int nrows,
                      int nthreads)
{
   if (op < 0 || op > 4 || type < 0 || type > T_F64 || nch < 1 || nch > 4) return -1;
   for (int k = 0; k < nrows; ++k)
      if (rows[k] < 0 || rows[k] >= H) return -1;
   const int left = -(tw / 2), top = -(th / 2);
   set_threads(nthreads);
   if (type == T_F64) {
      #pragma omp parallel for schedule(dynamic, 1)
      for (int k = 0; k < nrows; ++k) {
         int y0_ = rows[k], y1_ = rows[k] + 1;
         { int y0 = y0_, y1 = y1_; PROX_BODY(double, fma, sqrt, elem_dd) }
      }
   } else {
      #pragma omp parallel for schedule(dynamic, 1)
      for (int k = 0; k < nrows; ++k) {
         int y0_ = rows[k], y1_ = rows[k] + 1;
         { int y0 = y0_, y1 = y1_; PROX_BODY(float, fmaf, sqrtf, elem_f) }
      }
   }
   return 0;
}

/* ------------------------------------------------------------------------------------------
 * orc_prox_f64: the same window sums with every term formed from the exact element value
 * and accumulated in double; Norm computed in double.  dst is double, pitch W*nch.
 * (Ground truth for float sources: products of two floats are exact in double.)
 * ------------------------------------------------------------------------------------------ */
int orc_prox_f64(int op, int type, int nch,
                 const void *src, int W, int H, int sstep,
                 const void *tpl, int tw, int th, int tstep,
                 double *dst, int nthreads)
{
   if (op < 0 || op > 4 || type < 0 || type > T_F64 || nch < 1 || nch > 4) return -1;
   const int left = -(tw / 2), top = -(th / 2);
   set_threads(nthreads);
   #pragma omp parallel for schedule(dynamic, 1)
   for (int y = 0; y < H; ++y) {
      for (int x = 0; x < W; ++x) {
         for (int c = 0; c < nch; ++c) {
            double sum = 0, ss = 0, tt = 0;
            for (int h = 0; h < th; ++h) {
               const unsigned char *trow = (const unsigned char *)tpl + (size_t)h * tstep;
               int py = y + h + top;
               const unsigned char *srow = (py >= 0 && py < H)
                  ? (const unsigned char *)src + (size_t)py * sstep : NULL;
               for (int w = 0; w < tw; ++w) {
                  int px = x + w + left;
                  double t = elem_d(trow, type, w * nch + c);
                  double s = (srow && px >= 0 && px < W) ? elem_d(srow, type, px * nch + c) : 0.0;
                  switch (op) {
                  case OP_SSD: case OP_SSD_NORM: sum += (s - t) * (s - t); break;
                  case OP_SAD: sum += fabs(s - t); break;
                  default: sum += s * t; break;
                  }
                  ss += s * s;
                  tt += t * t;
               }
            }
            if (op == OP_SSD_NORM || op == OP_CC_NORM) sum = sum / sqrt(ss * tt);
            dst[((size_t)y * W + x) * nch + c] = sum;
         }
      }
   }
   return 0;
}

/* ------------------------------------------------------------------------------------------
 * orc_prox_i64: exact integer window sums for integer sources (the values the reference
 * would produce with unbounded precision).  which: 0 SSD, 1 SAD, 2 CC, 3 sum of s^2 over
 * the zero-padded window, 4 sum of t^2 (constant).  dst is int64, pitch W*nch.
 * ------------------------------------------------------------------------------------------ */
int orc_prox_i64(int which, int type, int nch,
                 const void *src, int W, int H, int sstep,
                 const void *tpl, int tw, int th, int tstep,
                 int64_t *dst, int nthreads)
{
   if (type < 0 || type >= T_F32 || which < 0 || which > 4 || nch < 1 || nch > 4) return -1;
   const int left = -(tw / 2), top = -(th / 2);
   set_threads(nthreads);
   #pragma omp parallel for schedule(dynamic, 1)
   for (int y = 0; y < H; ++y) {
      for (int x = 0; x < W; ++x) {
         for (int c = 0; c < nch; ++c) {
            __int128 acc = 0;
            for (int h = 0; h < th; ++h) {
               const unsigned char *trow = (const unsigned char *)tpl + (size_t)h * tstep;
               int py = y + h + top;
               const unsigned char *srow = (py >= 0 && py < H)
                  ? (const unsigned char *)src + (size_t)py * sstep : NULL;
               for (int w = 0; w < tw; ++w) {
                  int px = x + w + left;
                  int64_t t = elem_i(trow, type, w * nch + c);
                  int64_t s = (srow && px >= 0 && px < W) ? elem_i(srow, type, px * nch + c) : 0;
                  __int128 d = (__int128)s - t;
                  switch (which) {
                  case 0: acc += d * d; break;
                  case 1: acc += d < 0 ? -d : d; break;
                  case 2: acc += (__int128)s * t; break;
                  case 3: acc += (__int128)s * s; break;
                  case 4: acc += (__int128)t * t; break;
                  }
               }
            }
            dst[((size_t)y * W + x) * nch + c] = (int64_t)acc;
         }
      }
   }
   return 0;
}

/* ------------------------------------------------------------------------------------------
 * orc_integral_ref: the reference's 5-launch scan in its exact summation order, REAL = float
 * (dtype F32) or double (dtype F64).  sqr = 1 squares each element first (sqr_F32/sqr_F64).
 *   scan1 / sqr   cl-files/Integral.cl:39-161 (F32), 252-374 (F64):
 *        8-px serial prefix per work item; serial sum of the preceding 8-px run totals
 *        inside the 64-px group; serial 8-row vertical prefix inside the 8-row band.
 *   scan2 + scan3 cl-files/Integral.cl:164-204: serial sum of the 64-px group row totals to
 *        the left, added to every pixel of the group.
 *   scan4 + scan5 cl-files/Integral.cl:207-247: serial sum of the 8-row band column totals
 *        above, added to every pixel of the band.
 *   group counts  C++/programs/Integral.cpp:28-30 with C++/programs/WorkGroup.h:100-130.
 * The reference reads past the row end when W is not a multiple of 64 (and past the last
 * row when H is not a multiple of 8); those values only reach outputs that are never
 * written, so they are taken as 0 here.  dst has pitch dstep bytes.
 * ------------------------------------------------------------------------------------------ */
#define INTEGRAL_BODY(REALT, FMA)                                                             \
   {                                                                                          \
      const int G = 64, B = 8;                                                                \
      int NGW = (W + G - 1) / G, NGH = (H + B - 1) / B;                                       \
      int Wp = NGW * G, Hp = NGH * B;                                                         \
      REALT *T1 = (REALT *)calloc((size_t)Wp * Hp, sizeof(REALT));                            \
      REALT *D = (REALT *)calloc((size_t)Wp * Hp, sizeof(REALT));                             \
      if (!T1 || !D) { free(T1); free(D); return -2; }                                       \
      /* row pass inside each 64-px group */                                                  \
      for (int y = 0; y < Hp; ++y) {                                                          \
         const unsigned char *srow = y < H ? (const unsigned char *)src + (size_t)y * sstep : NULL; \
         for (int g = 0; g < NGW; ++g) {                                                      \
            REALT runtot[8];                                                                  \
            for (int k = 0; k < 8; ++k) {                                                     \
               REALT s = 0;                                                                   \
               for (int i = 0; i < 8; ++i) {                                                  \
                  int x = g * G + k * 8 + i;                                                  \
                  REALT v = (srow && x < W) ? (REALT)ELEMR(srow, type, x) : (REALT)0;         \
                  if (sqr) s = contract ? FMA(v, v, s) : s + v * v; else s += v;              \
                  T1[(size_t)y * Wp + x] = s;                                                 \
               }                                                                              \
               runtot[k] = s;                                                                 \
            }                                                                                 \
            for (int k = 0; k < 8; ++k) {                                                     \
               REALT p = 0;                                                                   \
               for (int j = 0; j < k; ++j) p += runtot[j];                                    \
               for (int i = 0; i < 8; ++i) {                                                  \
                  size_t o = (size_t)y * Wp + g * G + k * 8 + i;                              \
                  T1[o] = p + T1[o];                                                          \
               }                                                                              \
            }                                                                                 \
         }                                                                                    \
      }                                                                                       \
      /* vertical pass inside each 8-row band */                                              \
      for (int b = 0; b < NGH; ++b)                                                           \
         for (int x = 0; x < Wp; ++x) {                                                       \
            REALT s = 0;                                                                      \
            for (int i = 0; i < B; ++i) {                                                     \
               s += T1[(size_t)(b * B + i) * Wp + x];                                         \
               D[(size_t)(b * B + i) * Wp + x] = s;                                           \
            }                                                                                 \
         }                                                                                    \
      /* scan2 + scan3: group row totals to the left */                                       \
      for (int y = 0; y < H; ++y) {                                                           \
         REALT *vert = (REALT *)malloc(sizeof(REALT) * (NGW > 1 ? NGW - 1 : 1));              \
         for (int g = 0; g + 1 < NGW; ++g) {                                                  \
            REALT s = 0;                                                                      \
            for (int i = 0; i <= g; ++i) s += D[(size_t)y * Wp + (i + 1) * G - 1];            \
            vert[g] = s;                                                                      \
         }                                                                                    \
         for (int x = G; x < W; ++x) {                                                        \
            size_t o = (size_t)y * Wp + x;                                                    \
            D[o] = D[o] + vert[x / G - 1];                                                    \
         }                                                                                    \
         free(vert);                                                                          \
      }                                                                                       \
      /* scan4 + scan5: band column totals above */                                           \
      for (int x = 0; x < W; ++x) {                                                           \
         REALT *horiz = (REALT *)malloc(sizeof(REALT) * (NGH > 1 ? NGH - 1 : 1));             \
         for (int g = 0; g + 1 < NGH; ++g) {                                                  \
            REALT s = 0;                                                                      \
            for (int i = 0; i <= g; ++i) s += D[(size_t)((i + 1) * B - 1) * Wp + x];          \
            horiz[g] = s;                                                                     \
         }                                                                                    \
         for (int y = B; y < H; ++y) {                                                        \
            size_t o = (size_t)y * Wp + x;                                                    \
            D[o] = D[o] + horiz[y / B - 1];                                                   \
         }                                                                                    \
         free(horiz);                                                                         \
      }                                                                                       \
      for (int y = 0; y < H; ++y) {                                                           \
         REALT *drow = (REALT *)((unsigned char *)dst + (size_t)y * dstep);                  \
         for (int x = 0; x < W; ++x) drow[x] = D[(size_t)y * Wp + x];                         \
      }                                                                                       \
      free(T1); free(D);                                                                      \
   }

int orc_integral_ref(int sqr, int type, const void *src, int W, int H, int sstep,
                     int dtype, void *dst, int dstep, int contract)
{
   if (type < 0 || type > T_F64 || W <= 0 || H <= 0) return -1;
   if (dtype == T_F32) {
#define ELEMR elem_f
      INTEGRAL_BODY(float, fmaf)
#undef ELEMR
   } else if (dtype == T_F64) {
#define ELEMR elem_d
      INTEGRAL_BODY(double, fma)
#undef ELEMR
   } else {
      return -1;
   }
   return 0;
}

/* ------------------------------------------------------------------------------------------
 * orc_integral_exact: I(x,y) = sum_{j<=y} sum_{i<=x} f(S[j][i]), f = identity or square,
 * exact (int128 for integer sources, long double for float sources), rounded once to
 * double.  dst is double, pitch W.
 * ------------------------------------------------------------------------------------------ */
int orc_integral_exact(int sqr, int type, const void *src, int W, int H, int sstep, double *dst)
{
   if (type < 0 || type > T_F64 || W <= 0 || H <= 0) return -1;
   if (type < T_F32) {
      __int128 *col = (__int128 *)calloc((size_t)W, sizeof(__int128));
      if (!col) return -2;
      for (int y = 0; y < H; ++y) {
         const unsigned char *srow = (const unsigned char *)src + (size_t)y * sstep;
         __int128 rowacc = 0;
         for (int x = 0; x < W; ++x) {
            int64_t v = elem_i(srow, type, x);
            rowacc += sqr ? (__int128)v * v : (__int128)v;
            col[x] += rowacc;
            dst[(size_t)y * W + x] = (double)col[x];
         }
      }
      free(col);
   } else {
      long double *col = (long double *)calloc((size_t)W, sizeof(long double));
      if (!col) return -2;
      for (int y = 0; y < H; ++y) {
         const unsigned char *srow = (const unsigned char *)src + (size_t)y * sstep;
         long double rowacc = 0;
         for (int x = 0; x < W; ++x) {
            long double v = elem_d(srow, type, x);
            rowacc += sqr ? v * v : v;
            col[x] += rowacc;
            dst[(size_t)y * W + x] = (double)col[x];
         }
      }
      free(col);
   }
   return 0;
}

/* ------------------------------------------------------------------------------------------
 * orc_area_sum: window sum over [x - w/2, x - w/2 + w - 1] x [y - h/2, y - h/2 + h - 1]
 * clipped to the image, by four corners of an inclusive W x H integral
 * (cl-files/ImageProximityFFT.cl:55-84): Left = x - w/2 - 1, Top = y - h/2 - 1,
 * Right = min(x + w - w/2 - 1, W - 1), Bottom = min(y + h - h/2 - 1, H - 1);
 * corners with a negative index contribute 0.  Right or Bottom below 0 gives 0.
 * ------------------------------------------------------------------------------------------ */
int orc_area_sum(const double *I, int W, int H, int w, int h, double *dst)
{
   for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
         int L = x - w / 2 - 1, T = y - h / 2 - 1;
         int R = x + w - w / 2 - 1, Bm = y + h - h / 2 - 1;
         if (R > W - 1) R = W - 1;
         if (Bm > H - 1) Bm = H - 1;
         double br = (R >= 0 && Bm >= 0) ? I[(size_t)Bm * W + R] : 0.0;
         double tr = (R >= 0 && T >= 0) ? I[(size_t)T * W + R] : 0.0;
         double bl = (L >= 0 && Bm >= 0) ? I[(size_t)Bm * W + L] : 0.0;
         double tl = (L >= 0 && T >= 0) ? I[(size_t)T * W + L] : 0.0;
         dst[(size_t)y * W + x] = br - tr - (bl - tl);
      }
   return 0;
}

/* ------------------------------------------------------------------------------------------
 * orc_min_max_indx: Statistics::Min / Max with coordinates (want_max = 0 / 1) of a 1-channel
 * image, in the reference's order (its tie-breaks and its float partials).
 *
 *   cl-files/Statistics.cl:344-425 (REDUCE_POS_STD min_coord / max_coord; the _flush variant
 *   :436-507 visits the same pixels in the same order on images whose width is a multiple of
 *   256 and height a multiple of 16):
 *     work-groups of 16x16 work-items; work-item (lx, ly) of group (gxg, gyg) reads row
 *     gy = 16 gyg + ly at columns gx + 16 i, gx = 256 gxg + lx, i = 0..15 (:351-354); it is
 *     active when gx < W and gy < H (:357); it starts from the pixel at i = 0 and takes a later
 *     one only when strictly smaller (min) / larger (max), comparing in the source's own
 *     element type (type = SCALAR, :514-515; :359-370);
 *     the 256 results (lid = 16 ly + lx) are reduced by a halving tree, position lid taking
 *     position lid + k (k = 128, 64, .., 1) only when both are active and lid + k is strictly
 *     better (DO_REDUCE, :332-340; :379-417);
 *     position 0 is stored as float (global float *result, :346) with its coordinates,
 *     gid = gyg * ngx + gxg (:418-423).
 *   C++/programs/StatisticsHelpers.cpp:105-145 (ReduceMin / ReduceMax): start from group 0,
 *     take group g only when its float is strictly better; the value returned is that float.
 *   Group counts: C++/programs/WorkGroup.h:87-156 with 16 pixels per work-item along x
 *     (StatisticsHelpers.h:41): ngx = ceil(W / 256), ngy = ceil(H / 16).
 * NaN follows from the same comparisons (a NaN is never taken, and a NaN already held is
 * never replaced).  Returns 0.
 * ------------------------------------------------------------------------------------------ */
typedef struct { double v; int x, y, ok; } mm_item;

/* "b strictly better than a" in the element type: every element type converts to double
 * exactly, so comparing the doubles is comparing the elements */
static inline int mm_better(int want_max, double b, double a)
{
   return want_max ? b > a : b < a;
}

int orc_min_max_indx(int want_max, int type, const void *src, int W, int H, int sstep,
                     double *value, int *ix, int *iy)
{
   if (type < 0 || type > T_F64 || W <= 0 || H <= 0) return -1;
   const int ngx = (W + 255) / 256, ngy = (H + 15) / 16;
   float best = 0.0f;
   int bx = 0, by = 0;
   mm_item it[256];
   for (int gyg = 0; gyg < ngy; ++gyg)
      for (int gxg = 0; gxg < ngx; ++gxg) {
         for (int lid = 0; lid < 256; ++lid) {
            const int lx = lid & 15, ly = lid >> 4;
            const int gx = 256 * gxg + lx, gy = 16 * gyg + ly;
            mm_item *m = &it[lid];
            m->ok = gx < W && gy < H;
            if (!m->ok) continue;
            const unsigned char *srow = (const unsigned char *)src + (size_t)gy * sstep;
            m->v = elem_d(srow, type, gx);
            m->x = gx;
            m->y = gy;
            for (int i = 16; i < 256; i += 16)
               if (gx + i < W) {
                  const double px = elem_d(srow, type, gx + i);
                  if (mm_better(want_max, px, m->v)) { m->v = px; m->x = gx + i; }
               }
         }
         for (int k = 128; k >= 1; k >>= 1)
            for (int lid = 0; lid < k; ++lid)
               if (it[lid].ok && it[lid + k].ok && mm_better(want_max, it[lid + k].v, it[lid].v)) {
                  it[lid].v = it[lid + k].v;
                  it[lid].x = it[lid + k].x;
                  it[lid].y = it[lid + k].y;
               }
         const float part = (float)it[0].v;  /* RNE, as OpenCL's implicit conversion */
         if ((gxg == 0 && gyg == 0) || (want_max ? part > best : part < best)) {
            best = part; bx = it[0].x; by = it[0].y;
         }
      }
   *value = best; *ix = bx; *iy = by;
   return 0;
}

int orc_type_size(int t) { return type_size(t); }

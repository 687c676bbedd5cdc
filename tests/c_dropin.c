/* c_dropin.c -- plain C client of the template-matching C ABI, written only against
 * include/OpenCLIPP.h the way a user of the reference library writes it (the call sequence
 * of the reference's Test-OpenCLIPP.c and Bench: initialise, create images from host
 * buffers, run primitives, read back, release).  Compiled and linked by the CPU tests;
 * run by the GPU tests, which compare its output with the same calls made from Python.
 *
 * Usage: c_dropin <width> <height> <out.bin>
 * Writes, as raw float32, the SqrDistance and CrossCorr_Norm maps and the F32 integral of a
 * deterministic test image and template, then MaxIndx of the CrossCorr_Norm map as 3 floats.
 */
#include <stdio.h>
#include <stdlib.h>

#include "OpenCLIPP.h"

#define CHECK(call)                                                                  \
   do {                                                                              \
      ocipError e_ = (call);                                                         \
      if (e_ != CL_SUCCESS) {                                                        \
         fprintf(stderr, "%s failed: %s (%d)\n", #call, ocipGetErrorName(e_), e_);  \
         return 1;                                                                   \
      }                                                                              \
   } while (0)

int main(int argc, char **argv)
{
   if (argc != 4) {
      fprintf(stderr, "usage: %s width height out.bin\n", argv[0]);
      return 2;
   }
   const unsigned W = (unsigned)atoi(argv[1]), H = (unsigned)atoi(argv[2]);
   const unsigned TW = 16, TH = 16, OX = W / 3, OY = H / 4;
   float *src = malloc(sizeof(float) * W * H), *ssd = malloc(sizeof(float) * W * H);
   float *ncc = malloc(sizeof(float) * W * H), *integ = malloc(sizeof(float) * W * H);
   float *tpl = malloc(sizeof(float) * TW * TH);
   unsigned state = 12345u;
   for (unsigned i = 0; i < W * H; ++i) {
      state = state * 1664525u + 1013904223u;
      src[i] = (float)(state >> 8) / 16777216.0f;
   }
   for (unsigned y = 0; y < TH; ++y)
      for (unsigned x = 0; x < TW; ++x) tpl[y * TW + x] = src[(OY + y) * W + OX + x];

   ocipContext ctx;
   CHECK(ocipInitialize(&ctx, NULL, CL_DEVICE_TYPE_ALL));
   char name[128];
   CHECK(ocipGetDeviceName(name, sizeof name));
   fprintf(stderr, "device: %s\n", name);

   SImage simg = {W, H, W * sizeof(float), 1, F32};
   SImage stpl = {TW, TH, TW * sizeof(float), 1, F32};
   ocipImage hsrc, htpl, hssd, hncc, hint;
   CHECK(ocipCreateImage(&hsrc, simg, src, CL_MEM_READ_ONLY));
   CHECK(ocipCreateImage(&htpl, stpl, tpl, CL_MEM_READ_ONLY));
   CHECK(ocipCreateImage(&hssd, simg, ssd, CL_MEM_WRITE_ONLY));
   CHECK(ocipCreateImage(&hncc, simg, ncc, CL_MEM_WRITE_ONLY));
   CHECK(ocipCreateImage(&hint, simg, integ, CL_MEM_WRITE_ONLY));

   CHECK(ocipPrepareProximity(hsrc));
   CHECK(ocipSqrDistance(hsrc, htpl, hssd));
   CHECK(ocipCrossCorr_Norm(hsrc, htpl, hncc));

   ocipProgram integral;
   CHECK(ocipPrepareIntegral(&integral, hsrc));
   CHECK(ocipIntegral(integral, hsrc, hint));

   CHECK(ocipReadImage(hssd));
   CHECK(ocipReadImage(hncc));
   CHECK(ocipReadImage(hint));

   double best;
   int bx, by;
   CHECK(ocipMaxIndx(NULL, hncc, &best, &bx, &by));
   fprintf(stderr, "best match %g at (%d, %d), template cropped at (%u, %u)\n", best, bx, by,
           OX + TW / 2, OY + TH / 2);

   FILE *f = fopen(argv[3], "wb");
   if (!f) return 1;
   fwrite(ssd, sizeof(float), W * H, f);
   fwrite(ncc, sizeof(float), W * H, f);
   fwrite(integ, sizeof(float), W * H, f);
   float tail[3] = {(float)best, (float)bx, (float)by};
   fwrite(tail, sizeof(float), 3, f);
   fclose(f);

   CHECK(ocipReleaseProgram(integral));
   CHECK(ocipReleaseImage(hsrc));
   CHECK(ocipReleaseImage(htpl));
   CHECK(ocipReleaseImage(hssd));
   CHECK(ocipReleaseImage(hncc));
   CHECK(ocipReleaseImage(hint));
   CHECK(ocipUninitialize(ctx));
   free(src); free(ssd); free(ncc); free(integ); free(tpl);
   return (bx == (int)(OX + TW / 2) && by == (int)(OY + TH / 2)) ? 0 : 3;
}

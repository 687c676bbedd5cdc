// Runs every function of the CPU oracle (oracle/oracle.c), the CPU baseline
// (oracle/baseline.c) and the argument checks of openclipp_amd/csrc/validation.h on small
// inputs of every element type, channel count and edge shape, built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_sanitize.py (host code only, g++/gcc).  The
// sanitizers stop the program at the first error they find; the driver itself exits non-zero
// on the first unexpected return code.  Values are not checked here: tests/test_oracle.py pins
// them.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "validation.h"

extern "C" {
int orc_prox_ref(int op, int type, int nch, const void *src, int W, int H, int sstep,
                 const void *tpl, int tw, int th, int tstep, void *dst, int dstep, int contract,
                 int y0, int y1, int nthreads);
int orc_prox_ref_rows(int op, int type, int nch, const void *src, int W, int H, int sstep,
                      const void *tpl, int tw, int th, int tstep, void *dst, int dstep,
                      int contract, const int *rows, int nrows, int nthreads);
int orc_prox_f64(int op, int type, int nch, const void *src, int W, int H, int sstep,
                 const void *tpl, int tw, int th, int tstep, double *dst, int nthreads);
int orc_prox_i64(int which, int type, int nch, const void *src, int W, int H, int sstep,
                 const void *tpl, int tw, int th, int tstep, int64_t *dst, int nthreads);
int orc_integral_ref(int sqr, int type, const void *src, int W, int H, int sstep, int dtype,
                     void *dst, int dstep, int contract);
int orc_integral_exact(int sqr, int type, const void *src, int W, int H, int sstep,
                       double *dst);
int orc_area_sum(const double *I, int W, int H, int w, int h, double *dst);
int orc_min_max_indx(int want_max, int type, const void *src, int W, int H, int sstep,
                     double *value, int *ix, int *iy);
int orc_type_size(int t);
int base_cc_norm(const float *src, int W, int H, int sstep, const float *tpl, int tw, int th,
                 int tstep, float *dst, int dstep, int y0, int y1, int nthreads);
}

namespace {

int failures = 0;

void expect(bool ok, const char *what, int a, int b)
{
   if (!ok) {
      std::fprintf(stderr, "unexpected result: %s (%d, %d)\n", what, a, b);
      ++failures;
   }
}

uint32_t rng_state = 12345u;
uint32_t next_u32()
{
   rng_state = rng_state * 1664525u + 1013904223u;
   return rng_state;
}

// An image of `type` with `nch` channels, rows `pad` elements longer than needed (a pitched
// image), filled with values that exercise the type's range (including zero rows so that
// Norm windows of zeros occur), allocated to exactly Height * Step bytes.
std::vector<unsigned char> make_image(int type, int W, int H, int nch, int pad, int *step)
{
   const int esz = orc_type_size(type);
   *step = (W * nch + pad) * esz;
   std::vector<unsigned char> img((size_t)H * *step);
   for (int y = 0; y < H; ++y)
      for (int i = 0; i < W * nch + pad; ++i) {
         unsigned char *p = img.data() + (size_t)y * *step + (size_t)i * esz;
         const uint32_t r = (y % 5 == 4) ? 0u : next_u32();
         switch (type) {
         case 0: *p = (unsigned char)r; break;
         case 1: *(int8_t *)p = (int8_t)r; break;
         case 2: { uint16_t v = (uint16_t)r; std::memcpy(p, &v, 2); break; }
         case 3: { int16_t v = (int16_t)r; std::memcpy(p, &v, 2); break; }
         case 4: { uint32_t v = r; std::memcpy(p, &v, 4); break; }
         case 5: { int32_t v = (int32_t)r; std::memcpy(p, &v, 4); break; }
         case 6: { float v = (float)(r % 1000) / 999.0f; std::memcpy(p, &v, 4); break; }
         case 7: { double v = (double)(r % 1000) / 999.0; std::memcpy(p, &v, 8); break; }
         }
      }
   return img;
}

void proximity_cases()
{
   const int shapes[][4] = {{1, 1, 1, 1}, {7, 5, 3, 3}, {16, 9, 16, 9}, {20, 13, 4, 7},
                            {33, 17, 8, 2}};
   for (int type = 0; type < 8; ++type)
      for (int nch : {1, 2, 4})
         for (const auto &sh : shapes) {
            const int W = sh[0], H = sh[1], tw = sh[2], th = sh[3];
            int sstep, tstep;
            auto src = make_image(type, W, H, nch, 3, &sstep);
            auto tpl = make_image(type, tw, th, nch, 0, &tstep);
            const int real = type == 7 ? 8 : 4;
            std::vector<unsigned char> dst((size_t)H * W * nch * real);
            std::vector<double> d64((size_t)H * W * nch);
            std::vector<int64_t> i64((size_t)H * W * nch);
            for (int op = 0; op < 5; ++op) {
               for (int contract = 0; contract < 2; ++contract) {
                  int r = orc_prox_ref(op, type, nch, src.data(), W, H, sstep, tpl.data(), tw,
                                       th, tstep, dst.data(), W * nch * real, contract, 0, H, 1);
                  expect(r == 0, "orc_prox_ref", op, type);
               }
               const int rows[2] = {0, H - 1};
               int r = orc_prox_ref_rows(op, type, nch, src.data(), W, H, sstep, tpl.data(), tw,
                                         th, tstep, dst.data(), W * nch * real, 1, rows, 2, 1);
               expect(r == 0, "orc_prox_ref_rows", op, type);
               r = orc_prox_f64(op, type, nch, src.data(), W, H, sstep, tpl.data(), tw, th, tstep,
                                d64.data(), 1);
               expect(r == 0, "orc_prox_f64", op, type);
            }
            for (int which = 0; which < 5; ++which) {
               const int r = orc_prox_i64(which, type, nch, src.data(), W, H, sstep, tpl.data(),
                                          tw, th, tstep, i64.data(), 1);
               expect(r == (type < 6 ? 0 : -1), "orc_prox_i64", which, type);
            }
         }
   // rejected arguments return -1 without touching memory
   expect(orc_prox_ref(5, 6, 1, nullptr, 1, 1, 4, nullptr, 1, 1, 4, nullptr, 4, 1, 0, 1, 1) == -1,
          "orc_prox_ref bad op", 5, 0);
   expect(orc_prox_f64(0, 9, 1, nullptr, 1, 1, 4, nullptr, 1, 1, 4, nullptr, 1) == -1,
          "orc_prox_f64 bad type", 9, 0);
}

void integral_and_stats_cases()
{
   const int shapes[][2] = {{1, 1}, {5, 3}, {64, 9}, {257, 17}, {300, 40}};
   for (int type = 0; type < 8; ++type)
      for (const auto &sh : shapes) {
         const int W = sh[0], H = sh[1];
         int sstep;
         auto src = make_image(type, W, H, 1, 2, &sstep);
         std::vector<float> f32((size_t)W * H);
         std::vector<double> f64((size_t)W * H), area((size_t)W * H);
         for (int sqr = 0; sqr < 2; ++sqr) {
            expect(orc_integral_ref(sqr, type, src.data(), W, H, sstep, 6, f32.data(), W * 4, 1) == 0,
                   "orc_integral_ref F32", type, sqr);
            expect(orc_integral_ref(sqr, type, src.data(), W, H, sstep, 7, f64.data(), W * 8, 0) == 0,
                   "orc_integral_ref F64", type, sqr);
            expect(orc_integral_exact(sqr, type, src.data(), W, H, sstep, f64.data()) == 0,
                   "orc_integral_exact", type, sqr);
         }
         expect(orc_area_sum(f64.data(), W, H, 3, 5, area.data()) == 0, "orc_area_sum", W, H);
         expect(orc_area_sum(f64.data(), W, H, W + 2, H + 2, area.data()) == 0, "orc_area_sum big",
                W, H);
         for (int want_max = 0; want_max < 2; ++want_max) {
            double v;
            int ix = -1, iy = -1;
            const int r = orc_min_max_indx(want_max, type, src.data(), W, H, sstep, &v, &ix, &iy);
            expect(r == 0 && ix >= 0 && ix < W && iy >= 0 && iy < H, "orc_min_max_indx", ix, iy);
         }
      }
}

void baseline_cases()
{
   const int shapes[][4] = {{1, 1, 1, 1}, {37, 11, 5, 3}, {100, 20, 16, 16}, {130, 9, 33, 9}};
   for (const auto &sh : shapes) {
      const int W = sh[0], H = sh[1], tw = sh[2], th = sh[3];
      int sstep, tstep;
      auto src = make_image(6, W, H, 1, 1, &sstep);
      auto tpl = make_image(6, tw, th, 1, 0, &tstep);
      std::vector<float> dst((size_t)W * H);
      expect(base_cc_norm(reinterpret_cast<const float *>(src.data()), W, H, sstep,
                          reinterpret_cast<const float *>(tpl.data()), tw, th, tstep, dst.data(),
                          W * 4, 0, H, 1) == 0,
             "base_cc_norm", W, H);
   }
}

void validation_cases()
{
   // every descriptor field over a range that includes zero and huge values, and element
   // types past the last one (within the enum's value range): the checks must not overflow,
   // divide by zero or read anything
   const unsigned dims[] = {0u, 1u, 7u, 4096u, 0x7fffffffu, 0xffffffffu};
   const int types[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 15};
   int n = 0;
   for (unsigned w : dims)
      for (unsigned h : dims)
         for (unsigned step : dims)
            for (unsigned ch : {0u, 1u, 3u, 4u, 5u})
               for (int t : types) {
                  SImage a{w, h, step, ch, (SImage::EDataType)t};
                  SImage b{h, w, step, ch, (SImage::EDataType)((t + 2) % 9)};
                  n += ocip::check::create_image(a) == CL_SUCCESS;
                  (void)ocip::check::proximity(a, b, a);
                  (void)ocip::check::proximity_fft(a, b, a);
                  (void)ocip::check::fft(a, b);
                  if (ocip::check::elem_bytes(b.Type) > 0) (void)ocip::check::fft_layout(a, b);
                  (void)ocip::check::integral(a, b);
                  (void)ocip::check::precision(a, t);
               }
   expect(n > 0, "some descriptors valid", n, 0);
}

}  // namespace

int main()
{
   proximity_cases();
   integral_and_stats_cases();
   baseline_cases();
   validation_cases();
   if (failures) return 1;
   std::printf("sanitized run: ok\n");
   return 0;
}

// mfma.hip -- host side of the MFMA cross-correlation: tile geometry, the autotuner, the
// single-template launch and the batched launch (one persistent kernel per template-width
// bucket, captured once into a HIP graph and replayed).  Kernels: mfma_kernels.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "mfma_kernels.h"

namespace ocip {

using mfma::BatchDesc;
using mfma::BatchItem;
using mfma::BatchLaunch;
using mfma::SingleLaunch;
using mfma::u16x8;

namespace {

// work-queue counters of the batch kernels, cleared by a kernel (not a memset node) so that
// the recorded graph replays correctly
constexpr int kBatchCounters = 512;
__global__ void reset_counters(unsigned *c, int n)
{
   if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}

// CrossCorr_Norm from the CrossCorr map: every output C(x, y) of the map (written by the MFMA
// kernel) is scaled in place by 1 / sqrt(SS TT), SS(x, y) = sum of s^2 over the template window
// of (x, y) (origin (-tw/2, -th/2), zero outside the image) straight from the source image,
// TT = sum t^2 in FP64, the reciprocal square root in FP32 (inv_sqrt); C * (1 / sqrt(SS TT)) is
// one FP32 product.
// One wavefront owns kDenCols output columns and a band of kDenBand output rows:
//  * vertical pass, lane = column of the window segment [x0 - tw/2, x0 - tw/2 + kDenCols - 1
//    + tw): two FP64 running sums of s^2 from the band's first row, `lead` taking each row as
//    it enters the window and `lag` as it leaves, so that the column sum V = lead - lag is the
//    difference of two prefixes of one sequence (an all-zero column gives exactly 0).  V goes
//    to LDS as FP32 (relative error 2^-24), kDenRows output rows at a time;
//  * horizontal pass, lane = (output row, half of the columns): the same pair of running sums
//    along the row of V gives SS; the results go back through LDS and scale the map's rows
//    (loaded before the horizontal pass, so that their latency overlaps it).
// The rows are a stream of steps: step j takes row y0 - th/2 + j into `lead` (j >= 0), row
// y0 - th/2 + j - th into `lag` (j >= th) and completes output row y0 + j - (th - 1).  Steps
// run in batches of U, the loads of the next D - 1 batches in flight while a batch is summed;
// the stream starts at jstart <= 0 (steps before 0 load nothing) so that output rows fill the
// LDS block at batch boundaries.  No integral is formed: the source is read once from HBM or
// the caches (the halo columns and the lag rows are read a second time) and the map is read
// and written once.
constexpr int kDenRows = 32;     // output rows per LDS block
constexpr int kDenBand = 128;    // output rows per wavefront
constexpr int kDenCols = 64;     // output columns per wavefront (two halves of 32)
constexpr int kDenMaxChunks = 9; // 64-column chunks of the segment: tw <= 64 * 9 - 63

__device__ float kZeroPad = 0.f;   // never written: the zero read for positions outside the image

int den_chunks(int tw) { return (kDenCols - 1 + tw + 63) / 64; }
size_t den_lds_bytes(int tw) { return (size_t)kDenRows * (den_chunks(tw) * 64 + 1) * sizeof(float); }

// Offset that puts a buffer load out of range (the hardware returns 0): used for positions
// outside the image when the image holds fewer than kDenOob bytes.
constexpr int kDenOob = 0x40000000;

template <int NCH, bool BUF, bool V4>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NCH <= 3 ? 2 : 1, NCH <= 6 ? 2 : 1))) void ncc_den_src(const char *__restrict__ src, int sstep, int W,
                                                  int H, int tw, int th,
                                                  const double *__restrict__ tsums,
                                                  char *__restrict__ dst, int dstep, int bands)
{
   extern __shared__ float vbuf[];  // kDenRows x P floats; P odd: conflict-free row walks
   constexpr int P = 64 * NCH + 1;
   constexpr int U = 4;                          // steps per batch (divides kDenRows)
   constexpr int HALF = kDenCols / 2;           // outputs per lane, horizontal pass
   constexpr int BPB = kDenRows / U;            // output batches per LDS block (even)
   static_assert(BPB % 2 == 0, "two register sets of batch loads alternate within a block");
   // consecutive blocks (one XCD's run) are side by side along a band: neighbours share the
   // halo columns of the segment in L2
   const int b = xcd_tile(blockIdx.x, gridDim.x);
   const int cblks = gridDim.x / bands;
   const int band = b / cblks, cblk = b % cblks;
   const int lane = threadIdx.x;
   const int x0 = cblk * kDenCols, cx = tw / 2, cy = th / 2;
   const int y0 = band * kDenBand, n = min(kDenBand, H - y0);
   const double tt = tsums[0];
   int col[NCH], coff[NCH];
   bool cok[NCH];
#pragma unroll
   for (int k = 0; k < NCH; ++k) {
      const int c = x0 - cx + 64 * k + lane;
      cok[k] = c >= 0 && c < W && 64 * k + lane < kDenCols - 1 + tw;
      col[k] = min(max(c, 0), W - 1);
      coff[k] = cok[k] ? c * 4 : kDenOob;
   }
   const int pre = th - 1;                       // steps before the first output row
   // Rows the band reads: [rlo, rhi).  Every other row reads 0: below rlo (above the image,
   // steps before the stream, and the lag of steps j < th, which is always below y0 - cy)
   // through an out-of-range row offset, rhi and beyond (below the image, and steps past the
   // band's last output row) through the range of the buffer resource.
   const int rlo = max(0, y0 - cy), rhi = min(H, y0 - cy + pre + n);
   const __amdgpu_buffer_rsrc_t rsrc =
       __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(src), (short)0, BUF ? rhi * sstep : 0, 0x00020000);
   // D register sets of batch loads rotate with the batch (D - 1 batches in flight ahead of
   // the one being summed); jstart <= 0 with (pre - jstart) % U == 0 and a multiple b0 of D
   // batches before the first output row, so that every block starts on set 0
   constexpr int D = NCH <= 3 ? 4 : 2;
   static_assert(BPB % D == 0, "blocks hold whole rotations of the batch registers");
   int jstart = -((U - pre % U) % U);
   while (((pre - jstart) / U) % D) jstart -= U;
   const int b0 = (pre - jstart) / U;            // first batch that completes output rows

   // Every load is unconditional, so a whole batch is in flight at once.  BUF: a buffer load
   // whose offset (row + column, either kDenOob when outside) is out of range returns 0; the
   // row offset takes one scalar select.  Otherwise a position outside the image (or a step
   // that takes no row) reads the zero of kZeroPad (an address select, no branch around the
   // load).
   auto load_batch = [&](int bb, float (&L)[U][NCH], float (&G)[U][NCH]) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
         const int j = jstart + bb * U + u;
         const int lr = y0 - cy + j, gr = lr - th;
         if (BUF) {
            const int loff = lr >= rlo ? lr * sstep : kDenOob, goff = gr >= rlo ? gr * sstep : kDenOob;
#pragma unroll
            for (int k = 0; k < NCH; ++k) {
               const unsigned lo = (unsigned)coff[k] + (unsigned)loff, go = (unsigned)coff[k] + (unsigned)goff;
               L[u][k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)lo, 0, 0));
               G[u][k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)go, 0, 0));
            }
         } else {
            const bool live = j - pre < n;
            const bool lok = live && j >= 0 && lr >= 0 && lr < H;
            const bool gok = live && j >= th && gr >= 0 && gr < H;
            const float *lrow = reinterpret_cast<const float *>(src + (size_t)min(max(lr, 0), H - 1) * sstep);
            const float *grow = reinterpret_cast<const float *>(src + (size_t)min(max(gr, 0), H - 1) * sstep);
#pragma unroll
            for (int k = 0; k < NCH; ++k) {
               L[u][k] = *((lok && cok[k]) ? lrow + col[k] : &kZeroPad);
               G[u][k] = *((gok && cok[k]) ? grow + col[k] : &kZeroPad);
            }
         }
      }
   };

   double lead[NCH], lag[NCH];
#pragma unroll
   for (int k = 0; k < NCH; ++k) lead[k] = lag[k] = 0.0;

   // BUF: the map through a buffer resource too, lane = column at byte offset xoff (kDenOob
   // past the image's width); rows at and below H read 0 and take no store
   int xoff[kDenCols / 64];
#pragma unroll
   for (int hh = 0; hh < kDenCols / 64; ++hh) {
      const int x = x0 + 64 * hh + lane;
      xoff[hh] = x < W ? x * 4 : kDenOob;
   }
   // V4 (BUF, W a multiple of 4, 16-byte aligned map rows): lane = 4 columns 4 g4 .. 4 g4 + 3
   // of rows r4, r4 + 4, ..., 16-byte loads and stores
   static_assert(!V4 || (BUF && kDenCols == 64), "16-byte map access: buffer path, 64 columns");
   const int g4 = lane & 15, r4 = lane >> 4;
   const int xoff4 = x0 + 4 * g4 < W ? (x0 + 4 * g4) * 4 : kDenOob;

   // the map's values of output rows [ybase, ybase + nrows) (BUF: rows up to ybase + kDenRows,
   // the rows past nrows never stored): lane = column, unconditional loads, issued before the
   // block's batches so that waiting for them later does not wait for the newer batch loads
   float cval[kDenCols / 64][kDenRows];
   auto load_cval = [&](int ybase, int nrows) {
      const __amdgpu_buffer_rsrc_t mrsrc =
          __builtin_amdgcn_make_buffer_rsrc(dst, (short)0, BUF ? H * dstep : 0, 0x00020000);
      if (V4) {
#pragma unroll
         for (int k = 0; k < kDenRows / 4; ++k) {
            const auto q = __builtin_amdgcn_raw_buffer_load_b128(
                mrsrc, (int)((unsigned)xoff4 + (unsigned)((ybase + r4 + 4 * k) * dstep)), 0, 0);
#pragma unroll
            for (int e = 0; e < 4; ++e) cval[0][4 * k + e] = __uint_as_float(q[e]);
         }
         return;
      }
#pragma unroll
      for (int hh = 0; hh < kDenCols / 64; ++hh) {
         const int x = x0 + 64 * hh + lane;
#pragma unroll
         for (int rr = 0; rr < kDenRows; ++rr) {
            if (BUF) {
               cval[hh][rr] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                   mrsrc, (int)((unsigned)xoff[hh] + (unsigned)((ybase + rr) * dstep)), 0, 0));
            } else {
               const float *p = reinterpret_cast<const float *>(dst + (size_t)(ybase + rr) * dstep) + x;
               cval[hh][rr] = *((x < W && rr < nrows) ? p : &kZeroPad);
            }
         }
      }
   };

   // horizontal pass and scaling of the nrows rows of the LDS block (output rows ybase..)
   auto flush = [&](int ybase, int nrows) {
      __syncthreads();
      const int r = lane & 31, xs = (lane >> 5) * HALF;
      float res[HALF];
      if (r < nrows) {
         const float *row = vbuf + r * P + xs;
         double hl = 0.0, hg = 0.0;
         for (int i = 0; i < tw - 1; ++i) hl += (double)row[i];
         // the FP32 reciprocal square root of inv_sqrt for every output, the range that it
         // needs checked once per lane: a lane with any product outside it runs the same
         // running sums again with inv_sqrt
         const double hl0 = hl;
         double pmin = 1e30, pmax = 1e-30;
#pragma unroll
         for (int x = 0; x < HALF; ++x) {
            hl += (double)row[x + tw - 1];
            const double ss = hl - hg;
            hg += (double)row[x];
            const double p = ss * tt;
            res[x] = __builtin_amdgcn_rsqf((float)p);
            pmin = fmin(pmin, p);
            pmax = fmax(pmax, p);
         }
         if (!(pmin >= 1e-30 && pmax <= 1e30)) {
            hl = hl0;
            hg = 0.0;
#pragma unroll
            for (int x = 0; x < HALF; ++x) {
               hl += (double)row[x + tw - 1];
               const double ss = hl - hg;
               hg += (double)row[x];
               res[x] = inv_sqrt(ss * tt);
            }
         }
      }
      __syncthreads();
      if (r < nrows) {
#pragma unroll
         for (int x = 0; x < HALF; ++x) vbuf[r * P + xs + x] = res[x];
      }
      __syncthreads();
      if (V4) {
         const __amdgpu_buffer_rsrc_t srsrc =
             __builtin_amdgcn_make_buffer_rsrc(dst, (short)0, (ybase + nrows) * dstep, 0x00020000);
#pragma unroll
         for (int k = 0; k < kDenRows / 4; ++k) {
            const int rr = r4 + 4 * k;
            __attribute__((ext_vector_type(4))) unsigned q;
#pragma unroll
            for (int e = 0; e < 4; ++e)
               q[e] = __float_as_uint(cval[0][4 * k + e] * vbuf[rr * P + 4 * g4 + e]);
            __builtin_amdgcn_raw_buffer_store_b128(
                q, srsrc, (int)((unsigned)xoff4 + (unsigned)((ybase + rr) * dstep)), 0, 0);
         }
      } else if (BUF) {
         // a resource ending at row ybase + nrows drops the stores of the rows past the block
         const __amdgpu_buffer_rsrc_t srsrc =
             __builtin_amdgcn_make_buffer_rsrc(dst, (short)0, (ybase + nrows) * dstep, 0x00020000);
#pragma unroll
         for (int hh = 0; hh < kDenCols / 64; ++hh) {
#pragma unroll
            for (int rr = 0; rr < kDenRows; ++rr)
               __builtin_amdgcn_raw_buffer_store_b32(
                   __float_as_uint(cval[hh][rr] * vbuf[rr * P + 64 * hh + lane]), srsrc,
                   (int)((unsigned)xoff[hh] + (unsigned)((ybase + rr) * dstep)), 0, 0);
         }
      } else
#pragma unroll
      for (int hh = 0; hh < kDenCols / 64; ++hh) {
         const int x = x0 + 64 * hh + lane;
         if (x < W) {
#pragma unroll
            for (int rr = 0; rr < kDenRows; ++rr)
               if (rr < nrows)
                  reinterpret_cast<float *>(dst + (size_t)(ybase + rr) * dstep)[x] =
                      cval[hh][rr] * vbuf[rr * P + 64 * hh + lane];
         }
      }
      __syncthreads();
   };

   // vertical pass of batch bb: the running sums, and (OUT: batches bb >= b0) the column sums
   // V of completed output rows into the LDS block.  Before b0 every step is j < th - 1, whose
   // lag is 0, and completes no output row.  Rows of the last block past the band's end land
   // in LDS rows that the flush does not read.
   auto process = [&](auto out, int bb, const float (&L)[U][NCH], const float (&G)[U][NCH]) {
      constexpr bool OUT = decltype(out)::value;
      const int e0 = (bb - b0) * U;             // output row (band-relative) of step u = 0
#pragma unroll
      for (int u = 0; u < U; ++u) {
#pragma unroll
         for (int k = 0; k < NCH; ++k) {
            lead[k] = fma((double)L[u][k], (double)L[u][k], lead[k]);
            if (OUT) lag[k] = fma((double)G[u][k], (double)G[u][k], lag[k]);
         }
         if (OUT) {
#pragma unroll
            for (int k = 0; k < NCH; ++k)
               vbuf[((e0 + u) % kDenRows) * P + 64 * k + lane] = (float)(lead[k] - lag[k]);
         }
      }
   };

   float Ls[D][U][NCH], Gs[D][U][NCH];
#pragma unroll
   for (int d = 0; d < D - 1; ++d) load_batch(d, Ls[d], Gs[d]);
   // batches bb .. bb + D - 1 (bb a multiple of D): each step loads batch + D - 1 into the set
   // freed by the previous batch, then sums its own
   // (the scheduling barrier keeps each batch's loads ahead of the summing of the batch before
   // it: left to itself the compiler sinks them below, and every rotation then waits for all
   // loads in flight)
   auto rotation = [&](auto out, int bb) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
         load_batch(bb + d + D - 1, Ls[(d + D - 1) % D], Gs[(d + D - 1) % D]);
         __builtin_amdgcn_sched_barrier(0);
         process(out, bb + d, Ls[d], Gs[d]);
      }
   };
   for (int bb = 0; bb < b0; bb += D) rotation(std::false_type(), bb);
   // blocks of kDenRows output rows, BPB (even) batches each
   for (int first = 0; first < n; first += kDenRows) {
      const int nrows = min(kDenRows, n - first);
      load_cval(y0 + first, nrows);
      const int bbase = b0 + first / U;
      for (int i = 0; i < BPB; i += D) rotation(std::true_type(), bbase + i);
      flush(y0 + first, nrows);
   }
}

// Wider templates (tw > 64 * kDenMaxChunks - 63): one output per thread, four corners of the
// FP64 integral each, scaling the CrossCorr map in place.
__global__ __launch_bounds__(256) void ncc_scale_direct(const double *__restrict__ integ, int W,
                                                        int H, int tw, int th,
                                                        const double *__restrict__ tsums,
                                                        char *__restrict__ dst, int dstep)
{
   const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
   if (k >= (long long)W * H) return;
   const int y = (int)(k / W), x = (int)(k - (long long)y * W);
   float *d = reinterpret_cast<float *>(dst + (size_t)y * dstep) + x;
   *d = *d * inv_sqrt(window_sqsum(integ, W, H, x, y, tw, th) * tsums[0]);
}

int kc_for(int tw) { return tw <= 97 ? std::max(2, (tw + 31 + 15) / 16) : 6; }
int nub_for(int tw) { return tw <= 97 ? 1 : (tw + mfma::kBlockW - 1) / mfma::kBlockW; }
int planes_for(int precision) { return precision == mfma::kBF16x3 ? 2 : 1; }

struct Geometry {
   int ldsp, vb;
   size_t lds;
};

// LDS rows of lcols_for(KC, NT) 16-bit values at a pitch that is an odd multiple of 16 bytes,
// so that any 8 consecutive rows (the 8 lanes of one 128-byte phase of an A-fragment read)
// start on different 16-byte bank groups; as many template rows per band as fit in 160 KiB.
Geometry geometry(int precision, int KC, int NT, int th_max)
{
   Geometry g;
   const int lcols = mfma::lcols_for(KC, NT);
   g.ldsp = (lcols + 7) / 8 * 8;
   while ((g.ldsp * 2) % 32 != 16) g.ldsp += 8;
   const long rows_fit = (long)mfma::kLdsMax / ((long)planes_for(precision) * g.ldsp * 2);
   g.vb = (int)std::min<long>(th_max, rows_fit - (mfma::kTileRows - 1));
   g.lds = g.vb >= 1 ? (size_t)planes_for(precision) * (mfma::kTileRows + g.vb - 1) * g.ldsp * 2 : 0;
   return g;
}

// Row-pair kernel: chunks of 16 source columns per 16 output columns, and the LDS for
// 126 + 2 vb staged rows (vb steps of two template rows per band) or for the 129 x
// (16 NT + 4) FP32 tile of the epilogue, whichever is larger.
int kc2_for(int tw) { return tw <= 97 ? (tw + 15 + 15) / 16 : (mfma::kBlockW + 15 + 15) / 16; }

Geometry pair_geometry(int KC, int NT, int th, long lds_cap = mfma::kLdsMax, int planes = 1)
{
   Geometry g;
   const int lcols = mfma::pair_lcols(KC, NT);
   g.ldsp = (lcols + 7) / 8 * 8;
   while ((g.ldsp * 2) % 32 != 16) g.ldsp += 8;
   const long rows_fit = lds_cap / ((long)planes * g.ldsp * 2);
   const int S = (th + 1) / 2;
   g.vb = (int)std::min<long>(S, (rows_fit - (mfma::kTileRows - 2)) / 2);
   const size_t staging = (size_t)planes * (mfma::kTileRows - 2 + 2 * g.vb) * g.ldsp * 2;
   const size_t epi = (size_t)(mfma::kTileRows + 1) * mfma::pair_tile_pitch(NT) * sizeof(float);
   g.lds = g.vb >= 1 ? std::max(staging, epi) : 0;
   return g;
}

// Strip kernel: a ring of kRingRows LDS rows (pitch = 8 mod 16 elements: the 16-byte A reads of
// 16 consecutive rows fall on distinct bank groups) and the 4 x 64 FP32 row buffer.
constexpr int kStripNT = 4;

int ring_ldsp(int lcols) { return (lcols + 7) / 16 * 16 + 8; }

size_t strip_lds(int KC)
{
   const int lcols = mfma::pair_lcols(KC, kStripNT);
   return (size_t)mfma::kRingRows * ring_ldsp(lcols) * 2 + 4 * 16 * kStripNT * sizeof(float);
}

// rows per step of the ring loads: one 8-column item per thread
int strip_q_rows(int KC) { return 256 / (mfma::pair_lcols(KC, kStripNT) / 8); }

// The ring schedule of crosscorr_mfma_strip for runs of 1, 2 and 3 tiles (longer runs repeat
// the steady state of the third): every row read at a step was written to LDS at an earlier
// step (or staged before the first), and no write replaces a row still to be read.
bool ring_schedule_ok(int S, int q)
{
   const int T = mfma::kPairRows, R = mfma::kRingRows, D = mfma::kRingDepth;
   for (int K = 1; K <= 3; ++K) {
      const int grows = T * (K - 1) + 2 * S + mfma::kTileRows - 2;
      std::vector<int> written(grows, -2), slot(R, -1);
      const int fill = std::min(grows, mfma::kRingFill);
      for (int g = 0; g < fill; ++g) {
         written[g] = -1;
         slot[g % R] = g;
      }
      int ih = fill;
      std::vector<std::pair<int, int>> batch(K * S, {0, 0});
      for (int k = 0; k < K; ++k)
         for (int s = 0; s < S; ++s) {
            const int n = k * S + s, live = T * k + 2 * s;
            if (n >= D) {
               for (int g = batch[n - D].first; g < batch[n - D].second; ++g) {
                  if (slot[g % R] >= live) return false;
                  slot[g % R] = g;
                  written[g] = n;
               }
            }
            const int cnt = std::max(0, std::min(q, std::min(grows, live + R) - ih));
            batch[n] = {ih, ih + cnt};
            ih += cnt;
            for (int g = live; g < live + mfma::kTileRows; ++g)
               if (g >= grows || written[g] == -2 || written[g] >= n || slot[g % R] != g) return false;
         }
      for (int n = std::max(0, K * S - D); n < K * S; ++n)
         if (batch[n].second > batch[n].first) return false;
      if (ih != grows) return false;
   }
   return true;
}

bool strip_fits(int precision, int tw, int th)
{
   if (precision == mfma::kBF16x3 || tw > 97) return false;
   const int KC = kc2_for(tw);
   return KC >= 2 && KC <= 6 && strip_lds(KC) <= (size_t)mfma::kLdsMax &&
          ring_schedule_ok(mfma::pair_frag_steps(th, 1), strip_q_rows(KC));
}

int device_cus()
{
   static int cus = 0;
   if (!cus) {
      int dev = 0;
      hip_check(hipGetDevice(&dev), "device");
      hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "device");
   }
   return cus;
}

void launch_single_p(int precision, hipStream_t s, const SingleLaunch &L)
{
   if (precision == mfma::kBF16) mfma::launch_single<mfma::kBF16>(s, L);
   else if (precision == mfma::kF16) mfma::launch_single<mfma::kF16>(s, L);
   else mfma::launch_single<mfma::kBF16x3>(s, L);
}

void launch_batch_p(int precision, hipStream_t s, const BatchLaunch &L)
{
   if (precision == mfma::kBF16) mfma::launch_batch<mfma::kBF16>(s, L);
   else if (precision == mfma::kF16) mfma::launch_batch<mfma::kF16>(s, L);
   else mfma::launch_batch<mfma::kBF16x3>(s, L);
}

void launch_pair_batch_p(int precision, hipStream_t s, const BatchLaunch &L)
{
   if (precision == mfma::kBF16) mfma::launch_pair_batch<mfma::kBF16>(s, L);
   else mfma::launch_pair_batch<mfma::kF16>(s, L);
}

void launch_pair_batch_prep_p(int precision, hipStream_t s, const BatchDesc *d, int n)
{
   if (precision == mfma::kBF16) mfma::launch_pair_batch_prep<mfma::kBF16>(s, d, n);
   else mfma::launch_pair_batch_prep<mfma::kF16>(s, d, n);
}

void launch_batch_prep_p(int precision, hipStream_t s, const BatchDesc *d, int n)
{
   if (precision == mfma::kBF16) mfma::launch_batch_prep<mfma::kBF16>(s, d, n);
   else if (precision == mfma::kF16) mfma::launch_batch_prep<mfma::kF16>(s, d, n);
   else mfma::launch_batch_prep<mfma::kBF16x3>(s, d, n);
}

// rows and base of the source allow 16-byte loads
bool rows_aligned(const View &S)
{
   return S.step % 16 == 0 && reinterpret_cast<uintptr_t>(S.ptr) % 16 == 0;
}

// ---------------------------------------------------------------------------------------
// Autotuner: the kernel variant for each (kind, precision, mode, template size, image size,
// row pitch and alignment, device).  Variants 2 and 4: crosscorr_mfma with NT = 2 or 4 32-column n-tiles per wave
// (64- or 128-column output tiles); 14 and 16: crosscorr_mfma_pair with NT = 4 or 6 16-column
// n-tiles (64- or 96-column tiles); 24: crosscorr_mfma_strip (NT = 4).  The first call with a
// new key times every candidate with HIP events on the call's own inputs and keeps the
// fastest; the table is kept for the process and, when the environment variable
// OCIP_AUTOTUNE_CACHE names a file, read from and appended to that file.  An entry is used
// only when it names one of the calling op's candidates (a variant it can run).
bool known_variant(int v) { return v == 2 || v == 4 || v == 14 || v == 16 || v == 24; }

std::map<std::string, int> &tune_table()
{
   static std::map<std::string, int> table;
   static bool loaded = false;
   if (!loaded) {
      loaded = true;
      if (const char *path = getenv("OCIP_AUTOTUNE_CACHE")) {
         if (FILE *f = fopen(path, "r")) {
            char key[256];
            int nt;
            while (fscanf(f, "%255s %d", key, &nt) == 2)
               if (known_variant(nt)) table[key] = nt;
            fclose(f);
         }
      }
   }
   return table;
}

int tuned_variant(const std::string &key, hipStream_t s, const std::function<void(int)> &run,
                  std::initializer_list<int> candidates)
{
   auto &table = tune_table();
   auto it = table.find(key);
   // a table or cache-file entry is used only when it names one of this call's candidates,
   // i.e. a variant whose preconditions this call meets; anything else is tuned again
   if (it != table.end()) {
      for (int c : candidates)
         if (c == it->second) return c;
      table.erase(it);
   }
   hipEvent_t e0, e1;
   hip_check(hipEventCreate(&e0), "event");
   hip_check(hipEventCreate(&e1), "event");
   float best_ms = 1e30f;
   int best = *candidates.begin();
   // the fastest of three timed runs per candidate (after one warm run): run-to-run spread
   // is a few per cent, as large as the gaps between the leading variants
   for (int nt : candidates) {
      run(nt);  // warm
      for (int rep = 0; rep < 3; ++rep) {
         hip_check(hipEventRecord(e0, s), "event");
         run(nt);
         hip_check(hipEventRecord(e1, s), "event");
         hip_check(hipEventSynchronize(e1), "event");
         float ms = 0;
         hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
         if (ms < best_ms) {
            best_ms = ms;
            best = nt;
         }
      }
   }
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   table[key] = best;
   if (const char *path = getenv("OCIP_AUTOTUNE_CACHE")) {
      if (FILE *f = fopen(path, "a")) {
         fprintf(f, "%s %d\n", key.c_str(), best);
         fclose(f);
      }
   }
   return best;
}

// The device the tuning was measured on: architecture name and CU count of the current GPU.
const std::string &device_tag()
{
   static std::string tag;
   if (tag.empty()) {
      int dev = 0;
      hip_check(hipGetDevice(&dev), "device");
      hipDeviceProp_t prop;
      hip_check(hipGetDeviceProperties(&prop, dev), "device");
      char buf[96];
      snprintf(buf, sizeof buf, "%s:cu%d", prop.gcnArchName, prop.multiProcessorCount);
      for (char *c = buf; *c; ++c)
         if (*c == ' ') *c = '_';  // one whitespace-free token per key in the cache file
      tag = buf;
   }
   return tag;
}

// Everything a variant's validity and speed depend on: the templates (`shape`: a single
// template's width and height, from which the fragment counts KC, KC2 and the width blocks
// follow, or a batch bucket's KC and tallest template), the image size, row pitch (the strip
// kernel's 32-bit offsets) and 16-byte alignment, and the device.
std::string tune_key(const char *kind, int precision, int mode, const char *shape, const View &S)
{
   char buf[256];
   snprintf(buf, sizeof buf, "%s:p%d:m%d:%s:%dx%d:s%d:a%d:%s", kind, precision, mode, shape,
            S.width, S.height, S.step, rows_aligned(S) ? 1 : 0, device_tag().c_str());
   return buf;
}

}  // namespace

// ======================================================================= single template
// Work buffer: fragments of crosscorr_mfma, then fragments of the row-pair kernels (one plane).
static size_t round256(size_t b) { return (b + 255) / 256 * 256; }

static size_t mfma_frag_bytes(int precision, int tw, int th)
{
   return round256((size_t)planes_for(precision) * nub_for(tw) * th * kc_for(tw) * 64 * sizeof(u16x8));
}

static size_t pair_frag_count(int tw, int th)
{
   const int nub = nub_for(tw);
   return (size_t)nub * mfma::pair_frag_steps(th, nub) * kc2_for(tw);
}

static size_t pair_frag_bytes(int precision, int tw, int th)
{
   return round256((size_t)planes_for(precision) * pair_frag_count(tw, th) * 64 * sizeof(u16x8));
}

bool crosscorr_mfma_needs_integral(int tw) { return den_chunks(tw) > kDenMaxChunks; }

size_t crosscorr_mfma_work_bytes(int precision, int, int, int tw, int th)
{
   return mfma_frag_bytes(precision, tw, th) + pair_frag_bytes(precision, tw, th);
}

// CrossCorr_Norm: the CrossCorr map scaled in place by 1 / sqrt(SS TT)
static void launch_ncc_scale(hipStream_t s, const View &S, int tw, int th, const double *tsums,
                             const double *integ, float *dst, int dstep)
{
   KernelTimer kt(s, "ncc_den");
   char *d = reinterpret_cast<char *>(dst);
   if (den_chunks(tw) <= kDenMaxChunks) {
      const int bands = (S.height + kDenBand - 1) / kDenBand;
      const int cblks = (S.width + kDenCols - 1) / kDenCols;
      const size_t lds = den_lds_bytes(tw);
      auto go = [&](auto kernel) {
         if (lds > 64 * 1024)
            hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                      "lds attribute");
         kernel<<<bands * cblks, 64, lds, s>>>(static_cast<const char *>(S.ptr), S.step, S.width,
                                               S.height, tw, th, tsums, d, dstep, bands);
      };
      // buffer loads and stores when every row offset the kernel forms (source rows -th - 64 ..
      // H + th + 64, map rows up to H + kDenRows) and two out-of-range offsets added stay
      // below 2^31
      const bool buf = (long long)(S.height + th + 64) * S.step <= kDenOob &&
                       (long long)(S.height + kDenRows) * dstep <= kDenOob;
      // 16-byte map loads and stores: whole 4-column groups, 16-byte aligned rows
      const bool v4 = buf && S.width % 4 == 0 && dstep % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(d) % 16 == 0;
#define OCIP_DEN_CASE(N) \
   case N: if (v4) go(ncc_den_src<N, true, true>); else if (buf) go(ncc_den_src<N, true, false>); \
           else go(ncc_den_src<N, false, false>); break;
      switch (den_chunks(tw)) {
      OCIP_DEN_CASE(1)
      OCIP_DEN_CASE(2)
      OCIP_DEN_CASE(3)
      OCIP_DEN_CASE(4)
      OCIP_DEN_CASE(5)
      OCIP_DEN_CASE(6)
      OCIP_DEN_CASE(7)
      OCIP_DEN_CASE(8)
      default: if (v4) go(ncc_den_src<9, true, true>); else if (buf) go(ncc_den_src<9, true, false>);
               else go(ncc_den_src<9, false, false>); break;
      }
#undef OCIP_DEN_CASE
   } else {
      if (!integ) fail(CL_INVALID_VALUE, "integral required for templates this wide");
      const long long n = (long long)S.width * S.height;
      ncc_scale_direct<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(integ, S.width, S.height, tw, th,
                                                                  tsums, d, dstep);
   }
   launch_check("ncc_den");
}

void launch_crosscorr_mfma(hipStream_t s, int precision, bool norm, const View &S, const View &T,
                           double *tsums, const double *integ, float *dst, int dstep, void *work,
                           bool may_tune)
{
   const int tw = T.width, th = T.height;
   if (tw <= 0 || th <= 0) fail(CL_INVALID_VALUE, "empty template");
   const int KC = kc_for(tw), KC2 = kc2_for(tw), nub = nub_for(tw);
   if (geometry(precision, KC, 4, th).vb < 1)
      fail(CL_INVALID_VALUE, "template too large for the MFMA tile");
   const bool pair_ok = true;   // every precision runs on the row-pair tile (split BF16: two planes)
   // the strip kernel's buffer loads and stores address source and output with 32-bit offsets
   // below kRingOob
   const bool strip_ok = strip_fits(precision, tw, th) &&
                         (long long)S.height * S.step <= (long long)mfma::kRingOob &&
                         (long long)S.height * dstep <= (long long)mfma::kRingOob;
   // the contraction is the same for CrossCorr and CrossCorr_Norm: one tuning key
   char shape[48];
   snprintf(shape, sizeof shape, "tw%d:th%d", tw, th);
   const std::string key = tune_key("single", precision, mfma::kModeCC, shape, S);
   // OCIP_MFMA_VARIANT (2, 4, 14, 16 or 24) names a variant and skips the autotuner
   int forced = 0;
   if (const char *e = getenv("OCIP_MFMA_VARIANT")) {
      const int v = atoi(e);
      if (known_variant(v) && (pair_ok || v < 10) && (strip_ok || v != 24)) forced = v;
   }
   // the variants whose preconditions this call meets, in the autotuner's order
   const std::initializer_list<int> cands_strip = {24, 16, 14, 4, 2}, cands_pair = {16, 14, 4, 2},
                                    cands_plain = {4, 2};
   const std::initializer_list<int> &cands = strip_ok ? cands_strip : pair_ok ? cands_pair
                                                                              : cands_plain;
   // a tuned variant that this call cannot run counts as untuned (0: prepare every candidate's
   // fragments, then tune again)
   int chosen = forced;
   if (!chosen) {
      const auto known = tune_table().find(key);
      if (known != tune_table().end())
         for (int c : cands)
            if (c == known->second) chosen = c;
   }
   // without tuning: the row-pair kernel with 96-column tiles (within noise of the fastest
   // variant from 32x32 to 64x64 templates)
   if (!chosen && !may_tune) chosen = pair_ok ? 16 : 4;

   char *wk = static_cast<char *>(work);
   const size_t nfrag = (size_t)nub * th * KC;
   u16x8 *f0 = reinterpret_cast<u16x8 *>(wk);
   u16x8 *f1 = precision == mfma::kBF16x3 ? f0 + nfrag * 64 : nullptr;
   u16x8 *g0 = reinterpret_cast<u16x8 *>(wk + mfma_frag_bytes(precision, tw, th));
   u16x8 *g1 = precision == mfma::kBF16x3 ? g0 + pair_frag_count(tw, th) * 64 : nullptr;
   {
      // fragments of the chosen variant (of every candidate before the first tuning)
      KernelTimer kt(s, "template_prep");
      if (chosen < 10) {
         const BatchDesc d{static_cast<const char *>(T.ptr), tw, th, T.step, KC, nub, f0, f1, tsums};
         if (precision == mfma::kBF16) mfma::launch_single_prep<mfma::kBF16>(s, d);
         else if (precision == mfma::kF16) mfma::launch_single_prep<mfma::kF16>(s, d);
         else mfma::launch_single_prep<mfma::kBF16x3>(s, d);
      }
      if (pair_ok && (chosen == 0 || chosen > 10)) {
         const BatchDesc d{static_cast<const char *>(T.ptr), tw, th, T.step, KC2, nub, g0, g1,
                           tsums};
         if (precision == mfma::kBF16) mfma::launch_pair_prep<mfma::kBF16>(s, d);
         else if (precision == mfma::kF16) mfma::launch_pair_prep<mfma::kF16>(s, d);
         else mfma::launch_pair_prep<mfma::kBF16x3>(s, d);
      }
   }
   // staged columns start at x0 - tw/2 + 64 ub + 8 g with x0 a multiple of 64 (crosscorr_mfma)
   // or 16 (row-pair kernels): 16-byte aligned when tw/2 is a multiple of 4 and the rows are
   // 16-byte aligned
   const int vec = ((tw / 2) % 4 == 0 && rows_aligned(S)) ? 1 : 0;
   auto run = [&](int variant) {
      SingleLaunch L;
      L.mode = mfma::kModeCC;
      if (variant == 24) {
         const int lcols = mfma::pair_lcols(KC2, kStripNT);
         const int tiles_x = (S.width + 16 * kStripNT - 1) / (16 * kStripNT);
         const int tiles_y = (S.height + mfma::kPairRows - 1) / mfma::kPairRows;
         L.a = mfma::TileArgs{static_cast<const float *>(S.ptr), S.width, S.height, S.step, tw,
                              th, nub, g0, nullptr, tsums, integ, dst, dstep, 0, vec};
         L.KC = KC2;
         L.NT = kStripNT;
         L.ldsp = ring_ldsp(lcols);
         L.lds = strip_lds(KC2);
         // two workgroups per CU fit the LDS; OCIP_STRIP_WORKGROUPS sets another count
         int wgs = 2 * device_cus();
         if (const char *e = getenv("OCIP_STRIP_WORKGROUPS")) wgs = std::max(1, atoi(e));
         L.grid = dim3(std::min(tiles_x * tiles_y, wgs), tiles_x);
         if (precision == mfma::kBF16) mfma::launch_strip<mfma::kBF16>(s, L, tiles_y, strip_q_rows(KC2));
         else mfma::launch_strip<mfma::kF16>(s, L, tiles_y, strip_q_rows(KC2));
         return;
      }
      if (variant > 10) {
         const int NT = variant - 10;
         const Geometry g = pair_geometry(KC2, NT, th, mfma::kLdsMax, planes_for(precision));
         L.a = mfma::TileArgs{static_cast<const float *>(S.ptr), S.width, S.height, S.step, tw,
                              th, nub, g0, g1, tsums, integ, dst, dstep, g.vb, vec};
         L.KC = KC2;
         L.NT = NT;
         L.ldsp = g.ldsp;
         L.lds = g.lds;
         L.grid = dim3((S.width + 16 * NT - 1) / (16 * NT),
                       (S.height + mfma::kPairRows - 1) / mfma::kPairRows);
         if (precision == mfma::kBF16) mfma::launch_pair<mfma::kBF16>(s, L);
         else if (precision == mfma::kF16) mfma::launch_pair<mfma::kF16>(s, L);
         else mfma::launch_pair<mfma::kBF16x3>(s, L);
         return;
      }
      const int NT = variant;
      const Geometry g = geometry(precision, KC, NT, th);
      L.a = mfma::TileArgs{static_cast<const float *>(S.ptr), S.width, S.height, S.step, tw, th,
                           nub, f0, f1, tsums, integ, dst, dstep, g.vb, vec};
      L.KC = KC;
      L.NT = NT;
      L.ldsp = g.ldsp;
      L.lds = g.lds;
      L.grid = dim3((S.width + 32 * NT - 1) / (32 * NT),
                    (S.height + mfma::kTileRows - 1) / mfma::kTileRows);
      launch_single_p(precision, s, L);
   };
   const int variant = chosen ? chosen : tuned_variant(key, s, run, cands);
   {
      KernelTimer kt(s, "crosscorr_mfma");
      run(variant);
   }
   if (norm) launch_ncc_scale(s, S, tw, th, tsums, integ, dst, dstep);
}

// ======================================================================= many sources
// Split-BF16 CrossCorr maps (anchor (tw/2, th/2), zero outside the image, no Norm) of several
// sources of one size with one template: crosscorr_mfma (NT = 2 or 4) or the row-pair kernel
// (NT = 4 or 6), the autotuner's choice, the images of up to kZImages sources in one launch.
// The template's fragments, in both layouts, are prepared once by prep_crosscorr_split.
static_assert(kSplitImages == mfma::kZImages, "sources per launch");
size_t crosscorr_split_frag_bytes(int tw, int th)
{
   return mfma_frag_bytes(mfma::kBF16x3, tw, th) + pair_frag_bytes(mfma::kBF16x3, tw, th);
}

// rows per LDS band for the wider of the two tile widths the autotuner tries (NT = 4)
static Geometry split_geometry(int tw, int th, int NT = 4)
{
   const Geometry g = geometry(mfma::kBF16x3, kc_for(tw), NT, th);
   if (g.vb < 1) fail(CL_INVALID_VALUE, "template too large for the MFMA tile");
   return g;
}

void prep_crosscorr_split(hipStream_t s, const View &T, void *frags, double *sums)
{
   const int tw = T.width, th = T.height, KC = kc_for(tw), KC2 = kc2_for(tw), nub = nub_for(tw);
   (void)split_geometry(tw, th);
   u16x8 *f0 = static_cast<u16x8 *>(frags);
   const BatchDesc d{static_cast<const char *>(T.ptr), tw, th, T.step, KC, nub, f0,
                     f0 + (size_t)nub * th * KC * 64, sums};
   mfma::launch_single_prep<mfma::kBF16x3>(s, d);
   // the row-pair layout after it (both preps write the same sum of squares)
   u16x8 *g0 = reinterpret_cast<u16x8 *>(static_cast<char *>(frags) +
                                         mfma_frag_bytes(mfma::kBF16x3, tw, th));
   const BatchDesc e{static_cast<const char *>(T.ptr), tw, th, T.step, KC2, nub, g0,
                     g0 + pair_frag_count(tw, th) * 64, sums};
   mfma::launch_pair_prep<mfma::kBF16x3>(s, e);
}

void launch_crosscorr_split(hipStream_t s, int n, const float *const *src, const int *sstep,
                            float *const *dst, int dstep, int W, int H, int tw, int th,
                            const void *frags, const double *sums)
{
   const int KC = kc_for(tw), KC2 = kc2_for(tw), nub = nub_for(tw);
   const int vb = split_geometry(tw, th).vb;
   const u16x8 *f0 = static_cast<const u16x8 *>(frags);
   const u16x8 *f1 = f0 + (size_t)nub * th * KC * 64;
   const u16x8 *g0 = reinterpret_cast<const u16x8 *>(static_cast<const char *>(frags) +
                                                     mfma_frag_bytes(mfma::kBF16x3, tw, th));
   const u16x8 *g1 = g0 + pair_frag_count(tw, th) * 64;
   // one launch per kZImages sources: variant 2 or 4, crosscorr_mfma with NT = 2 or 4 column
   // tiles per workgroup; 14 or 16, the row-pair kernel with NT = 4 or 6
   auto run = [&](int variant) {
      const bool pair = variant > 10;
      const int NT = pair ? variant - 10 : variant;
      const Geometry g = pair ? pair_geometry(KC2, NT, th, mfma::kLdsMax, 2)
                              : split_geometry(tw, th, NT);
      if (g.vb < 1) fail(CL_INVALID_VALUE, "template too large for the MFMA tile");
      for (int i0 = 0; i0 < n; i0 += mfma::kZImages) {
         SingleLaunch L;
         L.z.n = std::min(mfma::kZImages, n - i0);
         // 16-byte staging loads: tw/2 a multiple of 4 and every source's rows 16-byte aligned
         int vec = (tw / 2) % 4 == 0;
         for (int k = 0; k < L.z.n; ++k) {
            L.z.src[k] = src[i0 + k];
            L.z.dst[k] = dst[i0 + k];
            L.z.sstep[k] = sstep[i0 + k];
            if (reinterpret_cast<uintptr_t>(src[i0 + k]) % 16 || sstep[i0 + k] % 16) vec = 0;
         }
         L.mode = mfma::kModeCC;
         L.NT = NT;
         L.ldsp = g.ldsp;
         L.lds = g.lds;
         if (pair) {
            L.a = mfma::TileArgs{src[i0], W, H, sstep[i0], tw, th, nub, g0, g1, sums, nullptr,
                                 dst[i0], dstep, g.vb, vec};
            L.KC = KC2;
            L.grid = dim3((W + 16 * NT - 1) / (16 * NT), (H + mfma::kPairRows - 1) / mfma::kPairRows);
            mfma::launch_pair<mfma::kBF16x3>(s, L);
         } else {
            L.a = mfma::TileArgs{src[i0], W, H, sstep[i0], tw, th, nub, f0, f1, sums, nullptr,
                                 dst[i0], dstep, vb, vec};
            L.KC = KC;
            L.grid = dim3((W + 32 * NT - 1) / (32 * NT), (H + mfma::kTileRows - 1) / mfma::kTileRows);
            mfma::launch_single<mfma::kBF16x3>(s, L);
         }
      }
   };
   // the autotuner's choice for this shape and number of sources (each candidate rewrites the
   // same outputs)
   char shape[64];
   snprintf(shape, sizeof shape, "tw%d:th%d:n%d", tw, th, std::min(n, (int)mfma::kZImages));
   const View v0{src[0], W, H, sstep[0], 1, SImage::F32};
   const int variant = tuned_variant(tune_key("split", mfma::kBF16x3, mfma::kModeCC, shape, v0), s,
                                     run, {16, 14, 4, 2});
   KernelTimer kt(s, "crosscorr_mfma");
   run(variant);
}

// ======================================================================= exact integer sums
// 8-bit integer images through the row-pair kernel (kModeInt): the exact cross term of
// CrossCorr, CrossCorr_Norm, SqrDistance and SqrDistance_Norm on BF16 MFMA, and the window
// sums of s^2 from ss_box8.  Limits: templates up to 97 x 97 (one template column block, and
// the rows ss_box8 stages), one row's products below 2^24 (the flush period is at least one
// step) and the whole sum below 2^31 (int32), the LDS tiles.  OCIP_INT_MFMA=0 turns the
// path off (the VALU kernel prox_tile_int then computes the same maps).
static int int_max_product(int type) { return type == SImage::U8 ? 255 * 255 : 128 * 128; }

// Window sums of s^2 of an 8-bit image, exact in int32, into the map `dst` (as int32 bits):
// SS(x, y) = sum over the template window of (x, y) (origin (-tw/2, -th/2), zero outside the
// image).  One wavefront per 64 output columns and a band of kBoxBand output rows, streaming
// down the band's th - 1 + kBoxBand source rows.  For each row:
//  * the squares q of its 63 + tw columns from x0 - tw/2, two adjacent columns per lane (NCH
//    chunks of 128 columns; zero outside the image through the range of a buffer resource);
//  * their exclusive prefix along the row, P[i] = sum of q over columns < i: a wavefront
//    inclusive scan of the lanes' pair sums (DPP row shifts and row broadcasts), the chunk
//    carry through readlane; integers, exact;
//  * for output column x0 + x (lane x from here on), R(x) = P[x + tw] - P[x], both prefixes
//    fetched from the lanes that hold them with ds_bpermute;
//  * a per-lane ring of the last th rows' R in LDS: V += R - R(th rows up); from the th-th
//    row on, V is the sum of the output row th - 1 rows up, stored as one 256-byte row segment.
// Integer arithmetic throughout (every value below 2^31, checked by the host), so the sums are
// exact in any order.
constexpr int kBoxBand = 64;
constexpr int kBoxOob = 0x40000000;   // a byte offset past any image this kernel takes

// Inclusive scans across the 64 lanes of N values at once (the N scans interleaved, so that
// each DPP step's wait on the previous one is filled by the others): within rows of 16 lanes
// (lanes without a source add 0), then rows 1 and 3 take the last lane of rows 0 and 2, rows
// 2 and 3 the last lane of row 1.
template <int N>
__device__ __forceinline__ void wave_incl_scan(int (&x)[N])
{
#pragma unroll
   for (int k = 0; k < N; ++k) x[k] += __builtin_amdgcn_update_dpp(0, x[k], 0x111, 0xf, 0xf, false);
#pragma unroll
   for (int k = 0; k < N; ++k) x[k] += __builtin_amdgcn_update_dpp(0, x[k], 0x112, 0xf, 0xf, false);
#pragma unroll
   for (int k = 0; k < N; ++k) x[k] += __builtin_amdgcn_update_dpp(0, x[k], 0x114, 0xf, 0xf, false);
#pragma unroll
   for (int k = 0; k < N; ++k) x[k] += __builtin_amdgcn_update_dpp(0, x[k], 0x118, 0xf, 0xf, false);
#pragma unroll
   for (int k = 0; k < N; ++k) x[k] += __builtin_amdgcn_update_dpp(0, x[k], 0x142, 0xa, 0xf, false);
#pragma unroll
   for (int k = 0; k < N; ++k) x[k] += __builtin_amdgcn_update_dpp(0, x[k], 0x143, 0xc, 0xf, false);
}

template <int NCH, bool SIGNED, bool BST>
__global__ __launch_bounds__(64) void ss_box8(const unsigned char *__restrict__ src, int W,
                                              int H, int sstep, int tw, int th,
                                              char *__restrict__ dst, int dstep, int bands)
{
   extern __shared__ int ring[];   // th x 64: lane's R of the last th rows
   constexpr int U = 8;            // rows per batch of loads (two batches in flight)
   const int lane = threadIdx.x;
   // consecutive blocks (one XCD's run) side by side along a band share halo columns in L2
   const int b = xcd_tile(blockIdx.x, gridDim.x);
   const int cblks = gridDim.x / bands;
   const int band = b / cblks, cblk = b % cblks;
   const int x0 = cblk * 64, y0 = band * kBoxBand;
   const int cx = tw / 2, cy = th / 2;
   const int n = min(kBoxBand, H - y0);
   const int steps = n + th - 1;
   // lane holds columns x0 - cx + 128 k + 2 lane + e (e = 0, 1) of chunk k
   int coff[NCH][2];
#pragma unroll
   for (int k = 0; k < NCH; ++k)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
         const int c = x0 - cx + 128 * k + 2 * lane + e;
         coff[k][e] = c >= 0 && c < W ? c : kBoxOob;
      }
   const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
       const_cast<unsigned char *>(src), (short)0, H * sstep, 0x00020000);
   // (bytes kept as loaded until used, so that the loop carries no conversion that would wait
   // for the loads in flight).  A row outside the image takes the row offset 2 kBoxOob, a
   // column outside it the column offset kBoxOob: every sum with either is past the range (and
   // below 2^32); rows past the band's last step load real rows whose sums are never used.
   auto load = [&](int j0, unsigned char (&v)[U][NCH][2]) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
         const int sy = y0 - cy + j0 + u;
         const unsigned roff = (unsigned)sy < (unsigned)H ? (unsigned)(sy * sstep) : 2u * kBoxOob;
#pragma unroll
         for (int k = 0; k < NCH; ++k)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
               const unsigned o = roff + (unsigned)coff[k][e];
               v[u][k][e] = __builtin_amdgcn_raw_buffer_load_b8(rsrc, (int)o, 0, 0);
            }
      }
   };
   // P[x] at lane x / 2 of chunk 0 (element x % 2); P[x + tw] at lane ((x + tw) % 128) / 2
   // of chunk (x + tw) / 128 (element (x + tw) % 2)
   const int ia = lane, ib = lane + tw;
   const int la = (ia >> 1) * 4, lb = ((ib & 127) >> 1) * 4;
   const bool oa = ia & 1, ob = ib & 1, cb = ib >= 128;
   int *const myring = ring + lane;
   int V = 0, slot = 0;
   // BST (maps of fewer than kBoxOob bytes): buffer stores, the column offset kBoxOob past
   // the image's width dropping a lane's store; otherwise stores through a pointer
   int *orow = reinterpret_cast<int *>(dst + (size_t)y0 * dstep) + x0 + lane;
   const bool xin = x0 + lane < W;
   const __amdgpu_buffer_rsrc_t drsrc =
       __builtin_amdgcn_make_buffer_rsrc(dst, (short)0, BST ? H * dstep : 0, 0x00020000);
   const int xoff = xin ? (x0 + lane) * 4 : kBoxOob;
   int yoff = y0 * dstep;
   unsigned char va[U][NCH][2], vb[U][NCH][2];
   auto process = [&](int j0, const unsigned char (&v)[U][NCH][2]) {
      // the batch's row sums first (independent rows: their scans and lane exchanges
      // interleave, two rows at a time; rows past the band read real rows or zeros, and their
      // sums are never used), then the ring and the running sums in order
      int R[U];
#pragma unroll
      for (int u = 0; u < U; u += 2) {
         int q0[2 * NCH], t[2 * NCH], inc[2 * NCH];
#pragma unroll
         for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int k = 0; k < NCH; ++k) {
               int e0 = v[u + h][k][0], e1 = v[u + h][k][1];
               if (SIGNED) {
                  e0 = (int)(signed char)e0;
                  e1 = (int)(signed char)e1;
               }
               q0[h * NCH + k] = e0 * e0;
               t[h * NCH + k] = inc[h * NCH + k] = e0 * e0 + e1 * e1;
            }
         wave_incl_scan<2 * NCH>(inc);
#pragma unroll
         for (int h = 0; h < 2; ++h) {
            int E0[NCH], E1[NCH];
            int carry = 0;
#pragma unroll
            for (int k = 0; k < NCH; ++k) {
               const int i = inc[h * NCH + k] + carry;
               E0[k] = i - t[h * NCH + k];
               E1[k] = E0[k] + q0[h * NCH + k];
               carry = __builtin_amdgcn_readlane(i, 63);
            }
            const int a0 = __builtin_amdgcn_ds_bpermute(la, E0[0]);
            const int a1 = __builtin_amdgcn_ds_bpermute(la, E1[0]);
            int b0 = __builtin_amdgcn_ds_bpermute(lb, E0[0]);
            int b1 = __builtin_amdgcn_ds_bpermute(lb, E1[0]);
            if constexpr (NCH == 2) {
               // P[x + tw] in chunk 1 for the lanes with x + tw >= 128
               const int c0 = __builtin_amdgcn_ds_bpermute(lb, E0[1]);
               const int c1 = __builtin_amdgcn_ds_bpermute(lb, E1[1]);
               b0 = cb ? c0 : b0;
               b1 = cb ? c1 : b1;
            }
            R[u + h] = (ob ? b1 : b0) - (oa ? a1 : a0);
         }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
         const int j = j0 + u;
         if (j >= steps) break;
         const int old = myring[slot * 64];
         myring[slot * 64] = R[u];
         V += R[u] - (j >= th ? old : 0);
         if (++slot == th) slot = 0;
         if (j >= th - 1) {
            if (BST) {
               __builtin_amdgcn_raw_buffer_store_b32(V, drsrc, (int)((unsigned)yoff + (unsigned)xoff), 0, 0);
               yoff += dstep;
            } else {
               if (xin) *orow = V;
               orow = reinterpret_cast<int *>(reinterpret_cast<char *>(orow) + dstep);
            }
         }
      }
   };
   load(0, va);
   for (int j0 = 0; j0 < steps; j0 += 2 * U) {
      load(j0 + U, vb);
      process(j0, va);
      load(j0 + 2 * U, va);
      process(j0 + U, vb);
   }
}

// chunks of 128 columns: P up to index 63 + tw (the 63 + tw staged columns and one past)
static int ss_box8_chunks(int tw) { return 63 + tw < 128 ? 1 : 2; }

bool int_mfma_ok(int op, const View &S, const View &T)
{
   const char *e = getenv("OCIP_INT_MFMA");
   if ((e && e[0] == '0') || op == kSAD) return false;
   if ((S.type != SImage::U8 && S.type != SImage::S8) || S.channels != 1 || T.type != S.type)
      return false;
   const int tw = T.width, th = T.height;
   // (templates of one column block; ss_box8: byte offsets of the image below kBoxOob)
   if (tw <= 0 || th <= 0 || tw > 97 || th > 97) return false;
   if ((long long)S.height * S.step >= kBoxOob) return false;
   const long long mp = int_max_product(S.type);
   if ((long long)tw * mp >= (1LL << 24) || (long long)tw * th * mp >= (1LL << 31)) return false;
   return pair_geometry(kc2_for(tw), 4, th).vb >= 1;
}

size_t int_mfma_work_bytes(int tw, int th) { return pair_frag_bytes(mfma::kBF16, tw, th); }

void launch_int_mfma(hipStream_t s, int op, const View &S, const View &T, double *tsums,
                     float *dst, int dstep, void *work)
{
   const int tw = T.width, th = T.height, KC2 = kc2_for(tw);
   const int src8 = S.type == SImage::U8 ? mfma::kSrcU8 : mfma::kSrcS8;
   if (op != kCC) {
      // the window sums of s^2 go to the map first; the MFMA kernel's epilogue reads them
      KernelTimer kt(s, "ss_box8");
      const int bands = (S.height + kBoxBand - 1) / kBoxBand;
      const int cblks = (S.width + 63) / 64;
      const size_t lds = (size_t)th * 64 * sizeof(int);   // th <= 97: below 64 KiB
      auto go = [&](auto kernel) {
         kernel<<<bands * cblks, 64, lds, s>>>(static_cast<const unsigned char *>(S.ptr), S.width,
                                               S.height, S.step, tw, th,
                                               reinterpret_cast<char *>(dst), dstep, bands);
      };
      const bool sgn = src8 == mfma::kSrcS8;
      const bool bst = (long long)S.height * dstep < kBoxOob;
#define OCIP_BOX_CASE(NCH)                                                                      \
   if (sgn) { if (bst) go(ss_box8<NCH, true, true>); else go(ss_box8<NCH, true, false>); }      \
   else { if (bst) go(ss_box8<NCH, false, true>); else go(ss_box8<NCH, false, false>); }
      if (ss_box8_chunks(tw) == 1) {
         OCIP_BOX_CASE(1)
      } else {
         OCIP_BOX_CASE(2)
      }
#undef OCIP_BOX_CASE
      launch_check("ss_box8");
   }
   u16x8 *g0 = static_cast<u16x8 *>(work);
   {
      KernelTimer kt(s, "template_prep");
      BatchDesc d{static_cast<const char *>(T.ptr), tw, th, T.step, KC2, 1, g0, nullptr, tsums};
      d.src8 = src8;
      mfma::launch_pair_prep<mfma::kBF16>(s, d);
   }
   const Geometry g = pair_geometry(KC2, 4, th);
   SingleLaunch L;
   L.mode = mfma::kModeInt;
   // 8-byte staging loads: staged columns start at 64 bx - tw/2 + 8 g, so tw/2 a multiple of 8,
   // and 8-byte aligned rows
   const int vec8 = (tw / 2) % 8 == 0 && S.step % 8 == 0 &&
                    reinterpret_cast<uintptr_t>(S.ptr) % 8 == 0;
   L.a = mfma::TileArgs{static_cast<const float *>(S.ptr), S.width, S.height, S.step, tw, th, 1,
                        g0, nullptr, tsums, nullptr, dst, dstep, g.vb, vec8};
   L.a.src8 = src8;
   L.a.op = op;
   // steps between the int32 flushes: each step adds one template row of products
   L.a.flush = (int)(((1LL << 24) - 1) / ((long long)tw * int_max_product(S.type)));
   L.KC = KC2;
   L.NT = 4;
   L.ldsp = g.ldsp;
   L.lds = g.lds;
   L.grid = dim3((S.width + 63) / 64, (S.height + mfma::kPairRows - 1) / mfma::kPairRows);
   KernelTimer kt(s, "crosscorr_mfma_int");
   mfma::launch_pair<mfma::kBF16>(s, L);
}

// ======================================================================= batch
struct BatchEntry {
   hipGraphExec_t exec = nullptr;
   DevBuf frags, sums, descs, items, counters, integ, aux;
   ~BatchEntry()
   {
      if (exec) (void)hipGraphExecDestroy(exec);
      frags.release();
      sums.release();
      descs.release();
      items.release();
      counters.release();
      integ.release();
      aux.release();
   }
};

struct BatchCache {
   std::map<std::string, BatchEntry *> entries;
   std::vector<std::string> order;  // insertion order, for eviction
};

BatchCache *batch_cache_create() { return new BatchCache(); }

void batch_cache_destroy(BatchCache *c)
{
   if (!c) return;
   for (auto &kv : c->entries) delete kv.second;
   delete c;
}

void run_batch_mfma(hipStream_t s, BatchCache *cache, int num_cus, int precision, bool ssd,
                    const View &S, const std::vector<View> &T, const std::vector<float *> &dst,
                    const std::vector<int> &dstep)
{
   const int n = (int)T.size();
   const int mode = ssd ? mfma::kModeSSDN : mfma::kModeNCC;
   // the cache key is everything the recorded graph depends on
   std::string key;
   {
      char buf[256];
      snprintf(buf, sizeof buf, "p%d m%d S%p %dx%d s%d |", precision, mode, S.ptr, S.width,
               S.height, S.step);
      key = buf;
      for (int i = 0; i < n; ++i) {
         snprintf(buf, sizeof buf, "%p %dx%d s%d %p s%d|", T[i].ptr, T[i].width, T[i].height,
                  T[i].step, (void *)dst[i], dstep[i]);
         key += buf;
      }
   }
   auto hit = cache->entries.find(key);
   if (hit != cache->entries.end()) {
      KernelTimer kt(s, "crosscorr_mfma_batch");
      hip_check(hipGraphLaunch(hit->second->exec, s), "graph launch");
      return;
   }

   // BF16 and FP16 run the row-pair tile (crosscorr_mfma_pair_batch; KC = kc2_for(tw) chunks,
   // a step = two template rows), split BF16 the tile of crosscorr_mfma (KC = kc_for(tw))
   const bool pair = precision != mfma::kBF16x3;
   auto kc_of = [&](int tw) { return pair ? kc2_for(tw) : kc_for(tw); };
   auto nfrag_of = [&](const View &t) {
      const int nub = nub_for(t.width);
      return (size_t)nub * (pair ? mfma::pair_frag_steps(t.height, nub) : t.height) * kc_of(t.width);
   };
   // the row-pair batch's LDS per workgroup: at most half a CU's, two workgroups per CU
   const long pair_cap = mfma::kLdsMax / 2;

   // ---- build: buckets by KC, templates in each bucket by decreasing cost
   std::map<int, std::vector<int>> buckets;
   for (int i = 0; i < n; ++i) buckets[kc_of(T[i].width)].push_back(i);
   for (auto &b : buckets)
      std::stable_sort(b.second.begin(), b.second.end(), [&](int a, int c) {
         return (long)T[a].height * nub_for(T[a].width) > (long)T[c].height * nub_for(T[c].width);
      });
   // one work-queue counter per bucket (and per XCD band on the row-pair tile)
   const int counters_per_bucket = pair ? mfma::kPairBatchBands : 1;
   if ((int)buckets.size() * counters_per_bucket > kBatchCounters)
      fail(CL_INVALID_VALUE, "too many template widths");
   const int planes = pair ? 1 : planes_for(precision);
   std::vector<size_t> frag_off(n);
   size_t nfrag_total = 0;
   for (int i = 0; i < n; ++i) {
      frag_off[i] = nfrag_total;
      nfrag_total += nfrag_of(T[i]);
   }
   BatchEntry *e = new BatchEntry();
   try {
      e->frags.ensure(planes * nfrag_total * 64 * sizeof(u16x8));
      e->sums.ensure((size_t)n * 2 * sizeof(double));
      e->descs.ensure((size_t)n * sizeof(BatchDesc));
      e->items.ensure((size_t)n * sizeof(BatchItem));
      e->counters.ensure(kBatchCounters * sizeof(unsigned));
      e->integ.ensure(excl_bytes(S.width, S.height));
      e->aux.ensure(integral_aux_bytes(S.width, S.height));

      u16x8 *f0 = e->frags.as<u16x8>();
      u16x8 *f1 = planes == 2 ? f0 + nfrag_total * 64 : nullptr;
      double *sums = e->sums.as<double>();
      std::vector<BatchDesc> descs(n);
      for (int i = 0; i < n; ++i)
         descs[i] = BatchDesc{static_cast<const char *>(T[i].ptr), T[i].width, T[i].height,
                              T[i].step, kc_of(T[i].width), nub_for(T[i].width),
                              f0 + frag_off[i] * 64, f1 ? f1 + frag_off[i] * 64 : nullptr,
                              sums + 2 * i};
      hip_check(hipMemcpy(e->descs.ptr, descs.data(), n * sizeof(BatchDesc),
                          hipMemcpyHostToDevice), "batch table");

      // per bucket: its items (in bucket order) and the rows per LDS band
      struct Bucket {
         int KC, first, count, th_max;
      };
      std::vector<Bucket> bl;
      std::vector<BatchItem> items;
      for (auto &b : buckets) {
         Bucket k{b.first, (int)items.size(), (int)b.second.size(), 0};
         for (int i : b.second) k.th_max = std::max(k.th_max, T[i].height);
         // rows (pair: steps) per LDS band for the widest tile a bucket may run with (NT = 4,
         // pair: 6), so that every tile width the autotuner tries fits
         const Geometry g = pair ? pair_geometry(k.KC, 6, k.th_max, pair_cap)
                                 : geometry(precision, k.KC, 4, k.th_max);
         if (g.vb < 1) fail(CL_INVALID_VALUE, "template too large for the MFMA tile");
         for (int i : b.second) {
            const int depth = pair ? (T[i].height + 1) / 2 : T[i].height;
            items.push_back(BatchItem{T[i].width, T[i].height, nub_for(T[i].width),
                                      std::min(depth, g.vb), descs[i].frag0,
                                      descs[i].frag1, sums + 2 * i, dst[i], dstep[i]});
         }
         bl.push_back(k);
      }
      hip_check(hipMemcpy(e->items.ptr, items.data(), n * sizeof(BatchItem),
                          hipMemcpyHostToDevice), "batch table");

      unsigned *counters = e->counters.as<unsigned>();
      const BatchItem *ditems = e->items.as<BatchItem>();
      auto bucket_launch = [&](size_t b, int NT) {
         const Bucket &k = bl[b];
         const Geometry g = pair ? pair_geometry(k.KC, NT, k.th_max, pair_cap)
                                 : geometry(precision, k.KC, NT, k.th_max);
         BatchLaunch L;
         L.src = static_cast<const float *>(S.ptr);
         L.W = S.width;
         L.H = S.height;
         L.sstep = S.step;
         L.vec = rows_aligned(S) ? 1 : 0;
         L.integ = e->integ.as<double>();
         L.items = ditems + k.first;
         L.nitems = k.count;
         const int tcols = pair ? 16 * NT : 32 * NT, trows = pair ? mfma::kPairRows : mfma::kTileRows;
         L.tiles_x = (S.width + tcols - 1) / tcols;
         L.tiles_y = (S.height + trows - 1) / trows;
         L.counter = counters + b * counters_per_bucket;
         L.KC = k.KC;
         L.NT = NT;
         L.mode = mode;
         L.ldsp = g.ldsp;
         L.lds = g.lds;
         const int per_cu = std::max(1, std::min(2, (int)(mfma::kLdsMax / std::max<size_t>(g.lds, 1))));
         L.grid = num_cus * per_cu;
         if (pair) launch_pair_batch_p(precision, s, L);
         else launch_batch_p(precision, s, L);
      };
      auto prep = [&]() {
         launch_integral(s, S, 0, true, e->integ.ptr, 0, -1, e->aux.ptr);
         if (pair) launch_pair_batch_prep_p(precision, s, e->descs.as<BatchDesc>(), n);
         else launch_batch_prep_p(precision, s, e->descs.as<BatchDesc>(), n);
      };
      auto reset = [&]() {
         reset_counters<<<1, kBatchCounters, 0, s>>>(counters, (int)bl.size() * counters_per_bucket);
         launch_check("reset_counters");
      };

      // eager pass: integral, fragments, and the autotuner's choice per bucket
      std::vector<int> nts;
      {
         KernelTimer kt(s, "crosscorr_mfma_batch");
         prep();
         for (size_t b = 0; b < bl.size(); ++b) {
            auto run = [&](int NT) {
               reset();
               bucket_launch(b, NT);
            };
            char shape[48];
            snprintf(shape, sizeof shape, "kc%d:thmax%d", bl[b].KC, bl[b].th_max);
            if (pair)
               nts.push_back(tuned_variant(tune_key("pair_batch", precision, mode, shape, S), s, run, {6, 4}));
            else
               nts.push_back(tuned_variant(tune_key("batch", precision, mode, shape, S), s, run, {4, 2}));
         }
      }
      // record the whole batch once
      hipGraph_t graph = nullptr;
      timing_suspend(true);
      hipError_t err = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
      if (err == hipSuccess) {
         try {
            prep();
            reset();
            for (size_t b = 0; b < bl.size(); ++b) bucket_launch(b, nts[b]);
         } catch (...) {
            hipGraph_t g2 = nullptr;
            (void)hipStreamEndCapture(s, &g2);
            if (g2) (void)hipGraphDestroy(g2);
            timing_suspend(false);
            throw;
         }
         err = hipStreamEndCapture(s, &graph);
      }
      timing_suspend(false);
      hip_check(err, "graph capture");
      err = hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      hip_check(err, "graph instantiate");
   } catch (...) {
      delete e;
      throw;
   }
   // bounded cache
   if (cache->order.size() >= 16) {
      (void)hipStreamSynchronize(s);
      delete cache->entries[cache->order.front()];
      cache->entries.erase(cache->order.front());
      cache->order.erase(cache->order.begin());
   }
   cache->entries[key] = e;
   cache->order.push_back(key);
   KernelTimer kt(s, "crosscorr_mfma_batch");
   hip_check(hipGraphLaunch(e->exec, s), "graph launch");
}

}  // namespace ocip

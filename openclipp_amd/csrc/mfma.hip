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
__global__ void reset_counters(unsigned *c, int n)
{
   if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}

// 1 / sqrt(SS TT) for every output of a CrossCorr_Norm map, straight from the source image:
// SS(x, y) = sum of s^2 over the template window of (x, y) (origin (-tw/2, -th/2), zero
// outside the image), TT = sum t^2 in FP64, the reciprocal square root in FP32 (inv_sqrt).
// One wavefront owns kDenCols output columns and a band of kDenBand output rows:
//  * vertical pass, lane = column of the window segment [x0 - tw/2, x0 - tw/2 + kDenCols - 1
//    + tw): two FP64 running sums of s^2 from the band's first row, `lead` taking each row as
//    it enters the window and `lag` as it leaves, so that the column sum V = lead - lag is the
//    difference of two prefixes of one sequence (an all-zero column gives exactly 0).  V goes
//    to LDS as FP32 (relative error 2^-24), kDenRows output rows at a time;
//  * horizontal pass, lane = (output row, half of the columns): the same pair of running sums
//    along the row of V gives SS; the results go back through LDS and leave as coalesced rows.
// The rows are a stream of steps: step j takes row y0 - th/2 + j into `lead` (j >= 0), row
// y0 - th/2 + j - th into `lag` (j >= th) and completes output row y0 + j - (th - 1).  Steps
// run in batches of U, the loads of batch b + 1 in flight while batch b is summed; the
// stream starts at jstart <= 0 (steps before 0 load nothing) so that output rows fill the
// LDS block at batch boundaries.  No integral is formed: the source is read from HBM about
// once (halo columns and the lag rows come from L2) and the map is written once.
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

template <int NCH, bool BUF>
__global__ __launch_bounds__(64) void ncc_den_src(const char *__restrict__ src, int sstep, int W,
                                                  int H, int tw, int th,
                                                  const double *__restrict__ tsums,
                                                  float *__restrict__ out, int bands)
{
   extern __shared__ float vbuf[];  // kDenRows x P floats; P odd: conflict-free row walks
   constexpr int P = 64 * NCH + 1;
   constexpr int U = NCH <= 2 ? 16 : (NCH <= 3 ? 8 : 4);   // steps per batch (divides kDenRows)
   constexpr int HALF = kDenCols / 2;           // outputs per lane, horizontal pass
   constexpr int BPB = kDenRows / U;            // output batches per LDS block
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
   const __amdgpu_buffer_rsrc_t rsrc =
       __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(src), (short)0, BUF ? H * sstep : 0, 0x00020000);
   const int pre = th - 1;                       // steps before the first output row
   const int jstart = -((U - pre % U) % U);      // (pre - jstart) % U == 0
   const int b0 = (pre - jstart) / U;            // first batch that completes output rows
   const int nb = b0 + (n + U - 1) / U;

   // Every load is unconditional, so a whole batch is in flight at once.  BUF: a buffer load
   // whose offset (row + column, either kDenOob when outside) is out of range returns 0.
   // Otherwise a position outside the image (or a step that takes no row) reads the zero of
   // kZeroPad (an address select, no branch around the load).
   auto load_batch = [&](int bb, float (&L)[U][NCH], float (&G)[U][NCH]) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
         const int j = jstart + bb * U + u;
         const bool live = j - pre < n;
         const int lr = y0 - cy + j, gr = lr - th;
         const bool lok = live && j >= 0 && lr >= 0 && lr < H;
         const bool gok = live && j >= th && gr >= 0 && gr < H;
         if (BUF) {
            const int loff = lok ? lr * sstep : kDenOob, goff = gok ? gr * sstep : kDenOob;
#pragma unroll
            for (int k = 0; k < NCH; ++k) {
               const unsigned lo = (unsigned)coff[k] + (unsigned)loff, go = (unsigned)coff[k] + (unsigned)goff;
               L[u][k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)lo, 0, 0));
               G[u][k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)go, 0, 0));
            }
         } else {
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

   // horizontal pass and stores of the nrows rows of the LDS block (output rows ybase..)
   auto flush = [&](int ybase, int nrows) {
      __syncthreads();
      const int r = lane & 31, xs = (lane >> 5) * HALF;
      float res[HALF];
      if (r < nrows) {
         const float *row = vbuf + r * P + xs;
         double hl = 0.0, hg = 0.0;
         for (int i = 0; i < tw - 1; ++i) hl += (double)row[i];
#pragma unroll
         for (int x = 0; x < HALF; ++x) {
            hl += (double)row[x + tw - 1];
            const double ss = hl - hg;
            hg += (double)row[x];
            res[x] = inv_sqrt(ss * tt);
         }
      }
      __syncthreads();
      if (r < nrows) {
#pragma unroll
         for (int x = 0; x < HALF; ++x) vbuf[r * P + xs + x] = res[x];
      }
      __syncthreads();
#pragma unroll
      for (int h = 0; h < kDenCols / 64; ++h) {
         const int x = x0 + 64 * h + lane;
         if (x < W)
            for (int rr = 0; rr < nrows; ++rr)
               out[(size_t)(ybase + rr) * W + x] = vbuf[rr * P + 64 * h + lane];
      }
      __syncthreads();
   };

   auto process = [&](int bb, const float (&L)[U][NCH], const float (&G)[U][NCH]) {
      const int e0 = (bb - b0) * U;             // output row (band-relative) of step u = 0
#pragma unroll
      for (int u = 0; u < U; ++u) {
#pragma unroll
         for (int k = 0; k < NCH; ++k) {
            lead[k] = fma((double)L[u][k], (double)L[u][k], lead[k]);
            lag[k] = fma((double)G[u][k], (double)G[u][k], lag[k]);
         }
         if (bb >= b0 && e0 + u < n) {
#pragma unroll
            for (int k = 0; k < NCH; ++k)
               vbuf[((e0 + u) % kDenRows) * P + 64 * k + lane] = (float)(lead[k] - lag[k]);
         }
      }
      if (bb >= b0) {
         const int done = min(e0 + U, n);        // output rows completed so far
         if ((bb - b0 + 1) % BPB == 0 || done == n) {
            const int first = (done - 1) / kDenRows * kDenRows;
            flush(y0 + first, done - first);
         }
      }
   };

   float La[U][NCH], Ga[U][NCH], Lb[U][NCH], Gb[U][NCH];
   load_batch(0, La, Ga);
   for (int bb = 0; bb < nb; bb += 2) {
      if (bb + 1 < nb) load_batch(bb + 1, Lb, Gb);
      process(bb, La, Ga);
      if (bb + 1 >= nb) break;
      if (bb + 2 < nb) load_batch(bb + 2, La, Ga);
      process(bb + 1, Lb, Gb);
   }
}

// Wider templates (tw > 64 * kDenMaxChunks - 63): one output per thread, four corners of the
// FP64 integral each.
__global__ __launch_bounds__(256) void ncc_inv_den_direct(const double *__restrict__ integ, int W,
                                                          int H, int tw, int th,
                                                          const double *__restrict__ tsums,
                                                          float *__restrict__ out)
{
   const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
   if (k >= (long long)W * H) return;
   const int y = (int)(k / W), x = (int)(k - (long long)y * W);
   out[k] = inv_sqrt(window_sqsum(integ, W, H, x, y, tw, th) * tsums[0]);
}

int kc_for(int tw) { return tw <= 97 ? std::max(2, (tw + 31 + 15) / 16) : 6; }
int nub_for(int tw) { return tw <= 97 ? 1 : (tw + mfma::kBlockW - 1) / mfma::kBlockW; }
int planes_for(int precision) { return precision == mfma::kBF16x3 ? 2 : 1; }

struct Geometry {
   int ldsp, vb;
   size_t lds;
};

// LDS rows of lcols_for(KC, NT) 16-bit values, pitch shifted by 16 bytes per row modulo 128
// so that the 32 rows read by one A fragment start on different banks; as many template rows
// per band as fit in 160 KiB.
Geometry geometry(int precision, int KC, int NT, int th_max)
{
   Geometry g;
   const int lcols = mfma::lcols_for(KC, NT);
   g.ldsp = (lcols + 7) / 8 * 8;
   while ((g.ldsp * 2) % 128 != 16) g.ldsp += 8;
   const long rows_fit = (long)mfma::kLdsMax / ((long)planes_for(precision) * g.ldsp * 2);
   g.vb = (int)std::min<long>(th_max, rows_fit - (mfma::kTileRows - 1));
   g.lds = g.vb >= 1 ? (size_t)planes_for(precision) * (mfma::kTileRows + g.vb - 1) * g.ldsp * 2 : 0;
   return g;
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
// Autotuner: the number of 32-column n-tiles per wave (NT = 2 or 4, i.e. 64- or 128-column
// output tiles) for each (kind, precision, KC, mode, image size, template height).  The
// first call with a new key times both variants with HIP events on the call's own inputs
// and keeps the faster; the table is kept for the process and, when the environment
// variable OCIP_AUTOTUNE_CACHE names a file, read from and appended to that file.
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
               if (nt == 2 || nt == 4) table[key] = nt;
            fclose(f);
         }
      }
   }
   return table;
}

int tuned_nt(const std::string &key, hipStream_t s, const std::function<void(int)> &run)
{
   auto &table = tune_table();
   auto it = table.find(key);
   if (it != table.end()) return it->second;
   hipEvent_t e0, e1;
   hip_check(hipEventCreate(&e0), "event");
   hip_check(hipEventCreate(&e1), "event");
   float best_ms = 1e30f;
   int best = 4;
   for (int nt : {4, 2}) {
      run(nt);  // warm
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

std::string tune_key(const char *kind, int precision, int KC, int mode, int W, int H, int th)
{
   char buf[160];
   snprintf(buf, sizeof buf, "%s:p%d:kc%d:m%d:%dx%d:th%d", kind, precision, KC, mode, W, H, th);
   return buf;
}

}  // namespace

// ======================================================================= single template
static size_t mfma_frag_bytes(int precision, int tw, int th)
{
   const size_t b = (size_t)planes_for(precision) * nub_for(tw) * th * kc_for(tw) * 64 * sizeof(u16x8);
   return (b + 255) / 256 * 256;
}

// fragments, then the W x H map of 1 / sqrt(SS TT) of the Norm
bool crosscorr_mfma_needs_integral(int tw) { return den_chunks(tw) > kDenMaxChunks; }

size_t crosscorr_mfma_work_bytes(int precision, int W, int H, int tw, int th)
{
   return mfma_frag_bytes(precision, tw, th) + (size_t)W * H * sizeof(float);
}

void launch_crosscorr_mfma(hipStream_t s, int precision, bool norm, const View &S, const View &T,
                           double *tsums, const double *integ, float *dst, int dstep, void *work)
{
   const int tw = T.width, th = T.height;
   if (tw <= 0 || th <= 0) fail(CL_INVALID_VALUE, "empty template");
   const int KC = kc_for(tw), nub = nub_for(tw);
   if (geometry(precision, KC, 4, th).vb < 1)
      fail(CL_INVALID_VALUE, "template too large for the MFMA tile");
   const size_t nfrag = (size_t)nub * th * KC;
   u16x8 *f0 = static_cast<u16x8 *>(work);
   u16x8 *f1 = precision == mfma::kBF16x3 ? f0 + nfrag * 64 : nullptr;
   {
      KernelTimer kt(s, "template_prep");
      const BatchDesc d{static_cast<const char *>(T.ptr), tw, th, T.step, KC, nub, f0, f1, tsums};
      if (precision == mfma::kBF16) mfma::launch_single_prep<mfma::kBF16>(s, d);
      else if (precision == mfma::kF16) mfma::launch_single_prep<mfma::kF16>(s, d);
      else mfma::launch_single_prep<mfma::kBF16x3>(s, d);
   }
   const int mode = norm ? mfma::kModeNCCMap : mfma::kModeCC;
   float *invden = nullptr;
   if (norm) {
      invden = reinterpret_cast<float *>(static_cast<char *>(work) + mfma_frag_bytes(precision, tw, th));
      KernelTimer kt(s, "ncc_inv_den");
      const int nch = den_chunks(tw);
      if (nch <= kDenMaxChunks) {
         const int bands = (S.height + kDenBand - 1) / kDenBand;
         const int cblks = (S.width + kDenCols - 1) / kDenCols;
         const size_t lds = den_lds_bytes(tw);
         auto go = [&](auto kernel) {
            if (lds > 64 * 1024)
               hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                         "lds attribute");
            kernel<<<bands * cblks, 64, lds, s>>>(static_cast<const char *>(S.ptr), S.step,
                                                  S.width, S.height, tw, th, tsums, invden, bands);
         };
         // buffer loads when every byte offset, and two out-of-range offsets added, stay
         // below 2^31
         const bool buf = (long long)S.height * S.step <= kDenOob;
#define OCIP_DEN_CASE(N) \
   case N: if (buf) go(ncc_den_src<N, true>); else go(ncc_den_src<N, false>); break;
         switch (nch) {
         OCIP_DEN_CASE(1)
         OCIP_DEN_CASE(2)
         OCIP_DEN_CASE(3)
         OCIP_DEN_CASE(4)
         OCIP_DEN_CASE(5)
         OCIP_DEN_CASE(6)
         OCIP_DEN_CASE(7)
         OCIP_DEN_CASE(8)
         default: if (buf) go(ncc_den_src<9, true>); else go(ncc_den_src<9, false>); break;
         }
#undef OCIP_DEN_CASE
      } else {
         if (!integ) fail(CL_INVALID_VALUE, "integral required for templates this wide");
         const long long n = (long long)S.width * S.height;
         ncc_inv_den_direct<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(integ, S.width, S.height,
                                                                        tw, th, tsums, invden);
      }
      launch_check("ncc_inv_den");
   }
   // staged columns start at x0 - tw/2 + 64 ub + 8 g with x0 a multiple of 64: 16-byte
   // aligned when tw/2 is a multiple of 4 and the rows are 16-byte aligned
   const int vec = ((tw / 2) % 4 == 0 && rows_aligned(S)) ? 1 : 0;
   auto run = [&](int NT) {
      const Geometry g = geometry(precision, KC, NT, th);
      SingleLaunch L;
      L.a = mfma::TileArgs{static_cast<const float *>(S.ptr), S.width, S.height, S.step, tw, th,
                           nub, f0, f1, tsums, integ, invden, dst, dstep, g.vb, vec};
      L.KC = KC;
      L.NT = NT;
      L.mode = mode;
      L.ldsp = g.ldsp;
      L.lds = g.lds;
      L.grid = dim3((S.width + 32 * NT - 1) / (32 * NT),
                    (S.height + mfma::kTileRows - 1) / mfma::kTileRows);
      launch_single_p(precision, s, L);
   };
   const int NT = tuned_nt(tune_key("single", precision, KC, mode, S.width, S.height, th), s, run);
   KernelTimer kt(s, "crosscorr_mfma");
   run(NT);
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

   // ---- build: buckets by KC, templates in each bucket by decreasing cost
   std::map<int, std::vector<int>> buckets;
   for (int i = 0; i < n; ++i) buckets[kc_for(T[i].width)].push_back(i);
   for (auto &b : buckets)
      std::stable_sort(b.second.begin(), b.second.end(), [&](int a, int c) {
         return (long)T[a].height * nub_for(T[a].width) > (long)T[c].height * nub_for(T[c].width);
      });
   if (buckets.size() > 64) fail(CL_INVALID_VALUE, "too many template widths");
   const int planes = planes_for(precision);
   std::vector<size_t> frag_off(n);
   size_t nfrag_total = 0;
   for (int i = 0; i < n; ++i) {
      frag_off[i] = nfrag_total;
      nfrag_total += (size_t)nub_for(T[i].width) * T[i].height * kc_for(T[i].width);
   }
   BatchEntry *e = new BatchEntry();
   try {
      e->frags.ensure(planes * nfrag_total * 64 * sizeof(u16x8));
      e->sums.ensure((size_t)n * 2 * sizeof(double));
      e->descs.ensure((size_t)n * sizeof(BatchDesc));
      e->items.ensure((size_t)n * sizeof(BatchItem));
      e->counters.ensure(64 * sizeof(unsigned));
      e->integ.ensure(excl_bytes(S.width, S.height));
      e->aux.ensure(integral_aux_bytes(S.width, S.height));

      u16x8 *f0 = e->frags.as<u16x8>();
      u16x8 *f1 = planes == 2 ? f0 + nfrag_total * 64 : nullptr;
      double *sums = e->sums.as<double>();
      std::vector<BatchDesc> descs(n);
      for (int i = 0; i < n; ++i)
         descs[i] = BatchDesc{static_cast<const char *>(T[i].ptr), T[i].width, T[i].height,
                              T[i].step, kc_for(T[i].width), nub_for(T[i].width),
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
         const Geometry g = geometry(precision, k.KC, 4, k.th_max);
         if (g.vb < 1) fail(CL_INVALID_VALUE, "template too large for the MFMA tile");
         for (int i : b.second)
            items.push_back(BatchItem{T[i].width, T[i].height, nub_for(T[i].width),
                                      std::min(T[i].height, g.vb), descs[i].frag0,
                                      descs[i].frag1, sums + 2 * i, dst[i], dstep[i]});
         bl.push_back(k);
      }
      hip_check(hipMemcpy(e->items.ptr, items.data(), n * sizeof(BatchItem),
                          hipMemcpyHostToDevice), "batch table");

      unsigned *counters = e->counters.as<unsigned>();
      const BatchItem *ditems = e->items.as<BatchItem>();
      auto bucket_launch = [&](size_t b, int NT) {
         const Bucket &k = bl[b];
         const Geometry g = geometry(precision, k.KC, NT, k.th_max);
         BatchLaunch L;
         L.src = static_cast<const float *>(S.ptr);
         L.W = S.width;
         L.H = S.height;
         L.sstep = S.step;
         L.vec = rows_aligned(S) ? 1 : 0;
         L.integ = e->integ.as<double>();
         L.items = ditems + k.first;
         L.nitems = k.count;
         L.tiles_x = (S.width + 32 * NT - 1) / (32 * NT);
         L.tiles_y = (S.height + mfma::kTileRows - 1) / mfma::kTileRows;
         L.counter = counters + b;
         L.KC = k.KC;
         L.NT = NT;
         L.mode = mode;
         L.ldsp = g.ldsp;
         L.lds = g.lds;
         const int per_cu = std::max(1, std::min(2, (int)(mfma::kLdsMax / std::max<size_t>(g.lds, 1))));
         L.grid = num_cus * per_cu;
         launch_batch_p(precision, s, L);
      };
      auto prep = [&]() {
         launch_integral(s, S, 0, true, e->integ.ptr, 0, -1, e->aux.ptr);
         launch_batch_prep_p(precision, s, e->descs.as<BatchDesc>(), n);
      };
      auto reset = [&]() {
         reset_counters<<<1, 64, 0, s>>>(counters, (int)bl.size());
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
            nts.push_back(tuned_nt(tune_key("batch", precision, bl[b].KC, mode, S.width, S.height,
                                            bl[b].th_max), s, run));
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

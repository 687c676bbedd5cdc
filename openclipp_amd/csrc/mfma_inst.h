// mfma_inst.h -- kernel instances and launchers of mfma_kernels.h for one precision
// (OCIP_MFMA_P), included once by each of mfma_bf16.hip, mfma_f16.hip, mfma_bf16x3.hip.
#include "mfma_kernels.h"

namespace ocip {
namespace mfma {

namespace {

constexpr int P = OCIP_MFMA_P;

// The kernels take up to 160 KiB of LDS; the attribute is set once per instance.
template <class K> void allow_lds(K kernel)
{
   static bool done = false;
   if (done) return;
   hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, kLdsMax),
             "lds attribute");
   done = true;
}

template <int KC, int NT, int MODE> void single(hipStream_t s, const SingleLaunch &L)
{
   auto k = crosscorr_mfma<P, KC, NT, MODE>;
   allow_lds(k);
   const int tiles = L.grid.x * L.grid.y;
   k<<<tiles * std::max(1, L.z.n), 256, L.lds, s>>>(L.a, L.ldsp, (int)L.grid.x, tiles, L.z);
}

template <int KC, int NT, int MODE> void batch(hipStream_t s, const BatchLaunch &L)
{
   auto k = crosscorr_mfma_batch<P, KC, NT, MODE>;
   allow_lds(k);
   k<<<L.grid, 256, L.lds, s>>>(L.src, L.W, L.H, L.sstep, L.vec, L.integ, L.items, L.nitems,
                                L.tiles_x, L.tiles_y, L.counter, L.ldsp);
}

template <int KC, int NT, int MODE> void pair(hipStream_t s, const SingleLaunch &L)
{
   auto k = crosscorr_mfma_pair<P, KC, NT, MODE>;
   allow_lds(k);
   const int tiles = L.grid.x * L.grid.y;
   k<<<tiles * std::max(1, L.z.n), 256, L.lds, s>>>(L.a, L.ldsp, (int)L.grid.x, tiles, L.z);
}

template <int KC> void pair_mode(hipStream_t s, const SingleLaunch &L)
{
   if (L.mode == kModeInt) {
      if constexpr (P == kBF16) pair<KC, 4, kModeInt>(s, L);
      else fail(CL_INVALID_VALUE, "exact integer sums take BF16 operands");
      return;
   }
   if (L.NT == 4) pair<KC, 4, kModeCC>(s, L); else pair<KC, 6, kModeCC>(s, L);
}

// the ring-buffer strip and the persistent batch kernel stage one operand plane
#if OCIP_MFMA_P != 3
template <int KC, bool VEC> void strip(hipStream_t s, const SingleLaunch &L, int tiles_y, int q_rows)
{
   auto k = crosscorr_mfma_strip<P, KC, 4, kModeCC, VEC>;
   allow_lds(k);
   k<<<L.grid.x, 256, L.lds, s>>>(L.a, L.ldsp, (int)L.grid.y, tiles_y, q_rows);
}

// 16-byte ring loads when every 4-column group of the source is wholly inside or outside the
// image: 16-byte aligned rows (L.a.vec), tw/2 and W multiples of 4
template <int KC> void strip_mode(hipStream_t s, const SingleLaunch &L, int tiles_y, int q_rows)
{
   if (L.a.vec && L.a.W % 4 == 0) strip<KC, true>(s, L, tiles_y, q_rows);
   else strip<KC, false>(s, L, tiles_y, q_rows);
}

template <int KC, int NT, int MODE> void pair_batch(hipStream_t s, const BatchLaunch &L)
{
   auto k = crosscorr_mfma_pair_batch<P, KC, NT, MODE>;
   allow_lds(k);
   k<<<L.grid, 256, L.lds, s>>>(L.src, L.W, L.H, L.sstep, L.vec, L.integ, L.items, L.nitems,
                                L.tiles_x, L.tiles_y, L.counter, L.ldsp);
}

template <int KC, int MODE> void pair_batch_nt(hipStream_t s, const BatchLaunch &L)
{
   if (L.NT == 4) pair_batch<KC, 4, MODE>(s, L); else pair_batch<KC, 6, MODE>(s, L);
}

template <int KC> void pair_batch_mode(hipStream_t s, const BatchLaunch &L)
{
   if (L.mode == kModeNCC) pair_batch_nt<KC, kModeNCC>(s, L); else pair_batch_nt<KC, kModeSSDN>(s, L);
}
#endif

template <int KC, int MODE> void single_nt(hipStream_t s, const SingleLaunch &L)
{
   if (L.NT == 2) single<KC, 2, MODE>(s, L); else single<KC, 4, MODE>(s, L);
}

template <int KC, int MODE> void batch_nt(hipStream_t s, const BatchLaunch &L)
{
   if (L.NT == 2) batch<KC, 2, MODE>(s, L); else batch<KC, 4, MODE>(s, L);
}

template <int KC> void single_mode(hipStream_t s, const SingleLaunch &L)
{
   single_nt<KC, kModeCC>(s, L);
}

template <int KC> void batch_mode(hipStream_t s, const BatchLaunch &L)
{
   if (L.mode == kModeNCC) batch_nt<KC, kModeNCC>(s, L); else batch_nt<KC, kModeSSDN>(s, L);
}

}  // namespace

template <> void launch_single<P>(hipStream_t s, const SingleLaunch &L)
{
   switch (L.KC) {
   case 2: single_mode<2>(s, L); break;
   case 3: single_mode<3>(s, L); break;
   case 4: single_mode<4>(s, L); break;
   case 5: single_mode<5>(s, L); break;
   case 6: single_mode<6>(s, L); break;
   case 7: single_mode<7>(s, L); break;
   case 8: single_mode<8>(s, L); break;
   default: fail(CL_INVALID_VALUE, "template width");
   }
   launch_check("crosscorr_mfma");
}

template <> void launch_batch<P>(hipStream_t s, const BatchLaunch &L)
{
   switch (L.KC) {
   case 2: batch_mode<2>(s, L); break;
   case 3: batch_mode<3>(s, L); break;
   case 4: batch_mode<4>(s, L); break;
   case 5: batch_mode<5>(s, L); break;
   case 6: batch_mode<6>(s, L); break;
   case 7: batch_mode<7>(s, L); break;
   case 8: batch_mode<8>(s, L); break;
   default: fail(CL_INVALID_VALUE, "template width");
   }
   launch_check("crosscorr_mfma_batch");
}

template <> void launch_pair<P>(hipStream_t s, const SingleLaunch &L)
{
   switch (L.KC) {
   case 1: pair_mode<1>(s, L); break;
   case 2: pair_mode<2>(s, L); break;
   case 3: pair_mode<3>(s, L); break;
   case 4: pair_mode<4>(s, L); break;
   case 5: pair_mode<5>(s, L); break;
   case 6: pair_mode<6>(s, L); break;
   case 7: pair_mode<7>(s, L); break;
   default: fail(CL_INVALID_VALUE, "template width");
   }
   launch_check("crosscorr_mfma_pair");
}

template <> void launch_pair_prep<P>(hipStream_t s, const BatchDesc &d)
{
   const int nfrag = d.nub * pair_frag_steps(d.th, d.nub) * d.KC;
   pair_prep<P><<<(nfrag + 3) / 4 + 1, 256, 0, s>>>(d);   // + 1: the sum of squares
   launch_check("pair_prep");
}

#if OCIP_MFMA_P != 3
template <> void launch_pair_batch<P>(hipStream_t s, const BatchLaunch &L)
{
   switch (L.KC) {
   case 1: pair_batch_mode<1>(s, L); break;
   case 2: pair_batch_mode<2>(s, L); break;
   case 3: pair_batch_mode<3>(s, L); break;
   case 4: pair_batch_mode<4>(s, L); break;
   case 5: pair_batch_mode<5>(s, L); break;
   case 6: pair_batch_mode<6>(s, L); break;
   case 7: pair_batch_mode<7>(s, L); break;
   default: fail(CL_INVALID_VALUE, "template width");
   }
   launch_check("crosscorr_mfma_pair_batch");
}

template <> void launch_pair_batch_prep<P>(hipStream_t s, const BatchDesc *descs, int n)
{
   batch_prep<P, true><<<dim3(kPrepBlocks, n), 256, 0, s>>>(descs);
   launch_check("batch_prep");
}

// grid.x = workgroups, grid.y = tiles_x
template <> void launch_strip<P>(hipStream_t s, const SingleLaunch &L, int tiles_y, int q_rows)
{
   switch (L.KC) {
   case 2: strip_mode<2>(s, L, tiles_y, q_rows); break;
   case 3: strip_mode<3>(s, L, tiles_y, q_rows); break;
   case 4: strip_mode<4>(s, L, tiles_y, q_rows); break;
   case 5: strip_mode<5>(s, L, tiles_y, q_rows); break;
   case 6: strip_mode<6>(s, L, tiles_y, q_rows); break;
   default: fail(CL_INVALID_VALUE, "template width");
   }
   launch_check("crosscorr_mfma_strip");
}
#endif

template <> void launch_single_prep<P>(hipStream_t s, const BatchDesc &d)
{
   const int nfrag = d.nub * d.th * d.KC;
   single_prep<P><<<(nfrag + 3) / 4 + 1, 256, 0, s>>>(d);   // + 1: the sum of squares
   launch_check("single_prep");
}

template <> void launch_batch_prep<P>(hipStream_t s, const BatchDesc *descs, int n)
{
   batch_prep<P, false><<<dim3(kPrepBlocks, n), 256, 0, s>>>(descs);
   launch_check("batch_prep");
}

}  // namespace mfma
}  // namespace ocip

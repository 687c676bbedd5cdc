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
   k<<<tiles, 256, L.lds, s>>>(L.a, L.ldsp, (int)L.grid.x, tiles);
}

template <int KC, int NT, int MODE> void batch(hipStream_t s, const BatchLaunch &L)
{
   auto k = crosscorr_mfma_batch<P, KC, NT, MODE>;
   allow_lds(k);
   k<<<L.grid, 256, L.lds, s>>>(L.src, L.W, L.H, L.sstep, L.vec, L.integ, L.items, L.nitems,
                                L.tiles_x, L.tiles_y, L.counter, L.ldsp);
}

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
   if (L.mode == kModeCC) single_nt<KC, kModeCC>(s, L); else single_nt<KC, kModeNCCMap>(s, L);
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

template <> void launch_single_prep<P>(hipStream_t s, const BatchDesc &d)
{
   const int nfrag = d.nub * d.th * d.KC;
   single_prep<P><<<(nfrag + 3) / 4, 256, 0, s>>>(d);
   launch_check("single_prep");
}

template <> void launch_batch_prep<P>(hipStream_t s, const BatchDesc *descs, int n)
{
   batch_prep<P><<<n, 256, 0, s>>>(descs);
   launch_check("batch_prep");
}

}  // namespace mfma
}  // namespace ocip

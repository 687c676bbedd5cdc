// Internal declarations shared by the C-ABI layer (capi.cpp) and the HIP kernels.
//
// Object model (reference equivalents in brackets):
//   Context  [COpenCL, C++/OpenCL.cpp:48-69]  one GPU + one in-order HIP stream, plus the
//            device workspaces the proximity ops reuse between calls.
//   Image    [Image/Buffer, C++/Image.cpp:147-165, C++/Buffer.cpp:100-194]  Height*Step
//            bytes of device memory, the caller's host pointer (not copied), an in-device flag.
//   Program  [Integral / Statistics programs, include/c++/Programs/Integral.h:32-61]
//            scratch buffers owned by an ocipProgram handle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>

#include "OpenCLIPP.h"

#ifndef CL_INVALID_BUFFER_SIZE
#define CL_INVALID_BUFFER_SIZE -61
#endif

namespace ocip {

// An error raised anywhere below the C ABI; the ABI layer returns `code`.
struct Error {
   cl_int code;
   const char *what;
};

[[noreturn]] inline void fail(cl_int code, const char *what) { throw Error{code, what}; }

// Failing HIP call: the error is mapped to an OpenCL code; with OCIP_VERBOSE set in the
// environment the HIP error text is printed to stderr.
void report_hip_error(hipError_t e, const char *what);
inline void hip_check(hipError_t e, const char *what)
{
   if (e != hipSuccess) {
      report_hip_error(e, what);
      fail(e == hipErrorOutOfMemory ? CL_MEM_OBJECT_ALLOCATION_FAILURE : CL_OUT_OF_RESOURCES, what);
   }
}

// Growable device buffer (contents are not preserved by `ensure`).
struct DevBuf {
   void *ptr = nullptr;
   size_t bytes = 0;
   void ensure(size_t n);
   void release();
   template <class T> T *as() const { return static_cast<T *>(ptr); }
};

struct Context;
struct BatchCache;

enum : uint32_t { kImageMagic = 0x474d494fu /*OIMG*/, kContextMagic = 0x5854434fu /*OCTX*/,
                  kProgramMagic = 0x4752504fu /*OPRG*/ };

// desc describes the device image the kernels see.  A 3-channel (colour) image is held on
// the device as 4 channels with the fourth set to 255 (reference ColorImage,
// C++/Image.cpp:243-275): host_desc keeps the caller's 3-channel layout, `stage` a device copy
// of it, and transfers convert between the two.
struct Image {
   uint32_t magic = kImageMagic;
   Context *ctx = nullptr;
   SImage desc{};
   SImage host_desc{};
   void *host = nullptr;
   void *dev = nullptr;
   void *stage = nullptr;   // colour images only
   size_t bytes = 0;        // bytes of the host image (Height * host Step)
   bool in_device = false;
};

// Colour conversions between the 3-channel host layout and the 4-channel device layout.
void launch_copy3to4(hipStream_t s, int type, const void *src, int sstep, void *dst, int dstep,
                     int width, int height);
void launch_copy4to3(hipStream_t s, int type, const void *src, int sstep, void *dst, int dstep,
                     int width, int height);

enum ProgramKind { kProgIntegral = 1, kProgStatistics = 2, kProgProximityFFT = 3, kProgFFT = 4 };

struct FftState;
FftState *fft_state_create();
void fft_state_destroy(FftState *s);

struct Program {
   uint32_t magic = kProgramMagic;
   Context *ctx = nullptr;
   int kind = 0;
   DevBuf scratch;
   FftState *fft = nullptr;   // kProgProximityFFT, kProgFFT: plans, twiddles, work buffers
};

// Per-template data the proximity kernels read: the template as F32 (or F64 / int32)
// with zero-padded rows, plus FP64 sum of squares.
struct TemplateBuf {
   DevBuf data;      // th rows x pitch elements
   DevBuf sums;      // [0] = sum t^2 (double), [1] = sum t (double)
};

struct Context {
   uint32_t magic = kContextMagic;
   int device = 0;
   hipStream_t stream = nullptr;
   std::string name;
   int num_cus = 0;
   TemplateBuf tmpl;      // template staging for the single-template ops
   DevBuf integ;          // FP64 exclusive sum-of-squares integral, (W+1) x (H+1)
   DevBuf integ_aux;      // integral scan aggregates
   DevBuf stats;          // reduction partials for MinIndx / MaxIndx, MFMA fragments
   DevBuf work;           // gradient workspace
   BatchCache *batch = nullptr;  // recorded batch graphs (mfma.hip)
};

// Kernel timing: construct before a launch, destroy after it.  Records a HIP event pair on
// `s` when timing is enabled (ocipEnableKernelTiming), nothing otherwise.
// While suspended (stream capture), no timing events are recorded.
void timing_suspend(bool suspend);
struct KernelTimer {
   KernelTimer(hipStream_t s, const char *name);
   ~KernelTimer();
   hipStream_t stream;
   const char *name;
   void *start = nullptr;
};

// ---------------------------------------------------------------------------------------
// Element types (SImage::EDataType)
inline int elem_size(int t)
{
   switch (t) {
   case SImage::U8: case SImage::S8: return 1;
   case SImage::U16: case SImage::S16: return 2;
   case SImage::U32: case SImage::S32: case SImage::F32: return 4;
   case SImage::F64: return 8;
   default: return 0;
   }
}
inline bool is_float(int t) { return t == SImage::F32 || t == SImage::F64; }
inline bool is_unsigned(int t) { return t == SImage::U8 || t == SImage::U16 || t == SImage::U32; }

// Proximity operations
enum ProxOp { kSSD = 0, kSSDNorm = 1, kSAD = 2, kCC = 3, kCCNorm = 4 };

// A device image view handed to kernel launchers.
struct View {
   const void *ptr;
   int width, height, step;   // step in bytes
   int channels, type;
};

// Internal exclusive FP64 integral (window sums for the Norm ops and gradients): H+1 rows of
// excl_pitch(W) doubles; entry (j, i) = sum over rows < j and columns < i (0 <= i <= W)
// sits at j * pitch + i + kExclOffset, so that entries 1..W of a row start on a 32-byte
// boundary and a row is written with full 32-byte stores.
constexpr int kExclOffset = 3;
__host__ __device__ inline size_t excl_pitch(int W) { return ((size_t)W + kExclOffset + 1 + 3) / 4 * 4; }
inline size_t excl_bytes(int W, int H) { return excl_pitch(W) * ((size_t)H + 1) * sizeof(double); }

// ---------------------------------------------------------------------------------------
// Kernel launchers (defined in the .hip files).  All enqueue on `s` and return immediately.

// Integral of f(src) with f = identity or square, FP64 sums.
//   out_type F32/F64 with inclusive same-size layout (public API), or
//   out_type = -1: the internal exclusive layout above (out_step ignored).
size_t integral_aux_bytes(int width, int height);
void launch_integral(hipStream_t s, const View &src, int channel, bool sqr, void *out,
                     int out_step, int out_type, void *aux);

// Template staging: converts the template to the kernels' compute type with a row pitch of
// `pitch` elements (zero padding), and writes FP64 sum(t^2), sum(t) to sums[0..1].
// kind 0: float, 1: double, 2: int32.
void launch_template_prep(hipStream_t s, const View &tpl, int channel, int pitch, int kind,
                          void *out, double *sums);

// Plain (VALU) proximity for any type/channel count.  `integ` is the exclusive FP64
// sum-of-squares integral of the source for the Norm ops (null for the others).
void launch_proximity(hipStream_t s, int op, const View &src, const View &tpl, void *dst,
                      int dst_step, const void *tstage, int tpitch, const double *tsums,
                      const double *integ, int channel);
// The plain Norm ops on F32 sources whose template fits one LDS band form the window sums
// of s^2 inside the tile kernel (launch_proximity with no integral)
bool prox_norm_from_tile(int op, const View &src, const View &tpl);

// MFMA cross-correlation (F32 1-channel source and template), precision 1..3.
// The template's fragments and sum of squares (tsums[0]) are built from T in the call.
// may_tune = false: no autotuning (no timing events); the variant tuned earlier for this
// shape, else a fixed default.
void launch_crosscorr_mfma(hipStream_t s, int precision, bool norm, const View &src,
                           const View &T, double *tsums, const double *integ, float *dst,
                           int dst_step, void *work, bool may_tune = true);
// The single-template MFMA Norm path forms its denominators from the source image; only
// templates wider than 513 columns read the FP64 sum-of-squares integral instead.
// Split-BF16 CrossCorr maps (no Norm) of n sources of one size (row pitches sstep[]) with one
// template, whose fragments (crosscorr_split_frag_bytes) prep_crosscorr_split wrote to
// `frags` (and sum t^2 to sums[0]); up to 16 sources per launch
constexpr int kSplitImages = 16;   // sources per launch of launch_crosscorr_split
size_t crosscorr_split_frag_bytes(int tw, int th);
void prep_crosscorr_split(hipStream_t s, const View &T, void *frags, double *sums);
void launch_crosscorr_split(hipStream_t s, int n, const float *const *src, const int *sstep,
                            float *const *dst, int dstep, int W, int H, int tw, int th,
                            const void *frags, const double *sums);
bool crosscorr_mfma_needs_integral(int tw);
size_t crosscorr_mfma_work_bytes(int precision, int width, int height, int tw, int th);

// Scratch of the entry point that allocates nothing (noalloc.hip): a device array of
// kNoAllocScratchBytes per GPU.  noalloc_scratch_init takes its address on `device` (which
// must be current) once; noalloc_scratch returns that address.
constexpr size_t kNoAllocScratchBytes = (size_t)8 << 20;
constexpr size_t kNoAllocSumsBytes = 256;   // template sums ahead of the fragments
void noalloc_scratch_init(int device);
void *noalloc_scratch(int device);

// Exact integer sums of 8-bit images on BF16 MFMA (CrossCorr, CrossCorr_Norm, SqrDistance,
// SqrDistance_Norm of U8/S8 1-channel images): int_mfma_ok tells whether the path takes the
// call; tsums[0] receives sum t^2.  The maps equal prox_tile_int's bit for bit.
bool int_mfma_ok(int op, const View &src, const View &T);
size_t int_mfma_work_bytes(int tw, int th);
void launch_int_mfma(hipStream_t s, int op, const View &src, const View &T, double *tsums,
                     float *dst, int dst_step, void *work);

// Batched MFMA CrossCorr_Norm (ssd = false) or SqrDistance_Norm (ssd = true) of F32 1-channel
// templates against one F32 1-channel source, recorded into a HIP graph on first use and
// replayed for identical arguments.  The cache belongs to the context.
BatchCache *batch_cache_create();
void batch_cache_destroy(BatchCache *c);
void run_batch_mfma(hipStream_t s, BatchCache *cache, int num_cus, int precision, bool ssd,
                    const View &S, const std::vector<View> &T, const std::vector<float *> &dst,
                    const std::vector<int> &dstep);

// Gradients (backward.hip) of L = sum_i [sum G1_i CrossCorr_Norm(S_i, t) + sum G2_i
// SqrDistance(S_i, t)] over n same-size F32 1-channel images: G1 / G2 null leaves that op out;
// dS[i] / dT may be null.  dT is the sum over the batch.
size_t backward_work_bytes(int n, int W, int H, int tw, int th, int tpitch);
void launch_backward_batch(hipStream_t s, int n, const View *S, const View *G1, const View *G2,
                           const float *tstage, int tpitch, int tw, int th, const double *tsums,
                           void *work, void *aux, char *const *dS, const int *dsstep, char *dT,
                           int dtstep);

// Proximity through 2-D FFTs (FP64): op kSSD, kSSDNorm, kCC or kCCNorm on an F32 1-channel
// source and template into an F32 dest.  integ: the exclusive FP64 sum-of-squares integral
// (all ops but kCC); tsums[0] = sum t^2.  fft_sizes gives the padded transform size or
// CL_INVALID_IMAGE_SIZE when the image is too big for the path.
cl_int fft_sizes(int W, int H, int tw, int th, int *Wb, int *Hb);
void fft_prepare(hipStream_t s, FftState *st, int W, int H, int tw, int th);
void launch_proximity_fft(hipStream_t s, FftState *st, int op, const View &S, const View &T,
                          const double *integ, const double *tsums, void *dst, int dstep);

// Generic 2-D FFT (F32 or F64): forward real W x H -> half spectrum (W/2 + 1) x H of
// interleaved complex values, inverse back, unnormalised.  fft_plain_sizes: CL_SUCCESS or
// CL_INVALID_IMAGE_SIZE (a length with a prime factor above 5, or beyond LDS).
cl_int fft_plain_sizes(int W, int H);
void fft_plain_prepare(hipStream_t s, FftState *st, int W, int H);
void launch_fft(hipStream_t s, FftState *st, bool forward, const View &real, const View &cplx);

// Per-channel sum (sqr = false) or sum of squares of an image into result[0..channels),
// blocking.  partials: sum_partials_bytes of device scratch.
size_t sum_partials_bytes(int width, int height);
void launch_sum(hipStream_t s, bool sqr, const View &src, void *partials, double *result);
// Min / max of a 1-channel image and the pixel holding it, in the reference's reduction order
// (its tie-breaks and float partials), blocking.
size_t minmax_partials_bytes(int width, int height);
void launch_minmax(hipStream_t s, bool want_max, const View &src, void *partials,
                   double *host_result, int *ix, int *iy);

}  // namespace ocip

/* OpenCLIPP.h -- C interface of the MI355X template-matching library.
 *
 * Drop-in for the template-matching subset of the reference C API
 * (include/OpenCLIPP.h).  Every declaration in the "reference interface"
 * block has the same name, argument list, argument meaning and error codes
 * as the reference declaration cited beside it, so C code written against
 * the reference compiles and links unchanged against libOpenCLIPP.so built
 * from this repository.  The implementation behind it is HIP on gfx950; no
 * OpenCL runtime is loaded or linked.  The OpenCL scalar types and error
 * constants are used only as plain integers, with the Khronos numeric values.
 *
 * The "extensions" block at the end is additive: it has no counterpart in
 * the reference and changes nothing in the reference interface.
 */
#ifndef OCIP_OPENCLIPP_H
#define OCIP_OPENCLIPP_H

#include <stdint.h>

/* cl_int, cl_device_type, cl_mem_flags and the CL_* constants.  Use the
 * system Khronos header when it exists (the values below are identical). */
#if defined(__has_include)
#  if __has_include(<CL/cl.h>)
#    define OCIP_HAVE_CL_H 1
#  endif
#endif

#ifdef OCIP_HAVE_CL_H
#  include <CL/cl.h>
#else
typedef int32_t  cl_int;
typedef uint64_t cl_ulong;
typedef cl_ulong cl_bitfield;
typedef cl_bitfield cl_device_type;
typedef cl_bitfield cl_mem_flags;
#  define CL_SUCCESS                         0
#  define CL_DEVICE_NOT_FOUND               -1
#  define CL_DEVICE_NOT_AVAILABLE           -2
#  define CL_MEM_OBJECT_ALLOCATION_FAILURE  -4
#  define CL_OUT_OF_RESOURCES               -5
#  define CL_OUT_OF_HOST_MEMORY             -6
#  define CL_IMAGE_FORMAT_NOT_SUPPORTED    -10
#  define CL_INVALID_VALUE                 -30
#  define CL_INVALID_DEVICE_TYPE           -31
#  define CL_INVALID_CONTEXT               -34
#  define CL_INVALID_MEM_OBJECT            -38
#  define CL_INVALID_IMAGE_SIZE            -40
#  define CL_INVALID_PROGRAM               -44
#  define CL_INVALID_OPERATION             -59
#  define CL_INVALID_BUFFER_SIZE           -61
#  define CL_DEVICE_TYPE_DEFAULT      (1 << 0)
#  define CL_DEVICE_TYPE_CPU          (1 << 1)
#  define CL_DEVICE_TYPE_GPU          (1 << 2)
#  define CL_DEVICE_TYPE_ACCELERATOR  (1 << 3)
#  define CL_DEVICE_TYPE_ALL          0xFFFFFFFF
#  define CL_MEM_READ_WRITE           (1 << 0)
#  define CL_MEM_WRITE_ONLY           (1 << 1)
#  define CL_MEM_READ_ONLY            (1 << 2)
#endif

/* Only the ocip* entry points are exported from the shared library. */
#if defined(OCIP_BUILD)
#  define ocip_API __attribute__((visibility("default")))
#else
#  define ocip_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef char ocipBool;
typedef unsigned int uint;

#include "SImage.h"
typedef struct SImage SImage;

typedef cl_int ocipError;                      /* reference include/OpenCLIPP.h:62 */
typedef struct _cl_context * ocipContext;      /* reference include/OpenCLIPP.h:64 */
typedef struct _cl_image   * ocipImage;        /* reference include/OpenCLIPP.h:65 */
typedef struct _cl_program * ocipProgram;      /* reference include/OpenCLIPP.h:66 */

/* ===================== reference interface (template-matching path) ===================== */

/* Context ---------------------------------------------------------------------------------
 * A context is one GPU plus one in-order HIP stream.  Creating a context makes it current.
 * PreferredPlatform is accepted and ignored; deviceType must include the GPU bit
 * (CL_DEVICE_TYPE_ALL, _DEFAULT or _GPU), otherwise CL_DEVICE_NOT_FOUND. */
ocipError ocip_API ocipInitialize(ocipContext * ContextPtr, const char * PreferredPlatform,
                                  cl_device_type deviceType);           /* ref OpenCLIPP.h:99  */
ocipError ocip_API ocipUninitialize(ocipContext Context);               /* ref OpenCLIPP.h:104 */
ocipError ocip_API ocipChangeContext(ocipContext Context);              /* ref OpenCLIPP.h:112 */
/* Kernels are compiled ahead of time into the library; the path is recorded and unused. */
void      ocip_API ocipSetCLFilesPath(const char * Path);               /* ref OpenCLIPP.h:118 */
ocip_API const char * ocipGetErrorName(ocipError Error);               /* ref OpenCLIPP.h:124 */
ocipError ocip_API ocipGetDeviceName(char * Name, uint BufferLength);   /* ref OpenCLIPP.h:130 */
ocipError ocip_API ocipFinish(void);                                    /* ref OpenCLIPP.h:135 */

/* Images ----------------------------------------------------------------------------------
 * Device memory of Height*Step bytes.  ImageData is kept (not copied) and used as the
 * host side of Send/Read; NULL makes a device-only image.  A source image that was never
 * sent is sent automatically on first use.  Send is asynchronous, Read blocks until all
 * earlier work of the context is done. */
ocipError ocip_API ocipCreateImage(ocipImage * ImagePtr, SImage Image, void * ImageData,
                                   cl_mem_flags flags);                 /* ref OpenCLIPP.h:159 */
ocipError ocip_API ocipSendImage(ocipImage Image);                      /* ref OpenCLIPP.h:170 */
ocipError ocip_API ocipReadImage(ocipImage Image);                      /* ref OpenCLIPP.h:177 */
ocipError ocip_API ocipReleaseImage(ocipImage Image);                   /* ref OpenCLIPP.h:182 */
ocipError ocip_API ocipReleaseProgram(ocipProgram Program);             /* ref OpenCLIPP.h:231 */

/* Integral --------------------------------------------------------------------------------
 * Inclusive 2-D prefix sum (or sum of squares), same size as Source, 1 channel.
 * Dest must be F32 or F64.  Sums are formed in FP64 (exact int64 for integer sources)
 * and rounded once to the Dest type. */
ocipError ocip_API ocipPrepareIntegral(ocipProgram * ProgramPtr, ocipImage Image); /* ref :639 */
ocipError ocip_API ocipIntegral(ocipProgram Program, ocipImage Source, ocipImage Dest); /* :642 */
ocipError ocip_API ocipSqrIntegral(ocipProgram Program, ocipImage Source, ocipImage Dest); /* :645 */

/* Image proximity -------------------------------------------------------------------------
 * Dest(x,y) over a Template centred at (x,y) (origin offset -tw/2, -th/2, C division),
 * source pixels outside the image read as 0, Dest the same size as Source.
 * Asynchronous: use ocipReadImage or ocipFinish to wait. */
ocipError ocip_API ocipPrepareProximity(ocipImage Image);                           /* ref :654 */
ocipError ocip_API ocipSqrDistance_Norm(ocipImage Source, ocipImage Template, ocipImage Dest); /* :657 */
ocipError ocip_API ocipSqrDistance(ocipImage Source, ocipImage Template, ocipImage Dest);      /* :660 */
ocipError ocip_API ocipAbsDistance(ocipImage Source, ocipImage Template, ocipImage Dest);      /* :663 */
ocipError ocip_API ocipCrossCorr(ocipImage Source, ocipImage Template, ocipImage Dest);        /* :666 */
ocipError ocip_API ocipCrossCorr_Norm(ocipImage Source, ocipImage Template, ocipImage Dest);   /* :669 */

/* Statistics: match localisation and template energy (reference include/OpenCLIPP.h:546-561) 
 * Minimum / maximum over all pixels of a 1-channel image and the first (row-major) pixel
 * holding it.  Program comes from ocipPrepareStatistics (may be NULL). */
ocipError ocip_API ocipPrepareStatistics(ocipProgram * ProgramPtr, ocipImage Image); /* ref :546 */
ocipError ocip_API ocipMinIndx(ocipProgram Program, ocipImage Source, double * Result,
                               int * IndexX, int * IndexY);             /* ref OpenCLIPP.h:560 */
ocipError ocip_API ocipMaxIndx(ocipProgram Program, ocipImage Source, double * Result,
                               int * IndexX, int * IndexY);             /* ref OpenCLIPP.h:561 */
/* Sum and sum of squares of every channel: Result must hold one double per channel
 * (reference Statistics::Sum / SumSqr over all channels, Statistics.cpp:350-370). */
ocipError ocip_API ocipSum(ocipProgram Program, ocipImage Source, double * Result);    /* ref :552 */
ocipError ocip_API ocipSumSqr(ocipProgram Program, ocipImage Source, double * Result); /* ref :553 */

/* FFT (reference include/OpenCLIPP.h:612-634) --------------------------------------------
 * 2-D transform of a 1-channel F32 (or F64) real image of 2-3-5-smooth width and height into
 * its half spectrum: a complex image (real, imaginary interleaved, same data type) of at least
 * Width/2+1 complex values per row and the same height; the inverse goes back.  Neither
 * direction is normalised (inverse(forward(x)) = Width*Height*x).  Computed in FP64. */
ocipError ocip_API ocipPrepareFFT(ocipProgram * ProgramPtr, ocipImage RealImage,
                                  ocipImage ComplexImage);                                /* ref :613 */
ocipError ocip_API ocipFFTForward(ocipProgram Program, ocipImage RealSource,
                                  ocipImage ComplexDest);                                 /* ref :625 */
ocipError ocip_API ocipFFTInverse(ocipProgram Program, ocipImage ComplexSource,
                                  ocipImage RealDest);                                    /* ref :634 */

/* Proximity through FFTs (reference include/OpenCLIPP.h:673-688) -------------------------
 * F32 1-channel images; faster than the direct ops for large templates.  The program comes
 * from ocipPrepareImageProximityFFT (required).  The transforms are computed in FP64. */
ocipBool  ocip_API ocipIsFFTAvailable(void);                                              /* ref :615 */
ocipError ocip_API ocipPrepareImageProximityFFT(ocipProgram * ProgramPtr, ocipImage Image,
                                                ocipImage Template);                     /* ref :676 */
ocipError ocip_API ocipSqrDistanceFFT(ocipProgram Program, ocipImage Source,
                                      ocipImage Template, ocipImage Dest);               /* ref :679 */
ocipError ocip_API ocipSqrDistanceFFT_Norm(ocipProgram Program, ocipImage Source,
                                           ocipImage Template, ocipImage Dest);          /* ref :682 */
ocipError ocip_API ocipCrossCorrFFT(ocipProgram Program, ocipImage Source,
                                    ocipImage Template, ocipImage Dest);                 /* ref :685 */
ocipError ocip_API ocipCrossCorrFFT_Norm(ocipProgram Program, ocipImage Source,
                                         ocipImage Template, ocipImage Dest);            /* ref :688 */

/* ===================== extensions (no counterpart in the reference) ===================== */

/* Arithmetic used by the cross-correlation sums. */
enum ocipPrecision
{
   ocipPrecisionF32    = 0,  /* FP32 VALU, same as ocipCrossCorr / ocipCrossCorr_Norm        */
   ocipPrecisionBF16   = 1,  /* BF16 MFMA products, FP32 accumulation                         */
   ocipPrecisionF16    = 2,  /* FP16 MFMA products, FP32 accumulation                         */
   ocipPrecisionBF16x3 = 3   /* split hi+lo BF16, three MFMA products, FP32 accumulation      */
};

/* ocipCrossCorr / ocipCrossCorr_Norm with an explicit arithmetic (F32 1-channel sources for
 * the MFMA precisions).  Norm denominators always come from the FP64 sum-of-squares
 * integral of Source and the FP64 sum of squares of Template. */
ocipError ocip_API ocipCrossCorrEx(ocipImage Source, ocipImage Template, ocipImage Dest,
                                   int Precision);
ocipError ocip_API ocipCrossCorr_NormEx(ocipImage Source, ocipImage Template, ocipImage Dest,
                                        int Precision);

/* ocipCrossCorr_NormEx that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (serving loops, code under HIP graph capture,
 * harnesses that forbid hipMalloc).  MFMA precisions only (ocipPrecisionF32 gives
 * CL_INVALID_VALUE), templates up to 513 columns (wider ones: CL_INVALID_VALUE) whose
 * fragments fit an 8 MiB per-GPU scratch array held until the process ends (bigger ones:
 * CL_INVALID_BUFFER_SIZE; a 64x64 BF16 template needs about 0.6 MiB).  No autotuning: the
 * kernel variant tuned for the same shape by an earlier ocipCrossCorr_NormEx call, else a
 * fixed one, so its map equals ocipCrossCorr_NormEx's once that call has run.  Calls that
 * overlap on one GPU share the array and race. */
ocipError ocip_API ocipCrossCorr_NormExNoAlloc(ocipImage Source, ocipImage Template,
                                               ocipImage Dest, int Precision);

/* Many templates of any sizes against one F32 1-channel source: Dests[i] receives the map of
 * Templates[i].  One persistent launch (replayed from a HIP graph) covers the whole batch.
 * Equivalent to Count calls of ocipCrossCorr_NormEx / ocipSqrDistance_Norm. */
ocipError ocip_API ocipCrossCorr_NormBatch(ocipImage Source, const ocipImage * Templates,
                                           ocipImage * Dests, uint Count, int Precision);
ocipError ocip_API ocipSqrDistance_NormBatch(ocipImage Source, const ocipImage * Templates,
                                             ocipImage * Dests, uint Count, int Precision);

/* Gradients of L = sum(GradDest * op(Source, Template)) for F32 1-channel images.
 * GradSource is Source-sized, GradTemplate is Template-sized (both F32, overwritten).
 * Either gradient output may be NULL to skip it.  The Norm statistics are recomputed from
 * the FP64 integral of squares, not stored by the forward call. */
ocipError ocip_API ocipCrossCorr_NormBackward(ocipImage Source, ocipImage Template,
                                              ocipImage GradDest, ocipImage GradSource,
                                              ocipImage GradTemplate);
ocipError ocip_API ocipSqrDistanceBackward(ocipImage Source, ocipImage Template,
                                           ocipImage GradDest, ocipImage GradSource,
                                           ocipImage GradTemplate);

/* Gradients over a batch of same-size F32 1-channel sources sharing one template, of
 *   L = sum_i [ sum(GradNcc[i] * CrossCorr_Norm(Sources[i], Template))
 *             + sum(GradSsd[i] * SqrDistance(Sources[i], Template)) ]
 * GradSources[i] (Source-sized) receives dL/dSources[i]; GradTemplate (Template-sized) receives
 * dL/dTemplate, summed over the batch.  GradNcc or GradSsd may be NULL (that term absent);
 * GradSources, its entries, and GradTemplate may be NULL (that gradient skipped).  The two
 * ops' gradients share their correlations (one map per image combines both).  Equivalent to
 * Count calls of ocipCrossCorr_NormBackward and ocipSqrDistanceBackward with the results
 * added. */
ocipError ocip_API ocipProximityBackwardBatch(const ocipImage * Sources, ocipImage Template,
                                              const ocipImage * GradNcc,
                                              const ocipImage * GradSsd,
                                              ocipImage * GradSources, ocipImage GradTemplate,
                                              uint Count);

/* Kernel timing (profiling aid).  While enabled, every kernel the library launches is
 * bracketed by HIP events on the stream it runs on.  ocipGetKernelTime waits for the queued
 * work and returns the total time (ms) and launch count of the kernels whose name starts
 * with KernelName since timing was last enabled. */
ocipError ocip_API ocipEnableKernelTiming(ocipBool Enable);
ocipError ocip_API ocipGetKernelTime(const char * KernelName, double * TotalMs, uint * Count);

#ifdef __cplusplus
}
#endif

#endif /* OCIP_OPENCLIPP_H */

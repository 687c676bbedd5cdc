// capi.cpp -- the extern "C" boundary of libOpenCLIPP.so.
//
// Every reference entry point keeps its name, arguments, asynchrony and error codes
// (include/OpenCLIPP.h of the reference; C layer OpenCLIPP/OpenCLIPP.cpp:45-728).  Errors
// raised below the ABI are ocip::Error values; the ABI returns their code, as the
// reference's H() macro returns cl::Error::err() (OpenCLIPP/OpenCLIPP.cpp:82).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "ocip_internal.h"
#include "validation.h"

using namespace ocip;

namespace {

// Current context: process-global and unguarded, as in the reference
// (OpenCLIPP/OpenCLIPP.cpp:73-75).  The ABI is not thread-safe.
Context *g_current = nullptr;
// Contexts not yet released: an image released after its context is freed only returns its
// memory (the reference leaves that case undefined).
std::vector<Context *> g_live;

bool live(Context *c)
{
   for (Context *p : g_live)
      if (p == c) return true;
   return false;
}

bool verbose()
{
   static const bool v = getenv("OCIP_VERBOSE") != nullptr;
   return v;
}

template <class F> cl_int guarded(F &&f)
{
   try {
      f();
   } catch (const Error &e) {
      if (verbose()) fprintf(stderr, "ocip: %s (%d)\n", e.what, e.code);
      return e.code;
   } catch (const std::bad_alloc &) {
      return CL_OUT_OF_HOST_MEMORY;
   }
   return CL_SUCCESS;
}

Context &current()
{
   if (!g_current) fail(CL_INVALID_CONTEXT, "no current context");
   return *g_current;
}

void set_device(const Context &c) { hip_check(hipSetDevice(c.device), "hipSetDevice"); }

Image &img(ocipImage h)
{
   Image *p = reinterpret_cast<Image *>(h);
   if (!p || p->magic != kImageMagic) fail(CL_INVALID_MEM_OBJECT, "not an image");
   return *p;
}

Program &prog(ocipProgram h, int kind)
{
   Program *p = reinterpret_cast<Program *>(h);
   if (!p || p->magic != kProgramMagic || p->kind != kind) fail(CL_INVALID_PROGRAM, "bad program");
   return *p;
}

View view(const Image &i)
{
   return View{i.dev, (int)i.desc.Width, (int)i.desc.Height, (int)i.desc.Step,
               (int)i.desc.Channels, (int)i.desc.Type};
}

// Host -> device copy of an image; colour images go through their 3-channel staging copy
// (ColorImage::Send, C++/Image.cpp:262-275).
void upload(Context &c, Image &i)
{
   void *to = i.stage ? i.stage : i.dev;
   hip_check(hipMemcpyAsync(to, i.host, i.bytes, hipMemcpyHostToDevice, c.stream), "send");
   if (i.stage)
      launch_copy3to4(c.stream, i.desc.Type, i.stage, i.host_desc.Step, i.dev, i.desc.Step,
                      i.desc.Width, i.desc.Height);
}

// Device -> host copy, blocking (ColorImage::Read, C++/Image.cpp:251-259).
void download(Context &c, Image &i)
{
   if (i.stage)
      launch_copy4to3(c.stream, i.desc.Type, i.dev, i.desc.Step, i.stage, i.host_desc.Step,
                      i.desc.Width, i.desc.Height);
   hip_check(hipMemcpyAsync(i.host, i.stage ? i.stage : i.dev, i.bytes, hipMemcpyDeviceToHost,
                            c.stream), "read");
   hip_check(hipStreamSynchronize(c.stream), "read sync");
}

// Lazy host->device transfer of an input (reference kernel_helpers.h:135, Buffer.cpp:190-194)
void send_if_needed(Context &c, Image &i)
{
   if (i.in_device) return;
   if (i.host) upload(c, i);
   i.in_device = true;
}

// Images must belong to the current context: kernels run on its stream.
void same_context(Context &c, const Image &i)
{
   if (i.ctx != &c) fail(CL_INVALID_CONTEXT, "image from another context");
}

void check_same_size(const Image &a, const Image &b)
{
   if (a.desc.Width != b.desc.Width || a.desc.Height != b.desc.Height)
      fail(CL_INVALID_IMAGE_SIZE, "different image sizes");
}

void throw_if(cl_int code, const char *what)
{
   if (code != CL_SUCCESS) fail(code, what);
}

// ---- proximity driver ---------------------------------------------------------------------

// FP64 exclusive sum-of-squares integral of channel c of S into c.integ.
const double *sqr_integral_for(Context &c, const Image &S, int channel)
{
   View v = view(S);
   c.integ.ensure(excl_bytes(v.width, v.height));
   c.integ_aux.ensure(integral_aux_bytes(v.width, v.height));
   launch_integral(c.stream, v, channel, true, c.integ.ptr, 0, -1, c.integ_aux.ptr);
   return c.integ.as<double>();
}

int tpitch_for(int tw) { return (tw + 15) / 16 * 16; }

void run_proximity(int op, ocipImage hs, ocipImage ht, ocipImage hd, int precision)
{
   Context &c = current();
   Image &S = img(hs), &T = img(ht), &D = img(hd);
   throw_if(check::proximity(S.desc, T.desc, D.desc), "proximity arguments");
   throw_if(check::precision(S.desc, precision), "precision");
   same_context(c, S); same_context(c, T); same_context(c, D);
   set_device(c);
   send_if_needed(c, S);
   send_if_needed(c, T);

   const int tw = T.desc.Width, th = T.desc.Height, nch = S.desc.Channels;
   const int type = S.desc.Type;
   const bool norm = op == kSSDNorm || op == kCCNorm;
   // compute type of the staged template: double for F64, int32 for 8/16-bit integers
   // (exact integer sums), float otherwise.
   const int kind = type == SImage::F64 ? 1 : (elem_size(type) <= 2 ? 2 : 0);
   const int esz = kind == 1 ? 8 : 4;
   const int tpitch = tpitch_for(tw);
   c.tmpl.data.ensure((size_t)tpitch * (th > 0 ? th : 1) * esz);
   c.tmpl.sums.ensure(2 * sizeof(double));

   // SqrDistance on 8-bit sources is formed from the cross term and the window sums of
   // squares, which come from the integral as for the Norm ops
   const bool ssd8 = op == kSSD && (type == SImage::U8 || type == SImage::S8);
   // 8-bit integer images: exact sums on BF16 MFMA where the path applies
   if (precision == ocipPrecisionF32 && int_mfma_ok(op, view(S), view(T))) {
      c.stats.ensure(int_mfma_work_bytes(tw, th));
      launch_int_mfma(c.stream, op, view(S), view(T), c.tmpl.sums.as<double>(),
                      static_cast<float *>(D.dev), D.desc.Step, c.stats.ptr);
      D.in_device = true;
      return;
   }
   for (int ch = 0; ch < nch; ++ch) {
      const bool mfma = precision != ocipPrecisionF32;
      const bool need_integ = norm ? (mfma ? crosscorr_mfma_needs_integral(tw)
                                           : !prox_norm_from_tile(op, view(S), view(T)))
                                   : ssd8 && !mfma;
      const double *integ = need_integ ? sqr_integral_for(c, S, ch) : nullptr;
      if (mfma) {
         size_t wb = crosscorr_mfma_work_bytes(precision, S.desc.Width, S.desc.Height, tw, th);
         c.stats.ensure(wb > 0 ? wb : 1);
         launch_crosscorr_mfma(c.stream, precision, norm, view(S), view(T),
                               c.tmpl.sums.as<double>(), integ, (float *)D.dev, D.desc.Step,
                               c.stats.ptr);
      } else {
         launch_template_prep(c.stream, view(T), ch, tpitch, kind, c.tmpl.data.ptr,
                              c.tmpl.sums.as<double>());
         launch_proximity(c.stream, op, view(S), view(T), D.dev, D.desc.Step, c.tmpl.data.ptr,
                          tpitch, c.tmpl.sums.as<double>(), integ, ch);
      }
   }
   D.in_device = true;
}

// Gradients (extensions; F32 1-channel images) of
//   L = sum_i [ sum(G1_i * CrossCorr_Norm(S_i, T)) + sum(G2_i * SqrDistance(S_i, T)) ]
// for a batch of same-size sources sharing one template: dS_i and dT (summed over the batch);
// g1 / g2 null: that op absent; ds entries and dt may be null.
void run_backward_batch(uint n, const ocipImage *hs, ocipImage ht, const ocipImage *g1,
                        const ocipImage *g2, const ocipImage *hds, ocipImage hdt)
{
   Context &c = current();
   if (n == 0) return;
   if (!hs || (!g1 && !g2)) fail(CL_INVALID_VALUE, "null source or gradient array");
   auto f32_1 = [](const Image &i) {
      return i.desc.Type == SImage::F32 && i.desc.Channels == 1;
   };
   Image &T = img(ht);
   if (!f32_1(T)) fail(CL_INVALID_VALUE, "F32 1-channel images");
   if (T.desc.Width == 0 || T.desc.Height == 0) fail(CL_INVALID_VALUE, "empty template");
   same_context(c, T);
   Image *dT = hdt ? &img(hdt) : nullptr;
   if (dT) {
      if (!f32_1(*dT)) fail(CL_INVALID_VALUE, "F32 1-channel images");
      check_same_size(T, *dT);
      same_context(c, *dT);
   }
   std::vector<Image *> S(n), G1(n), G2(n), dS(n);
   for (uint i = 0; i < n; ++i) {
      S[i] = &img(hs[i]);
      if (!f32_1(*S[i])) fail(CL_INVALID_VALUE, "F32 1-channel images");
      if (T.desc.Width > S[i]->desc.Width || T.desc.Height > S[i]->desc.Height)
         fail(CL_INVALID_VALUE, "template larger than source");
      if (i > 0) check_same_size(*S[0], *S[i]);
      same_context(c, *S[i]);
      G1[i] = g1 ? &img(g1[i]) : nullptr;
      G2[i] = g2 ? &img(g2[i]) : nullptr;
      dS[i] = hds && hds[i] ? &img(hds[i]) : nullptr;
      for (Image *m : {G1[i], G2[i], dS[i]}) {
         if (!m) continue;
         if (!f32_1(*m)) fail(CL_INVALID_VALUE, "F32 1-channel images");
         check_same_size(*S[i], *m);
         same_context(c, *m);
      }
   }
   set_device(c);
   send_if_needed(c, T);
   std::vector<View> vs(n), v1(n), v2(n);
   std::vector<char *> ds(n);
   std::vector<int> dss(n);
   for (uint i = 0; i < n; ++i) {
      send_if_needed(c, *S[i]);
      vs[i] = view(*S[i]);
      if (G1[i]) { send_if_needed(c, *G1[i]); v1[i] = view(*G1[i]); }
      if (G2[i]) { send_if_needed(c, *G2[i]); v2[i] = view(*G2[i]); }
      ds[i] = dS[i] ? static_cast<char *>(dS[i]->dev) : nullptr;
      dss[i] = dS[i] ? (int)dS[i]->desc.Step : 0;
   }
   const int tw = T.desc.Width, th = T.desc.Height;
   const int W = S[0]->desc.Width, H = S[0]->desc.Height;
   const int tpitch = tpitch_for(tw);
   c.tmpl.data.ensure((size_t)tpitch * th * sizeof(float));
   c.tmpl.sums.ensure(2 * sizeof(double));
   c.integ_aux.ensure(integral_aux_bytes(W, H));
   launch_template_prep(c.stream, view(T), 0, tpitch, 0, c.tmpl.data.ptr, c.tmpl.sums.as<double>());
   c.work.ensure(backward_work_bytes((int)n, W, H, tw, th, tpitch));
   launch_backward_batch(c.stream, (int)n, vs.data(), g1 ? v1.data() : nullptr,
                         g2 ? v2.data() : nullptr, c.tmpl.data.as<float>(), tpitch, tw, th,
                         c.tmpl.sums.as<double>(), c.work.ptr, c.integ_aux.ptr, ds.data(),
                         dss.data(), dT ? static_cast<char *>(dT->dev) : nullptr,
                         dT ? (int)dT->desc.Step : 0);
   for (uint i = 0; i < n; ++i)
      if (dS[i]) dS[i]->in_device = true;
   if (dT) dT->in_device = true;
}

// One image, one op (mode 0: CrossCorr_Norm, 1: SqrDistance).
void run_backward(int mode, ocipImage hs, ocipImage ht, ocipImage hg, ocipImage hds,
                  ocipImage hdt)
{
   run_backward_batch(1, &hs, ht, mode == 0 ? &hg : nullptr, mode == 1 ? &hg : nullptr, &hds,
                      hdt);
}

// Many templates against one source (extensions).  F32 precision runs the plain path per
// template; the MFMA precisions run one recorded graph for the whole batch.
void run_batch(bool ssd, ocipImage hs, const ocipImage *ht, ocipImage *hd, uint count,
               int precision)
{
   Context &c = current();
   Image &S = img(hs);
   if (count > 0 && (!ht || !hd)) fail(CL_INVALID_VALUE, "null template or dest array");
   throw_if(check::precision(S.desc, precision), "precision");
   if (S.desc.Type != SImage::F32 || S.desc.Channels != 1)
      fail(CL_INVALID_VALUE, "batches need F32 1-channel images");
   same_context(c, S);
   for (uint i = 0; i < count; ++i) {
      Image &T = img(ht[i]), &D = img(hd[i]);
      throw_if(check::proximity(S.desc, T.desc, D.desc), "proximity arguments");
      if (T.desc.Width == 0 || T.desc.Height == 0) fail(CL_INVALID_VALUE, "empty template");
      same_context(c, T);
      same_context(c, D);
   }
   if (count == 0) return;
   if (precision == ocipPrecisionF32) {
      for (uint i = 0; i < count; ++i)
         run_proximity(ssd ? kSSDNorm : kCCNorm, hs, ht[i], hd[i], ocipPrecisionF32);
      return;
   }
   set_device(c);
   send_if_needed(c, S);
   std::vector<View> tv;
   std::vector<float *> dv;
   std::vector<int> ds;
   for (uint i = 0; i < count; ++i) {
      Image &T = img(ht[i]), &D = img(hd[i]);
      send_if_needed(c, T);
      tv.push_back(view(T));
      dv.push_back(static_cast<float *>(D.dev));
      ds.push_back((int)D.desc.Step);
   }
   if (!c.batch) c.batch = batch_cache_create();
   run_batch_mfma(c.stream, c.batch, c.num_cus, precision, ssd, view(S), tv, dv, ds);
   for (uint i = 0; i < count; ++i) img(hd[i]).in_device = true;
}

}  // namespace

void ocip::report_hip_error(hipError_t e, const char *what)
{
   if (verbose()) fprintf(stderr, "ocip: HIP error in %s: %s\n", what, hipGetErrorString(e));
}

void DevBuf::ensure(size_t n)
{
   if (n <= bytes && ptr) return;
   release();
   void *p = nullptr;
   hip_check(hipMalloc(&p, n ? n : 1), "workspace allocation");
   ptr = p;
   bytes = n;
}

void DevBuf::release()
{
   if (ptr) {
      // the buffer may still be read by queued kernels
      (void)hipDeviceSynchronize();
      (void)hipFree(ptr);
   }
   ptr = nullptr;
   bytes = 0;
}

// ======================================================================= extern "C"
extern "C" {

ocipError ocip_API ocipInitialize(ocipContext *ContextPtr, const char *PreferredPlatform,
                                  cl_device_type deviceType)
{
   (void)PreferredPlatform;
   return guarded([&] {
      if (!ContextPtr) fail(CL_INVALID_VALUE, "null context pointer");
      const cl_device_type gpu_ok = CL_DEVICE_TYPE_GPU | CL_DEVICE_TYPE_DEFAULT;
      if (deviceType != (cl_device_type)CL_DEVICE_TYPE_ALL && !(deviceType & gpu_ok))
         fail(CL_DEVICE_NOT_FOUND, "only GPU devices");
      int n = 0;
      if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) fail(CL_DEVICE_NOT_FOUND, "no GPU");
      Context *c = new Context();
      c->device = 0;  // devices[0], as the reference (C++/OpenCL.cpp:57-62)
      try {
         hip_check(hipSetDevice(c->device), "hipSetDevice");
         hipDeviceProp_t p;
         hip_check(hipGetDeviceProperties(&p, c->device), "properties");
         c->name = p.name;
         if (p.gcnArchName[0]) c->name += std::string(" (") + p.gcnArchName + ")";
         c->num_cus = p.multiProcessorCount;
         hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream");
         noalloc_scratch_init(c->device);
      } catch (...) {
         delete c;
         throw;
      }
      g_current = c;
      g_live.push_back(c);
      *ContextPtr = reinterpret_cast<ocipContext>(c);
   });
}

ocipError ocip_API ocipUninitialize(ocipContext Context)
{
   ocip::Context *c = reinterpret_cast<ocip::Context *>(Context);
   if (!c || c->magic != kContextMagic || !live(c)) return CL_INVALID_CONTEXT;
   if (c == g_current) g_current = nullptr;
   for (size_t k = 0; k < g_live.size(); ++k)
      if (g_live[k] == c) { g_live.erase(g_live.begin() + k); break; }
   return guarded([&] {
      (void)hipSetDevice(c->device);
      (void)hipStreamSynchronize(c->stream);
      c->tmpl.data.release();
      c->tmpl.sums.release();
      c->integ.release();
      c->integ_aux.release();
      c->stats.release();
      c->work.release();
      batch_cache_destroy(c->batch);
      c->batch = nullptr;
      (void)hipStreamDestroy(c->stream);
      c->magic = 0;
      delete c;
   });
}

ocipError ocip_API ocipChangeContext(ocipContext Context)
{
   ocip::Context *c = reinterpret_cast<ocip::Context *>(Context);
   if (!c || !live(c)) return CL_INVALID_CONTEXT;
   g_current = c;
   return CL_SUCCESS;
}

void ocip_API ocipSetCLFilesPath(const char *Path) { (void)Path; }

ocip_API const char *ocipGetErrorName(ocipError Error)
{
   switch (Error) {
   case CL_SUCCESS: return "SUCCESS";
   case CL_DEVICE_NOT_FOUND: return "CL_DEVICE_NOT_FOUND";
   case CL_DEVICE_NOT_AVAILABLE: return "CL_DEVICE_NOT_AVAILABLE";
   case CL_MEM_OBJECT_ALLOCATION_FAILURE: return "CL_MEM_OBJECT_ALLOCATION_FAILURE";
   case CL_OUT_OF_RESOURCES: return "CL_OUT_OF_RESOURCES";
   case CL_OUT_OF_HOST_MEMORY: return "CL_OUT_OF_HOST_MEMORY";
   case CL_IMAGE_FORMAT_NOT_SUPPORTED: return "CL_IMAGE_FORMAT_NOT_SUPPORTED";
   case CL_INVALID_VALUE: return "CL_INVALID_VALUE";
   case CL_INVALID_DEVICE_TYPE: return "CL_INVALID_DEVICE_TYPE";
   case CL_INVALID_CONTEXT: return "CL_INVALID_CONTEXT";
   case CL_INVALID_MEM_OBJECT: return "CL_INVALID_MEM_OBJECT";
   case CL_INVALID_IMAGE_SIZE: return "CL_INVALID_IMAGE_SIZE";
   case CL_INVALID_PROGRAM: return "CL_INVALID_PROGRAM";
   case CL_INVALID_OPERATION: return "CL_INVALID_OPERATION";
   case CL_INVALID_BUFFER_SIZE: return "CL_INVALID_BUFFER_SIZE";
   default: return "Unknown error";
   }
}

ocipError ocip_API ocipGetDeviceName(char *Name, uint BufferLength)
{
   if (!g_current) return CL_INVALID_CONTEXT;
   const std::string &n = g_current->name;
   if (!Name || BufferLength == 0 || n.size() > BufferLength - 1) return CL_OUT_OF_RESOURCES;
   memcpy(Name, n.c_str(), n.size() + 1);
   return CL_SUCCESS;
}

ocipError ocip_API ocipFinish(void)
{
   if (!g_current) return CL_INVALID_CONTEXT;
   return guarded([&] {
      set_device(*g_current);
      hip_check(hipStreamSynchronize(g_current->stream), "finish");
   });
}

ocipError ocip_API ocipCreateImage(ocipImage *ImagePtr, SImage Image, void *ImageData,
                                   cl_mem_flags flags)
{
   if (!g_current) return CL_INVALID_CONTEXT;
   return guarded([&] {
      if (!ImagePtr) fail(CL_INVALID_VALUE, "null image pointer");
      throw_if(check::create_image(Image), "image description");
      const cl_mem_flags known = CL_MEM_READ_WRITE | CL_MEM_WRITE_ONLY | CL_MEM_READ_ONLY;
      if (flags & ~known) fail(CL_INVALID_VALUE, "flags");
      size_t bytes = (size_t)Image.Height * Image.Step;
      set_device(*g_current);
      ocip::Image *i = new ocip::Image();
      i->ctx = g_current;
      i->desc = Image;
      i->host_desc = Image;
      i->host = ImageData;
      i->bytes = bytes;
      size_t dev_bytes = bytes;
      if (Image.Channels == 3) {
         // OpenCLIPP.cpp:185-188: a ColorImage, 4 channels on the device (Image.cpp:243-249)
         i->desc.Channels = 4;
         i->desc.Step = Image.Width * 4 * (uint)check::elem_bytes(Image.Type);
         dev_bytes = (size_t)Image.Height * i->desc.Step;
         if (hipMalloc(&i->stage, bytes) != hipSuccess) {
            delete i;
            fail(CL_MEM_OBJECT_ALLOCATION_FAILURE, "image allocation");
         }
      }
      if (hipMalloc(&i->dev, dev_bytes) != hipSuccess) {
         if (i->stage) (void)hipFree(i->stage);
         delete i;
         fail(CL_MEM_OBJECT_ALLOCATION_FAILURE, "image allocation");
      }
      *ImagePtr = reinterpret_cast<ocipImage>(i);
   });
}

ocipError ocip_API ocipSendImage(ocipImage Image)
{
   return guarded([&] {
      ocip::Image &i = img(Image);
      if (!i.host) return;  // device-only image (C++/Buffer.cpp:158-159)
      set_device(*i.ctx);
      upload(*i.ctx, i);
      i.in_device = true;
   });
}

ocipError ocip_API ocipReadImage(ocipImage Image)
{
   return guarded([&] {
      ocip::Image &i = img(Image);
      if (!i.host) return;  // device-only image (C++/Buffer.cpp:125-126)
      set_device(*i.ctx);
      download(*i.ctx, i);
   });
}

ocipError ocip_API ocipReleaseImage(ocipImage Image)
{
   return guarded([&] {
      ocip::Image &i = img(Image);
      if (live(i.ctx)) {
         (void)hipSetDevice(i.ctx->device);
         // queued work may still use the memory: wait for the context's stream
         (void)hipStreamSynchronize(i.ctx->stream);
      }
      (void)hipFree(i.dev);
      if (i.stage) (void)hipFree(i.stage);
      i.magic = 0;
      delete &i;
   });
}

ocipError ocip_API ocipReleaseProgram(ocipProgram Program)
{
   ocip::Program *p = reinterpret_cast<ocip::Program *>(Program);
   if (!p || p->magic != kProgramMagic) return CL_INVALID_PROGRAM;
   return guarded([&] {
      if (live(p->ctx)) {
         (void)hipSetDevice(p->ctx->device);
         (void)hipStreamSynchronize(p->ctx->stream);
      }
      p->scratch.release();
      if (p->fft) fft_state_destroy(p->fft);
      p->magic = 0;
      delete p;
   });
}

// ---- integral ---------------------------------------------------------------------------
static ocipError prepare_program(ocipProgram *ProgramPtr, ocipImage Image, int kind)
{
   if (!g_current) return CL_INVALID_CONTEXT;
   return guarded([&] {
      if (!ProgramPtr) fail(CL_INVALID_VALUE, "null program pointer");
      ocip::Program *p = new ocip::Program();
      p->ctx = g_current;
      p->kind = kind;
      *ProgramPtr = reinterpret_cast<ocipProgram>(p);
      if (Image) {
         ocip::Image &i = img(Image);
         set_device(*g_current);
         if (kind == kProgIntegral)
            p->scratch.ensure(integral_aux_bytes(i.desc.Width, i.desc.Height));
         else
            p->scratch.ensure(minmax_partials_bytes(i.desc.Width, i.desc.Height));
      }
   });
}

ocipError ocip_API ocipPrepareIntegral(ocipProgram *ProgramPtr, ocipImage Image)
{
   return prepare_program(ProgramPtr, Image, kProgIntegral);
}

// Integral.cpp:97-141 (dest type switch), then PrepareFor + CheckSameSize (:145-146).
static ocipError run_integral(ocipProgram Program, ocipImage Source, ocipImage Dest, bool sqr)
{
   return guarded([&] {
      ocip::Program &p = prog(Program, kProgIntegral);
      ocip::Image &S = img(Source), &D = img(Dest);
      throw_if(check::integral(S.desc, D.desc), "integral arguments");
      Context &c = *p.ctx;
      same_context(c, S); same_context(c, D);
      set_device(c);
      send_if_needed(c, S);
      p.scratch.ensure(integral_aux_bytes(S.desc.Width, S.desc.Height));
      launch_integral(c.stream, view(S), 0, sqr, D.dev, D.desc.Step, D.desc.Type, p.scratch.ptr);
      D.in_device = true;
   });
}

ocipError ocip_API ocipIntegral(ocipProgram Program, ocipImage Source, ocipImage Dest)
{
   return run_integral(Program, Source, Dest, false);
}

ocipError ocip_API ocipSqrIntegral(ocipProgram Program, ocipImage Source, ocipImage Dest)
{
   return run_integral(Program, Source, Dest, true);
}

// ---- proximity --------------------------------------------------------------------------
ocipError ocip_API ocipPrepareProximity(ocipImage Image)
{
   if (!g_current) return CL_INVALID_CONTEXT;
   // Nothing to prepare: kernels are compiled ahead of time and three-channel images are
   // already four channels on the device. Only the handle is checked.
   return guarded([&] { (void)img(Image); });
}

#define OCIP_PROX(name, op)                                                                   \
   ocipError ocip_API name(ocipImage Source, ocipImage Template, ocipImage Dest)             \
   {                                                                                          \
      return guarded([&] { run_proximity(op, Source, Template, Dest, ocipPrecisionF32); });  \
   }

OCIP_PROX(ocipSqrDistance_Norm, kSSDNorm)
OCIP_PROX(ocipSqrDistance, kSSD)
OCIP_PROX(ocipAbsDistance, kSAD)
OCIP_PROX(ocipCrossCorr, kCC)
OCIP_PROX(ocipCrossCorr_Norm, kCCNorm)

ocipError ocip_API ocipCrossCorrEx(ocipImage Source, ocipImage Template, ocipImage Dest,
                                   int Precision)
{
   return guarded([&] { run_proximity(kCC, Source, Template, Dest, Precision); });
}

ocipError ocip_API ocipCrossCorr_NormEx(ocipImage Source, ocipImage Template, ocipImage Dest,
                                        int Precision)
{
   return guarded([&] { run_proximity(kCCNorm, Source, Template, Dest, Precision); });
}

// ocipCrossCorr_NormEx without any allocation: MFMA precisions, templates whose fragments fit
// the per-GPU scratch array (noalloc.hip), no autotuning (the variant tuned earlier for the
// shape, else the row-pair kernel).  Overlapping calls on one GPU race on that array.
ocipError ocip_API ocipCrossCorr_NormExNoAlloc(ocipImage Source, ocipImage Template,
                                               ocipImage Dest, int Precision)
{
   return guarded([&] {
      Context &c = current();
      Image &S = img(Source), &T = img(Template), &D = img(Dest);
      throw_if(check::proximity(S.desc, T.desc, D.desc), "proximity arguments");
      throw_if(check::precision(S.desc, Precision), "precision");
      if (Precision == ocipPrecisionF32)
         fail(CL_INVALID_VALUE, "no-allocation entry point: MFMA precisions only");
      const int tw = T.desc.Width, th = T.desc.Height;
      if (tw == 0 || th == 0) fail(CL_INVALID_VALUE, "empty template");
      if (crosscorr_mfma_needs_integral(tw))
         fail(CL_INVALID_VALUE, "no-allocation entry point: template wider than 513 columns");
      const size_t need = kNoAllocSumsBytes + crosscorr_mfma_work_bytes(
                              Precision, S.desc.Width, S.desc.Height, tw, th);
      if (need > kNoAllocScratchBytes)
         fail(CL_INVALID_BUFFER_SIZE, "no-allocation entry point: template needs more than "
                                      "the 8 MiB scratch array");
      same_context(c, S); same_context(c, T); same_context(c, D);
      set_device(c);
      send_if_needed(c, S);
      send_if_needed(c, T);
      char *base = static_cast<char *>(noalloc_scratch(c.device));
      launch_crosscorr_mfma(c.stream, Precision, true, view(S), view(T),
                            reinterpret_cast<double *>(base), nullptr, static_cast<float *>(D.dev),
                            D.desc.Step, base + kNoAllocSumsBytes, false);
      D.in_device = true;
   });
}

ocipError ocip_API ocipCrossCorr_NormBatch(ocipImage Source, const ocipImage *Templates,
                                           ocipImage *Dests, uint Count, int Precision)
{
   return guarded([&] { run_batch(false, Source, Templates, Dests, Count, Precision); });
}

ocipError ocip_API ocipSqrDistance_NormBatch(ocipImage Source, const ocipImage *Templates,
                                             ocipImage *Dests, uint Count, int Precision)
{
   return guarded([&] { run_batch(true, Source, Templates, Dests, Count, Precision); });
}

ocipError ocip_API ocipCrossCorr_NormBackward(ocipImage Source, ocipImage Template,
                                              ocipImage GradDest, ocipImage GradSource,
                                              ocipImage GradTemplate)
{
   return guarded([&] { run_backward(0, Source, Template, GradDest, GradSource, GradTemplate); });
}

ocipError ocip_API ocipSqrDistanceBackward(ocipImage Source, ocipImage Template,
                                           ocipImage GradDest, ocipImage GradSource,
                                           ocipImage GradTemplate)
{
   return guarded([&] { run_backward(1, Source, Template, GradDest, GradSource, GradTemplate); });
}

ocipError ocip_API ocipProximityBackwardBatch(const ocipImage *Sources, ocipImage Template,
                                              const ocipImage *GradNcc, const ocipImage *GradSsd,
                                              ocipImage *GradSources, ocipImage GradTemplate,
                                              uint Count)
{
   return guarded([&] {
      run_backward_batch(Count, Sources, Template, GradNcc, GradSsd, GradSources, GradTemplate);
   });
}

// ---- statistics -------------------------------------------------------------------------
ocipError ocip_API ocipPrepareStatistics(ocipProgram *ProgramPtr, ocipImage Image)
{
   return prepare_program(ProgramPtr, Image, kProgStatistics);
}

static ocipError run_minmax(ocipProgram Program, ocipImage Source, double *Result, int *IndexX,
                            int *IndexY, bool want_max)
{
   if (!g_current) return CL_INVALID_CONTEXT;
   return guarded([&] {
      ocip::Image &S = img(Source);
      // Statistics.cpp:246 Check1Channel
      if (S.desc.Channels != 1) fail(CL_INVALID_VALUE, "only 1 channel images are allowed");
      if (!Result || !IndexX || !IndexY) fail(CL_INVALID_VALUE, "null result pointer");
      Context &c = Program ? *prog(Program, kProgStatistics).ctx : current();
      same_context(c, S);
      set_device(c);
      send_if_needed(c, S);
      DevBuf &buf = Program ? reinterpret_cast<ocip::Program *>(Program)->scratch : c.stats;
      buf.ensure(minmax_partials_bytes(S.desc.Width, S.desc.Height));
      launch_minmax(c.stream, want_max, view(S), buf.ptr, Result, IndexX, IndexY);
   });
}

ocipError ocip_API ocipMinIndx(ocipProgram Program, ocipImage Source, double *Result,
                               int *IndexX, int *IndexY)
{
   return run_minmax(Program, Source, Result, IndexX, IndexY, false);
}

ocipError ocip_API ocipMaxIndx(ocipProgram Program, ocipImage Source, double *Result,
                               int *IndexX, int *IndexY)
{
   return run_minmax(Program, Source, Result, IndexX, IndexY, true);
}

// Statistics::Sum / SumSqr (Statistics.cpp:350-370): every channel, no channel check.
static ocipError run_sum(ocipProgram Program, ocipImage Source, double *Result, bool sqr)
{
   if (!g_current) return CL_INVALID_CONTEXT;
   return guarded([&] {
      ocip::Image &S = img(Source);
      if (!Result) fail(CL_INVALID_VALUE, "null result pointer");
      Context &c = Program ? *prog(Program, kProgStatistics).ctx : current();
      same_context(c, S);
      set_device(c);
      send_if_needed(c, S);
      DevBuf &buf = Program ? reinterpret_cast<ocip::Program *>(Program)->scratch : c.stats;
      buf.ensure(sum_partials_bytes(S.desc.Width, S.desc.Height));
      launch_sum(c.stream, sqr, view(S), buf.ptr, Result);
   });
}

// ---- proximity through FFTs ------------------------------------------------------------------
// OpenCLIPP/OpenCLIPP.cpp:630-640: the program object owns the transform state; the image and
// template only size it (any later call re-sizes it for its own images).
ocipError ocip_API ocipPrepareImageProximityFFT(ocipProgram *ProgramPtr, ocipImage Image,
                                                ocipImage Template)
{
   if (!g_current) return CL_INVALID_CONTEXT;
   return guarded([&] {
      if (!ProgramPtr) fail(CL_INVALID_VALUE, "null program pointer");
      ocip::Program *p = new ocip::Program();
      p->ctx = g_current;
      p->kind = kProgProximityFFT;
      p->fft = fft_state_create();
      *ProgramPtr = reinterpret_cast<ocipProgram>(p);
      if (Image && Template) {
         ocip::Image &S = img(Image), &T = img(Template);
         int Wb, Hb;
         throw_if(fft_sizes(S.desc.Width, S.desc.Height, T.desc.Width, T.desc.Height, &Wb, &Hb),
                  "image too big for the FFT path");
         set_device(*g_current);
         fft_prepare(g_current->stream, p->fft, S.desc.Width, S.desc.Height, T.desc.Width,
                     T.desc.Height);
      }
   });
}

static ocipError run_proximity_fft(int op, ocipProgram Program, ocipImage Source,
                                   ocipImage Template, ocipImage Dest)
{
   if (!g_current) return CL_INVALID_CONTEXT;
   return guarded([&] {
      ocip::Program &p = prog(Program, kProgProximityFFT);
      Context &c = *p.ctx;
      Image &S = img(Source), &T = img(Template), &D = img(Dest);
      throw_if(check::proximity_fft(S.desc, T.desc, D.desc), "proximity FFT arguments");
      int Wb, Hb;
      throw_if(fft_sizes(S.desc.Width, S.desc.Height, T.desc.Width, T.desc.Height, &Wb, &Hb),
               "image too big for the FFT path");
      same_context(c, S); same_context(c, T); same_context(c, D);
      set_device(c);
      send_if_needed(c, S);
      send_if_needed(c, T);
      const double *integ = op == kCC ? nullptr : sqr_integral_for(c, S, 0);
      const int tpitch = tpitch_for(T.desc.Width);
      c.tmpl.data.ensure((size_t)tpitch * T.desc.Height * sizeof(float));
      c.tmpl.sums.ensure(2 * sizeof(double));
      launch_template_prep(c.stream, view(T), 0, tpitch, 0, c.tmpl.data.ptr, c.tmpl.sums.as<double>());
      launch_proximity_fft(c.stream, p.fft, op, view(S), view(T), integ, c.tmpl.sums.as<double>(),
                           D.dev, (int)D.desc.Step);
      D.in_device = true;
   });
}

ocipError ocip_API ocipSqrDistanceFFT(ocipProgram Program, ocipImage Source, ocipImage Template,
                                      ocipImage Dest)
{
   return run_proximity_fft(kSSD, Program, Source, Template, Dest);
}

ocipError ocip_API ocipSqrDistanceFFT_Norm(ocipProgram Program, ocipImage Source,
                                           ocipImage Template, ocipImage Dest)
{
   return run_proximity_fft(kSSDNorm, Program, Source, Template, Dest);
}

ocipError ocip_API ocipCrossCorrFFT(ocipProgram Program, ocipImage Source, ocipImage Template,
                                    ocipImage Dest)
{
   return run_proximity_fft(kCC, Program, Source, Template, Dest);
}

ocipError ocip_API ocipCrossCorrFFT_Norm(ocipProgram Program, ocipImage Source,
                                         ocipImage Template, ocipImage Dest)
{
   return run_proximity_fft(kCCNorm, Program, Source, Template, Dest);
}

// ---- generic FFT -----------------------------------------------------------------------------
// OpenCLIPP/OpenCLIPP.cpp:606-619, C++/programs/FFT.cpp:140-283.
static void check_fft(const Image &R, const Image &C)
{
   throw_if(check::fft(R.desc, C.desc), "FFT image types");
   throw_if(fft_plain_sizes(R.desc.Width, R.desc.Height), "FFT size");
   throw_if(check::fft_layout(R.desc, C.desc), "FFT complex image layout");
}

ocipError ocip_API ocipPrepareFFT(ocipProgram *ProgramPtr, ocipImage RealImage,
                                  ocipImage ComplexImage)
{
   if (!g_current) return CL_INVALID_CONTEXT;
   return guarded([&] {
      if (!ProgramPtr) fail(CL_INVALID_VALUE, "null program pointer");
      ocip::Program *p = new ocip::Program();
      p->ctx = g_current;
      p->kind = kProgFFT;
      p->fft = fft_state_create();
      *ProgramPtr = reinterpret_cast<ocipProgram>(p);
      if (RealImage && ComplexImage) {
         ocip::Image &R = img(RealImage), &C = img(ComplexImage);
         check_fft(R, C);
         set_device(*g_current);
         fft_plain_prepare(g_current->stream, p->fft, R.desc.Width, R.desc.Height);
      }
   });
}

static ocipError run_fft(ocipProgram Program, ocipImage Source, ocipImage Dest, bool forward)
{
   if (!g_current) return CL_INVALID_CONTEXT;
   return guarded([&] {
      ocip::Program &p = prog(Program, kProgFFT);
      Context &c = *p.ctx;
      Image &S = img(Source), &D = img(Dest);
      Image &R = forward ? S : D, &C = forward ? D : S;
      check_fft(R, C);
      same_context(c, S); same_context(c, D);
      set_device(c);
      send_if_needed(c, S);
      launch_fft(c.stream, p.fft, forward, view(R), view(C));
      D.in_device = true;
   });
}

ocipError ocip_API ocipFFTForward(ocipProgram Program, ocipImage RealSource, ocipImage ComplexDest)
{
   return run_fft(Program, RealSource, ComplexDest, true);
}

ocipError ocip_API ocipFFTInverse(ocipProgram Program, ocipImage ComplexSource, ocipImage RealDest)
{
   return run_fft(Program, ComplexSource, RealDest, false);
}

// OpenCLIPP/OpenCLIPP.cpp:621-625: this library always has its FFT path.
ocipBool ocip_API ocipIsFFTAvailable() { return 1; }

ocipError ocip_API ocipSum(ocipProgram Program, ocipImage Source, double *Result)
{
   return run_sum(Program, Source, Result, false);
}

ocipError ocip_API ocipSumSqr(ocipProgram Program, ocipImage Source, double *Result)
{
   return run_sum(Program, Source, Result, true);
}

}  // extern "C"

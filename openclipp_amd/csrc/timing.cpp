// timing.cpp -- HIP-event timing of the library's kernels (ocipEnableKernelTiming).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "ocip_internal.h"

namespace ocip {

namespace {
struct Record {
   std::string name;
   hipEvent_t start, stop;
};
bool g_enabled = false;
bool g_suspended = false;
std::vector<Record> g_records;

void clear_records()
{
   for (Record &r : g_records) {
      (void)hipEventDestroy(r.start);
      (void)hipEventDestroy(r.stop);
   }
   g_records.clear();
}
}  // namespace

void timing_suspend(bool suspend) { g_suspended = suspend; }

KernelTimer::KernelTimer(hipStream_t s, const char *n) : stream(s), name(n)
{
   if (!g_enabled || g_suspended) return;
   hipEvent_t e;
   if (hipEventCreate(&e) != hipSuccess) return;
   if (hipEventRecord(e, s) != hipSuccess) {
      (void)hipEventDestroy(e);
      return;
   }
   start = e;
}

KernelTimer::~KernelTimer()
{
   if (!start) return;
   hipEvent_t e;
   if (hipEventCreate(&e) != hipSuccess || hipEventRecord(e, stream) != hipSuccess) {
      (void)hipEventDestroy(static_cast<hipEvent_t>(start));
      return;
   }
   g_records.push_back(Record{name, static_cast<hipEvent_t>(start), e});
}

}  // namespace ocip

extern "C" {

ocipError ocip_API ocipEnableKernelTiming(ocipBool Enable)
{
   ocip::clear_records();
   ocip::g_enabled = Enable != 0;
   return CL_SUCCESS;
}

ocipError ocip_API ocipGetKernelTime(const char *KernelName, double *TotalMs, uint *Count)
{
   if (!KernelName || !TotalMs || !Count) return CL_INVALID_VALUE;
   double total = 0;
   uint n = 0;
   const size_t len = strlen(KernelName);
   for (ocip::Record &r : ocip::g_records) {
      if (r.name.compare(0, len, KernelName) != 0) continue;
      if (hipEventSynchronize(r.stop) != hipSuccess) return CL_OUT_OF_RESOURCES;
      float ms = 0;
      if (hipEventElapsedTime(&ms, r.start, r.stop) != hipSuccess) return CL_OUT_OF_RESOURCES;
      total += ms;
      ++n;
   }
   *TotalMs = total;
   *Count = n;
   return CL_SUCCESS;
}

}  // extern "C"

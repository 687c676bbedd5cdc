// noalloc.hip -- the device scratch of ocipCrossCorr_NormExNoAlloc, the entry point that
// allocates nothing (for callers that cannot call hipMalloc: serving loops, code under graph
// capture).
//
// `scratch` is a file-scope device array of kNoAllocScratchBytes: the template's MFMA
// fragments and its FP64 sum of squares live there instead of in buffers the library
// allocates.  Every GPU the library runs on holds its own copy, at its own address, until the
// process ends.  In host code the array's name is only a host-side stand-in: its device
// address is taken once per device, with that device current, by hipGetSymbolAddress at
// context creation (noalloc_scratch_init), and host code uses that address alone.  Every call
// on one GPU shares that GPU's array and leaves the last call's contents in it, so calls that
// overlap on one GPU (two streams, two threads, two contexts on one device) race.
#include "ocip_internal.h"

namespace ocip {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices];   // device address of `scratch` on each device, once taken
}  // namespace

void noalloc_scratch_init(int device)
{
   if (device < 0 || device >= kMaxDevices) fail(CL_DEVICE_NOT_FOUND, "device index");
   if (g_scratch[device]) return;
   void *p = nullptr;
   hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch symbol address");
   if (!p) fail(CL_OUT_OF_RESOURCES, "scratch symbol address");
   g_scratch[device] = p;
}

void *noalloc_scratch(int device)
{
   if (device < 0 || device >= kMaxDevices || !g_scratch[device])
      fail(CL_INVALID_CONTEXT, "no scratch address for this device");
   return g_scratch[device];
}

}  // namespace ocip

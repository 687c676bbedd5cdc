// color.hip -- 3-channel (colour) images: the caller's 3-channel rows and the 4-channel device
// layout the kernels see (reference ColorImage with the copy3Cto4C / copy4Cto3C kernels,
// C++/Image.cpp:243-275, cl-files/Convert.cl:225-258).  The fourth channel is set to 255 when
// sending, as the reference does, and dropped when reading back.
#include "device_common.h"

namespace ocip {

namespace {

template <int TYPE>
__global__ __launch_bounds__(256) void copy3to4(const char *__restrict__ src, int sstep,
                                                char *__restrict__ dst, int dstep, int W, int H)
{
   typedef typename ElemT<TYPE>::type T;
   const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
   if (k >= (long long)W * H) return;
   const int y = (int)(k / W), x = (int)(k - (long long)y * W);
   const T *s = reinterpret_cast<const T *>(src + (size_t)y * sstep) + 3 * x;
   T *d = reinterpret_cast<T *>(dst + (size_t)y * dstep) + 4 * x;
   d[0] = s[0];
   d[1] = s[1];
   d[2] = s[2];
   d[3] = (T)255;
}

template <int TYPE>
__global__ __launch_bounds__(256) void copy4to3(const char *__restrict__ src, int sstep,
                                                char *__restrict__ dst, int dstep, int W, int H)
{
   typedef typename ElemT<TYPE>::type T;
   const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
   if (k >= (long long)W * H) return;
   const int y = (int)(k / W), x = (int)(k - (long long)y * W);
   const T *s = reinterpret_cast<const T *>(src + (size_t)y * sstep) + 4 * x;
   T *d = reinterpret_cast<T *>(dst + (size_t)y * dstep) + 3 * x;
   d[0] = s[0];
   d[1] = s[1];
   d[2] = s[2];
}

unsigned blocks_for(int W, int H) { return (unsigned)(((long long)W * H + 255) / 256); }

}  // namespace

void launch_copy3to4(hipStream_t s, int type, const void *src, int sstep, void *dst, int dstep,
                     int W, int H)
{
   if ((long long)W * H == 0) return;
   OCIP_TYPE_SWITCH(type, TT, {
      copy3to4<TT><<<blocks_for(W, H), 256, 0, s>>>(static_cast<const char *>(src), sstep,
                                                      static_cast<char *>(dst), dstep, W, H);
   });
   launch_check("copy3to4");
}

void launch_copy4to3(hipStream_t s, int type, const void *src, int sstep, void *dst, int dstep,
                     int W, int H)
{
   if ((long long)W * H == 0) return;
   OCIP_TYPE_SWITCH(type, TT, {
      copy4to3<TT><<<blocks_for(W, H), 256, 0, s>>>(static_cast<const char *>(src), sstep,
                                                      static_cast<char *>(dst), dstep, W, H);
   });
   launch_check("copy4to3");
}

}  // namespace ocip

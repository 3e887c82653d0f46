// runtime.hip -- device, memory, stream and event plumbing behind the C ABI.
//
// Replaces the CUDA runtime use of the reference device: the single stream of
// DeviceGlobalState (optix_visrtx.h:165), the frame's start/end events and
// completion callback (frame/Frame.cu:49-53,255,295-307), and the thrust
// fills used for frame clears (frame/Frame.cu:590-660) -- the latter as a
// plain fill kernel so that it can be recorded into a HIP graph.
#include <stdarg.h>
#include <string.h>

#include <algorithm>
#include <atomic>

#include "vrx_internal.h"

namespace vrx {

int device_cus()
{
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return 256;
  if (dev >= 0 && dev < 64) {
    const int n = cache[dev].load(std::memory_order_relaxed);
    if (n > 0)
      return n;
  }
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess || p.multiProcessorCount < 1)
    return 256;
  if (dev >= 0 && dev < 64)
    cache[dev].store(p.multiProcessorCount, std::memory_order_relaxed);
  return p.multiProcessorCount;
}


static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
  set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  return e == hipErrorOutOfMemory ? VRX_ERR_NOMEM : VRX_ERR_HIP;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride)
    p[i] = v;
}

__global__ void k_copy_u32(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n16; i += stride)
    dst[i] = src[i];
}

} // namespace vrx

using namespace vrx;

extern "C" {

const char *vrx_last_error(void) { return g_err; }

int vrx_version(void) { return 100; }

int vrx_device_count(int *count)
{
  VRX_ARG(count, "count");
  VRX_HIP(hipGetDeviceCount(count));
  return VRX_OK;
}

int vrx_set_device(int device)
{
  VRX_HIP(hipSetDevice(device));
  return VRX_OK;
}

int vrx_get_device(int *device)
{
  VRX_ARG(device, "device");
  VRX_HIP(hipGetDevice(device));
  return VRX_OK;
}

int vrx_device_info(int device, char *name, size_t name_len, int *num_cus, size_t *hbm_bytes)
{
  hipDeviceProp_t p;
  VRX_HIP(hipGetDeviceProperties(&p, device));
  if (name && name_len) {
    snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (num_cus)
    *num_cus = p.multiProcessorCount;
  if (hbm_bytes)
    *hbm_bytes = p.totalGlobalMem;
  return VRX_OK;
}

int vrx_malloc(void **d_ptr, size_t bytes)
{
  VRX_ARG(d_ptr, "d_ptr");
  *d_ptr = nullptr;
  if (bytes == 0)
    return VRX_OK;
  VRX_HIP(hipMalloc(d_ptr, bytes));
  return VRX_OK;
}

int vrx_free(void *d_ptr)
{
  if (d_ptr)
    VRX_HIP(hipFree(d_ptr));
  return VRX_OK;
}

int vrx_host_device_pointer(void *h_ptr, void **d_ptr)
{
  VRX_ARG(h_ptr && d_ptr, "pointers");
  VRX_HIP(hipHostGetDevicePointer(d_ptr, h_ptr, 0));
  return VRX_OK;
}

int vrx_host_malloc(void **h_ptr, size_t bytes)
{
  VRX_ARG(h_ptr, "h_ptr");
  *h_ptr = nullptr;
  if (bytes == 0)
    return VRX_OK;
  VRX_HIP(hipHostMalloc(h_ptr, bytes, hipHostMallocDefault));
  return VRX_OK;
}

int vrx_host_free(void *h_ptr)
{
  if (h_ptr)
    VRX_HIP(hipHostFree(h_ptr));
  return VRX_OK;
}

int vrx_pointer_kind(const void *p, int *kind)
{
  VRX_ARG(kind, "kind");
  *kind = VRX_PTR_HOST;
  if (!p)
    return VRX_OK;
  hipPointerAttribute_t a;
  memset(&a, 0, sizeof(a));
  const hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    // plain pageable host memory is unknown to the runtime: clear the error
    // so that the next call does not report it
    (void)hipGetLastError();
    return VRX_OK;
  }
  int cur = 0;
  VRX_HIP(hipGetDevice(&cur));
  if (a.type == hipMemoryTypeManaged || a.isManaged)
    *kind = VRX_PTR_MANAGED;
  else if (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeArray)
    *kind = a.device == cur ? VRX_PTR_DEVICE : VRX_PTR_FOREIGN;
  return VRX_OK;
}

static hipMemcpyKind kind_of(int kind)
{
  return kind == VRX_H2D ? hipMemcpyHostToDevice
                         : (kind == VRX_D2H ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
}

int vrx_memcpy(void *dst, const void *src, size_t bytes, int kind, vrx_stream s)
{
  if (bytes == 0)
    return VRX_OK;
  VRX_ARG(dst && src, "dst/src");
  VRX_HIP(hipMemcpyAsync(dst, src, bytes, kind_of(kind), (hipStream_t)s));
  return VRX_OK;
}

int vrx_memcpy_sync(void *dst, const void *src, size_t bytes, int kind)
{
  if (bytes == 0)
    return VRX_OK;
  VRX_ARG(dst && src, "dst/src");
  VRX_HIP(hipMemcpy(dst, src, bytes, kind_of(kind)));
  return VRX_OK;
}

int vrx_fill_u32(void *d_ptr, uint32_t value, size_t count, vrx_stream s)
{
  if (count == 0)
    return VRX_OK;
  VRX_ARG(d_ptr, "d_ptr");
  size_t blocks = (count + 255) / 256;
  if (blocks > 65536)
    blocks = 65536;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)s,
      (uint32_t *)d_ptr, value, count);
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

int vrx_copy_kernel(void *d_dst, const void *d_src, size_t bytes, vrx_stream s)
{
  VRX_ARG(d_dst && d_src, "dst/src");
  VRX_ARG(bytes % 16 == 0, "bytes must be a multiple of 16");
  size_t n16 = bytes / 16;
  if (n16 == 0)
    return VRX_OK;
  size_t blocks = std::min<size_t>((n16 + 255) / 256, 256 * 64);
  hipLaunchKernelGGL(k_copy_u32, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)s,
      (const uint4 *)d_src, (uint4 *)d_dst, n16);
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

int vrx_stream_create(vrx_stream *s)
{
  VRX_ARG(s, "s");
  hipStream_t st;
  VRX_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  *s = (vrx_stream)st;
  return VRX_OK;
}

int vrx_stream_destroy(vrx_stream s)
{
  if (s)
    VRX_HIP(hipStreamDestroy((hipStream_t)s));
  return VRX_OK;
}

int vrx_stream_sync(vrx_stream s)
{
  VRX_HIP(hipStreamSynchronize((hipStream_t)s));
  return VRX_OK;
}

int vrx_device_sync(void)
{
  VRX_HIP(hipDeviceSynchronize());
  return VRX_OK;
}

int vrx_event_create(vrx_event *e)
{
  VRX_ARG(e, "e");
  hipEvent_t ev;
  VRX_HIP(hipEventCreate(&ev));
  *e = (vrx_event)ev;
  return VRX_OK;
}

int vrx_event_destroy(vrx_event e)
{
  if (e)
    VRX_HIP(hipEventDestroy((hipEvent_t)e));
  return VRX_OK;
}

int vrx_event_record(vrx_event e, vrx_stream s)
{
  VRX_HIP(hipEventRecord((hipEvent_t)e, (hipStream_t)s));
  return VRX_OK;
}

int vrx_event_sync(vrx_event e)
{
  VRX_HIP(hipEventSynchronize((hipEvent_t)e));
  return VRX_OK;
}

int vrx_event_query(vrx_event e)
{
  hipError_t r = hipEventQuery((hipEvent_t)e);
  if (r == hipSuccess)
    return 1;
  if (r == hipErrorNotReady)
    return 0;
  return hip_fail(r, "hipEventQuery", __FILE__, __LINE__);
}

int vrx_event_elapsed_ms(vrx_event start, vrx_event end, float *ms)
{
  VRX_ARG(ms, "ms");
  VRX_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
  return VRX_OK;
}

int vrx_launch_host_func(vrx_stream s, void (*fn)(void *), void *user)
{
  VRX_ARG(fn, "fn");
  VRX_HIP(hipLaunchHostFunc((hipStream_t)s, fn, user));
  return VRX_OK;
}

} // extern "C"

// Runtime part of the C ABI: device selection, memory, streams, events and
// graph capture.  These replace the ctypes wrapper around libamdhip64 in
// pyfr/backends/hip/driver.py:155-482.
#include "frhip_common.h"

#include <cstring>

namespace frhip {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

__global__ void k_fill_f64(double *p, long n, double v)
{
    const long i = (long)blockIdx.x*blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace frhip

using namespace frhip;

extern "C" {

const char *frhip_last_error(void) { return g_last_error.c_str(); }

int frhip_init(int device)
{
    FR_API_BEGIN
    int n = 0;
    FR_HIP(hipGetDeviceCount(&n));
    FR_CHECK(device >= 0 && device < n, "invalid device id");
    FR_HIP(hipSetDevice(device));
    FR_HIP(hipFree(nullptr));
    FR_HIP((hipError_t)red_scratch_init(device));
    return 0;
    FR_API_END
}

int frhip_device_info(char *name, size_t len, size_t *free_b, size_t *total_b)
{
    FR_API_BEGIN
    int dev = 0;
    FR_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    FR_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && len) {
        std::snprintf(name, len, "%s (%s, %d CUs)", prop.name,
                      prop.gcnArchName, prop.multiProcessorCount);
    }
    size_t f = 0, t = 0;
    FR_HIP(hipMemGetInfo(&f, &t));
    if (free_b) *free_b = f;
    if (total_b) *total_b = t;
    return 0;
    FR_API_END
}

int frhip_malloc(void **ptr, size_t nbytes)
{
    FR_API_BEGIN
    FR_CHECK(ptr, "null output pointer");
    *ptr = nullptr;
    FR_HIP(hipMalloc(ptr, nbytes ? nbytes : 256));
    return 0;
    FR_API_END
}

int frhip_free(void *ptr)
{
    FR_API_BEGIN
    if (ptr) FR_HIP(hipFree(ptr));
    return 0;
    FR_API_END
}

int frhip_host_malloc(void **ptr, size_t nbytes)
{
    FR_API_BEGIN
    FR_CHECK(ptr, "null output pointer");
    FR_HIP(hipHostMalloc(ptr, nbytes ? nbytes : 256, hipHostMallocDefault));
    return 0;
    FR_API_END
}

int frhip_host_free(void *ptr)
{
    FR_API_BEGIN
    if (ptr) FR_HIP(hipHostFree(ptr));
    return 0;
    FR_API_END
}

int frhip_memcpy_htod(void *dst, const void *src, size_t nbytes, void *stream)
{
    FR_API_BEGIN
    FR_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyHostToDevice,
                          as_stream(stream)));
    FR_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
    FR_API_END
}

int frhip_memcpy_dtoh(void *dst, const void *src, size_t nbytes, void *stream)
{
    FR_API_BEGIN
    FR_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToHost,
                          as_stream(stream)));
    FR_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
    FR_API_END
}

int frhip_fill_f64(void *ptr, long n, double v, void *stream)
{
    FR_API_BEGIN
    FR_CHECK(ptr && n >= 0, "invalid fill arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_fill_f64, dim3(cdiv(n, 256)), dim3(256), 0,
                       as_stream(stream), (double *)ptr, n, v);
    FR_LAUNCHED();
    return 0;
    FR_API_END
}

int frhip_memcpy_dtoh_async(void *dst, const void *src, size_t nbytes,
                            void *stream)
{
    FR_API_BEGIN
    FR_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToHost,
                          as_stream(stream)));
    return 0;
    FR_API_END
}

int frhip_memcpy_dtod(void *dst, const void *src, size_t nbytes, void *stream)
{
    FR_API_BEGIN
    FR_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToDevice,
                          as_stream(stream)));
    return 0;
    FR_API_END
}

int frhip_memset(void *ptr, int value, size_t nbytes, void *stream)
{
    FR_API_BEGIN
    FR_HIP(hipMemsetAsync(ptr, value, nbytes, as_stream(stream)));
    return 0;
    FR_API_END
}

int frhip_memcpy2d(void *dst, size_t dpitch, const void *src, size_t spitch,
                   size_t width, size_t height, int kind, void *stream)
{
    FR_API_BEGIN
    FR_CHECK(kind >= 0 && kind <= 2, "invalid copy kind");
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost
                                      : hipMemcpyDeviceToDevice;
    FR_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, k,
                            as_stream(stream)));
    if (kind != 2) FR_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
    FR_API_END
}

int frhip_stream_create(void **stream)
{
    FR_API_BEGIN
    hipStream_t s;
    FR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return 0;
    FR_API_END
}

int frhip_stream_destroy(void *stream)
{
    FR_API_BEGIN
    if (stream) FR_HIP(hipStreamDestroy(as_stream(stream)));
    return 0;
    FR_API_END
}

int frhip_stream_sync(void *stream)
{
    FR_API_BEGIN
    FR_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
    FR_API_END
}

int frhip_device_sync(void)
{
    FR_API_BEGIN
    FR_HIP(hipDeviceSynchronize());
    return 0;
    FR_API_END
}

int frhip_event_create(void **event)
{
    FR_API_BEGIN
    hipEvent_t e;
    FR_HIP(hipEventCreate(&e));
    *event = e;
    return 0;
    FR_API_END
}

int frhip_event_destroy(void *event)
{
    FR_API_BEGIN
    if (event) FR_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return 0;
    FR_API_END
}

int frhip_stream_wait_event(void *stream, void *event)
{
    FR_API_BEGIN
    FR_HIP(hipStreamWaitEvent(as_stream(stream),
                              static_cast<hipEvent_t>(event), 0));
    return 0;
    FR_API_END
}

int frhip_event_record(void *event, void *stream)
{
    FR_API_BEGIN
    FR_HIP(hipEventRecord(static_cast<hipEvent_t>(event), as_stream(stream)));
    return 0;
    FR_API_END
}

int frhip_event_sync(void *event)
{
    FR_API_BEGIN
    FR_HIP(hipEventSynchronize(static_cast<hipEvent_t>(event)));
    return 0;
    FR_API_END
}

int frhip_event_query(void *event, int *done)
{
    FR_API_BEGIN
    hipError_t e = hipEventQuery(static_cast<hipEvent_t>(event));
    if (e == hipErrorNotReady) {
        *done = 0;
        return 0;
    }
    FR_HIP(e);
    *done = 1;
    return 0;
    FR_API_END
}

int frhip_event_elapsed(float *ms, void *start, void *end)
{
    FR_API_BEGIN
    FR_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                               static_cast<hipEvent_t>(end)));
    return 0;
    FR_API_END
}

int frhip_graph_begin(void *stream)
{
    FR_API_BEGIN
    FR_CHECK(stream, "graph capture needs a non-default stream");
    FR_HIP(hipStreamBeginCapture(as_stream(stream),
                                 hipStreamCaptureModeThreadLocal));
    return 0;
    FR_API_END
}

int frhip_graph_end(void *stream, void **graph_exec)
{
    FR_API_BEGIN
    hipGraph_t g;
    FR_HIP(hipStreamEndCapture(as_stream(stream), &g));
    hipGraphExec_t ge;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    FR_HIP(e);
    *graph_exec = ge;
    return 0;
    FR_API_END
}

int frhip_graph_launch(void *graph_exec, void *stream)
{
    FR_API_BEGIN
    FR_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec),
                          as_stream(stream)));
    return 0;
    FR_API_END
}

int frhip_graph_destroy(void *graph_exec)
{
    FR_API_BEGIN
    if (graph_exec)
        FR_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
    return 0;
    FR_API_END
}

}  // extern "C"

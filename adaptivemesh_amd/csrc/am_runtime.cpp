// Error reporting, geometry set-up and thin device-memory helpers of the C ABI.
#include <stdarg.h>
#include <stdio.h>

#include <string>

#include "am_common.hpp"

namespace am {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int check_launch(const char *what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error((int)e, "%s: %s", what, hipGetErrorString(e));
    return AM_OK;
}

DepthScales depth_scales(const am_geometry &g)
{
    DepthScales s;
    for (int d = 0; d < 9; d++) {
        s.inv_dx[d] = (float)(1u << d) / g.outer_dx;
        s.dt[d] = g.outer_dt / (float)(1u << d);
    }
    return s;
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

static int hip_status(hipError_t e, const char *what)
{
    if (e != hipSuccess) return set_error((int)e, "%s: %s", what, hipGetErrorString(e));
    return AM_OK;
}

}  // namespace am

extern "C" {

const char *am_last_error_string(void) { return am::g_last_error.c_str(); }
int am_version(void) { return 100; }

int am_geometry_init(am_geometry *g, int32_t S, int32_t P, int32_t fd_order, float outer_size, float cfl)
{
    if (!g) return am::set_error(AM_ERR_INVALID, "am_geometry_init: null geometry");
    if (S <= 0 || P <= 0 || (S & 1)) return am::set_error(AM_ERR_INVALID, "am_geometry_init: bad S=%d P=%d", S, P);
    if (fd_order <= 0) fd_order = 2 * P;
    if ((fd_order & 1) || fd_order > 2 * P || fd_order > 6)
        return am::set_error(AM_ERR_INVALID, "am_geometry_init: fd_order %d must be even and <= min(2P, 6)", fd_order);
    g->size_domain = S;
    g->padding_domain = P;
    g->fd_order = fd_order;
    g->outer_size = outer_size;
    /* constexpr float outer_dx = outer_size / size_domain (AMR.h:16) */
    g->outer_dx = outer_size / (float)(uint32_t)S;
    /* constexpr float outer_dt = outer_dx * .4f (AMR.h:17), CFL as a parameter */
    g->outer_dt = g->outer_dx * cfl;
    return AM_OK;
}

int am_geometry_supported(const am_geometry *g)
{
    if (!g) return 0;
    const int s = g->size_domain, p = g->padding_domain, q = g->fd_order / 2;
    if (s == 16 && p == 3 && q >= 1 && q <= 3) return 1;
    if ((s == 16 || s == 28) && p == 2 && q >= 1 && q <= 2) return 1;
    return 0;
}

int am_device_alloc(void **ptr, size_t bytes)
{
    if (!ptr) return am::set_error(AM_ERR_INVALID, "am_device_alloc: null ptr");
    const hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return am::set_error(AM_ERR_ALLOC, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return AM_OK;
}
int am_device_free(void *ptr) { return am::hip_status(hipFree(ptr), "hipFree"); }
int am_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
    return am::hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream),
                          "hipMemcpyAsync H2D");
}
int am_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
    int rc = am::hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream),
                            "hipMemcpyAsync D2H");
    if (rc) return rc;
    return am::hip_status(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}
int am_memset_async(void *dst, int value, size_t bytes, void *stream)
{
    return am::hip_status(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream), "hipMemsetAsync");
}
int am_stream_sync(void *stream) { return am::hip_status(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize"); }
int am_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream)
{
    return am::hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream),
                          "hipMemcpyAsync(d2d)");
}

int am_stream_create(void **stream)
{
    if (!stream) return am::set_error(AM_ERR_INVALID, "am_stream_create: null output");
    hipStream_t st = nullptr;
    if (int rc = am::hip_status(hipStreamCreate(&st), "hipStreamCreate")) return rc;
    *stream = (void *)st;
    return AM_OK;
}

int am_stream_destroy(void *stream)
{
    if (!stream) return AM_OK;
    return am::hip_status(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
}

int am_graph_begin(void *stream)
{
    if (!stream) return am::set_error(AM_ERR_INVALID, "am_graph_begin: capture needs a non-default stream");
    return am::hip_status(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal),
                          "hipStreamBeginCapture");
}

int am_graph_end(void *stream, void **exec)
{
    if (!exec) return am::set_error(AM_ERR_INVALID, "am_graph_end: null output");
    hipGraph_t g = nullptr;
    if (int rc = am::hip_status(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture")) return rc;
    hipGraphExec_t e = nullptr;
    const int rc = am::hip_status(hipGraphInstantiate(&e, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    hipGraphDestroy(g);
    if (rc) return rc;
    *exec = (void *)e;
    return AM_OK;
}

int am_graph_launch(void *exec, void *stream)
{
    if (!exec) return am::set_error(AM_ERR_INVALID, "am_graph_launch: null graph");
    return am::hip_status(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream), "hipGraphLaunch");
}

int am_graph_destroy(void *exec)
{
    if (!exec) return AM_OK;
    return am::hip_status(hipGraphExecDestroy((hipGraphExec_t)exec), "hipGraphExecDestroy");
}

int am_device_sync(void) { return am::hip_status(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

int32_t am_word_bytes(int32_t format)
{
    switch (format) {
    case AM_WORD_CF3: return 4;
    case AM_WORD_F32: return 12;
    case AM_WORD_F16: return 6;
    default: return -1;
    }
}

int32_t am_device_cu_count(void)
{
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return n;
}

int am_event_create(void **ev)
{
    if (!ev) return am::set_error(AM_ERR_INVALID, "am_event_create: null");
    return am::hip_status(hipEventCreate((hipEvent_t *)ev), "hipEventCreate");
}
int am_event_record(void *ev, void *stream)
{
    return am::hip_status(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream), "hipEventRecord");
}
int am_event_elapsed_ms(void *start, void *stop, float *ms)
{
    if (!ms) return am::set_error(AM_ERR_INVALID, "am_event_elapsed_ms: null");
    if (int rc = am::hip_status(hipEventSynchronize((hipEvent_t)stop), "hipEventSynchronize")) return rc;
    return am::hip_status(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
}
int am_event_destroy(void *ev) { return am::hip_status(hipEventDestroy((hipEvent_t)ev), "hipEventDestroy"); }
int am_event_create_ordering(void **ev)
{
    if (!ev) return am::set_error(AM_ERR_INVALID, "am_event_create_ordering: null");
    return am::hip_status(hipEventCreateWithFlags((hipEvent_t *)ev, hipEventDisableTiming),
                          "hipEventCreateWithFlags");
}
int am_stream_wait_event(void *stream, void *ev)
{
    if (!ev) return am::set_error(AM_ERR_INVALID, "am_stream_wait_event: null event");
    return am::hip_status(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)ev, 0), "hipStreamWaitEvent");
}

}  // extern "C"

// runtime.cpp — device context, error reporting and device-memory entry points.
// Replaces the roles of CudaContext (src/gpu/CudaContext.cpp:83-256),
// DeviceMemoryPool (src/gpu/DeviceMemoryPool.hpp:64-104) and HostDeviceMirror
// (src/gpu/HostDeviceMirror.hpp:121-151).
#include "mi_common.hpp"

#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace mi
{

namespace
{
thread_local char g_err[1024] = "";
std::once_flag g_init_once;
int g_init_rc = MI_OK;
int g_device = 0;
hipStream_t g_stream = nullptr;
} // namespace

int set_error(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

void clear_error() { g_err[0] = '\0'; }

static int init_device(int dev)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return set_error(MI_ERR_NO_DEVICE, "no HIP device available (%s)",
                         e == hipSuccess ? "count 0" : hipGetErrorString(e));
    if (dev < 0 || dev >= count)
        return set_error(MI_ERR_BAD_ARG, "device %d out of range [0,%d)", dev, count);
    hipDeviceProp_t prop;
    MI_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(MI_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950", dev,
                         prop.gcnArchName);
    MI_HIP(hipSetDevice(dev));
    MI_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    g_device = dev;
    return MI_OK;
}

bool column_order_enabled()
{
    static const bool on = [] {
        const char* e = std::getenv("SIMALL_AMD_CSR_ORDER");
        return !(e && std::strcmp(e, "1") == 0);
    }();
    return on;
}

bool graphs_enabled()
{
    static const bool on = [] {
        const char* e = std::getenv("SIMALL_AMD_NO_GRAPH");
        return !(e && std::strcmp(e, "1") == 0);
    }();
    return on;
}

bool sell_delta_enabled()
{
    static const bool on = [] {
        const char* e = std::getenv("SIMALL_AMD_NO_SELL16");
        return !(e && std::strcmp(e, "1") == 0);
    }();
    return on;
}

bool sell_enabled()
{
    static const bool on = [] {
        const char* e = std::getenv("SIMALL_AMD_NO_SELL");
        return !(e && std::strcmp(e, "1") == 0);
    }();
    return on;
}

int ensure_device()
{
    std::call_once(g_init_once, [] { g_init_rc = init_device(0); });
    if (g_init_rc != MI_OK)
        return set_error(g_init_rc, "device initialisation failed earlier in this process");
    hipError_t e = hipSetDevice(g_device);
    if (e != hipSuccess)
        return set_error(MI_ERR_DEVICE, "hipSetDevice: %s", hipGetErrorString(e));
    return MI_OK;
}

hipStream_t default_stream() { return g_stream; }

} // namespace mi

using namespace mi;

extern "C" {

MI_API int mi_get_abi_version(void) { return MI_ABI_VERSION; }

MI_API const char* mi_last_error(void) { return g_err; }

MI_API int mi_init(int device)
{
    clear_error();
    int rc = MI_OK;
    bool ran = false;
    std::call_once(g_init_once, [&] {
        ran = true;
        g_init_rc = init_device(device);
        rc = g_init_rc;
    });
    if (!ran) {
        if (g_init_rc != MI_OK)
            return set_error(g_init_rc, "device initialisation failed earlier in this process");
        if (device != g_device)
            return set_error(MI_ERR_BAD_ARG, "already initialised on device %d", g_device);
    }
    return rc;
}

MI_API int mi_device_count(int* count)
{
    if (!count)
        return set_error(MI_ERR_BAD_ARG, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *count = (e == hipSuccess) ? c : 0;
    return MI_OK;
}

MI_API int mi_device_synchronize(void)
{
    MI_TRY(ensure_device());
    MI_HIP(hipStreamSynchronize(default_stream()));
    MI_HIP(hipDeviceSynchronize());
    return MI_OK;
}

MI_API int mi_device_alloc(size_t bytes, void** ptr)
{
    if (!ptr)
        return set_error(MI_ERR_BAD_ARG, "null output pointer");
    *ptr = nullptr;
    MI_TRY(ensure_device());
    if (bytes == 0)
        return MI_OK;
    MI_HIP(hipMalloc(ptr, bytes));
    return MI_OK;
}

MI_API int mi_device_free(void* ptr)
{
    if (!ptr)
        return MI_OK;
    MI_TRY(ensure_device());
    MI_HIP(hipFree(ptr));
    return MI_OK;
}

MI_API int mi_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    MI_TRY(ensure_device());
    if (bytes == 0)
        return MI_OK;
    if (!dst || !src)
        return set_error(MI_ERR_BAD_ARG, "null pointer");
    MI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, default_stream()));
    MI_HIP(hipStreamSynchronize(default_stream()));
    return MI_OK;
}

MI_API int mi_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    MI_TRY(ensure_device());
    if (bytes == 0)
        return MI_OK;
    if (!dst || !src)
        return set_error(MI_ERR_BAD_ARG, "null pointer");
    MI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, default_stream()));
    MI_HIP(hipStreamSynchronize(default_stream()));
    return MI_OK;
}

MI_API int mi_memcpy_d2d(void* dst, const void* src, size_t bytes)
{
    MI_TRY(ensure_device());
    if (bytes == 0)
        return MI_OK;
    if (!dst || !src)
        return set_error(MI_ERR_BAD_ARG, "null pointer");
    MI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, default_stream()));
    MI_HIP(hipStreamSynchronize(default_stream()));
    return MI_OK;
}

MI_API int mi_memset_device(void* dst, int value, size_t bytes)
{
    MI_TRY(ensure_device());
    if (bytes == 0)
        return MI_OK;
    if (!dst)
        return set_error(MI_ERR_BAD_ARG, "null pointer");
    MI_HIP(hipMemsetAsync(dst, value, bytes, default_stream()));
    MI_HIP(hipStreamSynchronize(default_stream()));
    return MI_OK;
}

} // extern "C"

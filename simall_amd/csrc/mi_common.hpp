// mi_common.hpp — shared host/device helpers for the gfx950 solve backend.
#pragma once

#include <vector>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>

#include "simall_amd.h"

namespace mi
{

// ---------------------------------------------------------------- errors
// Thread-local message behind mi_last_error(); set_error returns the code so
// call sites can `return set_error(...)`.
int set_error(int code, const char* fmt, ...);
void clear_error();

#define MI_HIP(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return ::mi::set_error(MI_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,                \
                                   hipGetErrorString(e_), __FILE__, __LINE__);                  \
    } while (0)

#define MI_TRY(call)                                                                            \
    do {                                                                                        \
        int rc_ = (call);                                                                       \
        if (rc_ != MI_OK)                                                                       \
            return rc_;                                                                         \
    } while (0)

// Initialises the HIP runtime once per process and checks for a device.
int ensure_device();
// Row blocks of stream matrices are stored in column order (the faster
// gathers) unless the environment sets SIMALL_AMD_CSR_ORDER=1, which keeps
// CSR order (for comparisons; results are identical either way).
bool column_order_enabled();
bool sell_enabled();
// Library stream used when a caller passes a null stream.
hipStream_t default_stream();

// ---------------------------------------------------------------- buffers
// Owning device allocation; movable, not copyable.
template <class T> struct DevBuf
{
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count)
    {
        if (count == n && p)
            return MI_OK;
        release();
        if (count == 0)
            return MI_OK;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return set_error(MI_ERR_DEVICE, "hipMalloc(%zu bytes) failed: %s", count * sizeof(T),
                             hipGetErrorString(e));
        }
        n = count;
        return MI_OK;
    }
    int upload(const T* h, size_t count, hipStream_t s)
    {
        MI_TRY(alloc(count));
        if (count)
            MI_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
        return MI_OK;
    }
};

// ---------------------------------------------------------------- launch
constexpr int kBlock = 256;  // 4 wavefronts of 64
constexpr int kWave = 64;

inline unsigned grid_for(int64_t work, int per_block)
{
    int64_t g = (work + per_block - 1) / per_block;
    return unsigned(g < 1 ? 1 : g);
}

// Number of blocks for grid-stride reductions: enough to fill 256 CUs with
// several blocks each, and a fixed upper bound so partial buffers are small.
constexpr int kMaxReduceBlocks = 2048;
inline unsigned reduce_grid(int64_t n, int per_thread = 4)
{
    int64_t g = (n + int64_t(kBlock) * per_thread - 1) / (int64_t(kBlock) * per_thread);
    if (g < 1)
        g = 1;
    if (g > kMaxReduceBlocks)
        g = kMaxReduceBlocks;
    return unsigned(g);
}

// ------------------------------------------------------- device reductions
#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// Sums NV values over a block of kBlock threads; the result is valid in
// thread 0.  `lds` must hold NV * (kBlock / kWave) doubles.
template <int NV> __device__ __forceinline__ void block_sum(double (&v)[NV], double* lds)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
#pragma unroll
    for (int j = 0; j < NV; ++j)
        v[j] = wave_sum(v[j]);
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < NV; ++j)
            lds[j * (kBlock / kWave) + wid] = v[j];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < kBlock / kWave; ++w)
                s += lds[j * (kBlock / kWave) + w];
            v[j] = s;
        }
    }
    __syncthreads();
}

// Grid-wide deterministic reductions are split in two launches: each block of
// the producing kernel stores its NV block sums (block_partials), and a
// one-block finalize kernel folds the G partials in a fixed order
// (fold_partials).  The kernel boundary publishes the partials; no block
// issues a device-scope fence, which on a multi-XCD device would write back
// the XCD's whole L2 every time.
template <int NV> __device__ __forceinline__ void block_partials(double (&v)[NV], double* partials)
{
    __shared__ double lds[NV * (kBlock / kWave)];
    block_sum<NV>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < NV; ++j)
            partials[size_t(j) * gridDim.x + blockIdx.x] = v[j];
}

// Totals over G partial blocks, valid in thread 0 of a one-block launch.
template <int NV>
__device__ __forceinline__ void fold_partials(const double* partials, int G, double (&tot)[NV])
{
    __shared__ double lds[NV * (kBlock / kWave)];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        double s = 0.0;
        for (int b = threadIdx.x; b < G; b += kBlock)
            s += partials[size_t(j) * G + b];
        tot[j] = s;
    }
    block_sum<NV>(tot, lds);
}

#endif

// Replays of a fixed launch sequence (a preconditioner apply): one
// executable HIP graph per (input, output, stream), captured on first use.
// The owner clears the cache whenever the buffers or constants the launches
// use change.  The sequences captured here contain kernel launches only.
struct GraphCache
{
    struct Entry
    {
        const void* in;
        const void* out;
        hipStream_t s;
        hipGraphExec_t exec;
    };
    std::vector<Entry> entries;
    GraphCache() = default;
    GraphCache(const GraphCache&) = delete;
    GraphCache& operator=(const GraphCache&) = delete;
    ~GraphCache() { clear(); }
    void clear()
    {
        for (Entry& e : entries)
            (void)hipGraphExecDestroy(e.exec);
        entries.clear();
    }
    template <class F> int run(const void* in, const void* out, hipStream_t s, F&& launches)
    {
        for (const Entry& e : entries)
            if (e.in == in && e.out == out && e.s == s) {
                MI_HIP(hipGraphLaunch(e.exec, s));
                return MI_OK;
            }
        hipGraph_t g = nullptr;
        MI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = launches();
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (rc != MI_OK || ec != hipSuccess || !g) {
            if (g)
                (void)hipGraphDestroy(g);
            return rc != MI_OK ? rc : set_error(MI_ERR_DEVICE, "graph capture failed");
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        MI_HIP(ei);
        entries.push_back(Entry{in, out, s, exec});
        MI_HIP(hipGraphLaunch(exec, s));
        return MI_OK;
    }
};

} // namespace mi

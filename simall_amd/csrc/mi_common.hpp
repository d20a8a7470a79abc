// mi_common.hpp — shared host/device helpers for the gfx950 solve backend.
#pragma once

#include <algorithm>
#include <vector>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <tuple>
#include <type_traits>
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
// Whether square SELL-64 forms get 16-bit column offsets (default on;
// SIMALL_AMD_NO_SELL16=1 keeps int32 columns, for comparisons).
bool sell_delta_enabled();
// Graph replays (V-cycles, chunks of batched iterations) are on unless the
// environment sets SIMALL_AMD_NO_GRAPH=1, which launches the same kernels
// directly: rocprofv3's kernel trace has ended the host process with SIGSEGV
// inside the runtime below hipGraph calls, so traced runs launch directly.
bool graphs_enabled();
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

// This thread's share of G partials: p[t] + p[t + kBlock] + ... (ascending,
// t = threadIdx.x).  The loads are issued kFoldBatch at a time before any is
// added, so a one-block fold waits for G / (kFoldBatch kBlock) memory round
// trips instead of G / kBlock; slots past G add +0.0, which leaves the sum
// (never -0.0: it starts at +0.0) bit for bit unchanged.
constexpr int kFoldBatch = 8;
__device__ __forceinline__ double strided_sum(const double* __restrict__ p, int G)
{
    double s = 0.0;
    for (int b0 = threadIdx.x; b0 < G; b0 += kFoldBatch * kBlock) {
        double v[kFoldBatch];
#pragma unroll
        for (int q = 0; q < kFoldBatch; ++q) {
            const int b = b0 + q * kBlock;
            v[q] = b < G ? p[b] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kFoldBatch; ++q)
            s += v[q];
    }
    return s;
}

// Totals over G partial blocks, valid in thread 0 of a one-block launch.
template <int NV>
__device__ __forceinline__ void fold_partials(const double* partials, int G, double (&tot)[NV])
{
    __shared__ double lds[NV * (kBlock / kWave)];
#pragma unroll
    for (int j = 0; j < NV; ++j)
        tot[j] = strided_sum(partials + size_t(j) * G, G);
    block_sum<NV>(tot, lds);
}

#endif

// Kernel launches that either run at once on `s` or, while `g` is set, are
// appended to that graph as kernel nodes, each depending on the previous one.
// Graphs are built node by node rather than by stream capture: nothing is
// enqueued while a graph is built, and a graph can only ever hold kernels.
struct Launcher
{
    hipStream_t s = nullptr;
    hipGraph_t g = nullptr;
    hipGraphNode_t last = nullptr;

    // k(args...) on a grid x block launch without dynamic LDS.
    template <class... P, class... A>
    int operator()(void (*k)(P...), dim3 grid, dim3 block, A&&... args)
    {
        static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
        std::tuple<std::decay_t<P>...> t(static_cast<std::decay_t<P>>(std::forward<A>(args))...);
        return go(reinterpret_cast<const void*>(k), grid, block, t,
                  std::index_sequence_for<P...>{});
    }

  private:
    template <class Tup, size_t... I>
    int go(const void* f, dim3 grid, dim3 block, Tup& t, std::index_sequence<I...>)
    {
        void* params[sizeof...(I) > 0 ? sizeof...(I) : 1] = {static_cast<void*>(&std::get<I>(t))...};
        if (!g) {
            MI_HIP(hipLaunchKernel(f, grid, block, params, 0, s));
            return MI_OK;
        }
        hipKernelNodeParams kp{};
        kp.func = const_cast<void*>(f);
        kp.gridDim = grid;
        kp.blockDim = block;
        kp.sharedMemBytes = 0;
        kp.kernelParams = params;
        kp.extra = nullptr;
        hipGraphNode_t node = nullptr;
        MI_HIP(hipGraphAddKernelNode(&node, g, last ? &last : nullptr, last ? 1 : 0, &kp));
        last = node;
        return MI_OK;
    }
};

// Builds an executable graph from `launches(Launcher&)` (kernel nodes only).
template <class F> int build_graph(hipStream_t s, F&& launches, hipGraphExec_t* out)
{
    *out = nullptr;
    hipGraph_t g = nullptr;
    MI_HIP(hipGraphCreate(&g, 0));
    Launcher L{s, g, nullptr};
    const int rc = launches(L);
    if (rc != MI_OK) {
        (void)hipGraphDestroy(g);
        return rc;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MI_HIP(ei);
    *out = exec;
    return MI_OK;
}

// Replays of a fixed launch sequence (a preconditioner apply): one
// executable HIP graph per (input, output, stream), built on first use.
// The owner clears the cache whenever the buffers or constants the launches
// use change.
struct GraphCache
{
    // A captured sequence is replayed only for the stream and the pointers it
    // was captured with (every buffer its launches read or write).
    static constexpr int kKey = 10;
    struct Entry
    {
        const void* key[kKey];
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
    // launches(Launcher&) issues the sequence through the launcher it is given
    template <class F> int run(const void* in, const void* out, hipStream_t s, F&& launches)
    {
        const void* key[kKey] = {in, out};
        return run_keyed(key, s, std::forward<F>(launches));
    }
    template <class F> int run_keyed(const void* const (&key)[kKey], hipStream_t s, F&& launches)
    {
        if (!graphs_enabled()) {
            Launcher L{s};
            return launches(L);
        }
        for (const Entry& e : entries)
            if (e.s == s && std::equal(key, key + kKey, e.key)) {
                MI_HIP(hipGraphLaunch(e.exec, s));
                return MI_OK;
            }
        hipGraphExec_t exec = nullptr;
        MI_TRY(build_graph(s, launches, &exec));
        Entry e{};
        std::copy(key, key + kKey, e.key);
        e.s = s;
        e.exec = exec;
        entries.push_back(e);
        MI_HIP(hipGraphLaunch(exec, s));
        return MI_OK;
    }
};

} // namespace mi

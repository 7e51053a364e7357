// Shared types of librooftop: complex arithmetic, error state, device memory accounting and the context.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "rooftop.h"

namespace rt {

// ---------------------------------------------------------------------------------------------------------------
// Complex numbers: plain POD pairs so buffers are interleaved (re, im) like MATLAB's interleaved complex API.
// ---------------------------------------------------------------------------------------------------------------
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T x, y;
};
using c128 = cplx<double>;
using c64 = cplx<float>;

template <typename T>
__host__ __device__ __forceinline__ cplx<T> mk(T x, T y) { return cplx<T>{x, y}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator*(T s, cplx<T> a) { return {s * a.x, s * a.y}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> conjg(cplx<T> a) { return {a.x, -a.y}; }
// conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
    return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
template <typename T>
__host__ __device__ __forceinline__ T abs2(cplx<T> a) { return a.x * a.x + a.y * a.y; }

template <typename To, typename From>
__host__ __device__ __forceinline__ cplx<To> cvt(cplx<From> a) { return {static_cast<To>(a.x), static_cast<To>(a.y)}; }

// ---------------------------------------------------------------------------------------------------------------
// Errors: every failure becomes an Error carrying a status, the reference's identifier and a message; the C ABI
// catches it, stores it thread-locally and returns the status.
// ---------------------------------------------------------------------------------------------------------------
struct Error : std::runtime_error {
    int status;
    std::string id;
    Error(int s, std::string i, const std::string &m) : std::runtime_error(m), status(s), id(std::move(i)) {}
};

[[noreturn]] void fail(int status, const char *id, const char *fmt, ...);
void set_last_error(int status, const std::string &id, const std::string &message);

#define RT_HIP(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            ::rt::fail(RT_EHIP, "rooftop:hip", "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                                           \
    } while (0)

#define RT_LAUNCH_CHECK() RT_HIP(hipGetLastError())

// ---------------------------------------------------------------------------------------------------------------
// Context: one device, one stream, a byte ceiling (the Budget of the reference) and the allocations it holds.
// ---------------------------------------------------------------------------------------------------------------
struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    // A second stream (default priority, created on first use) on which the matvec runs every other chunk of its
    // right-hand sides beside the main stream, and the untimed events that order the two.
    hipStream_t side = nullptr;
    hipEvent_t sync_events[4] = {};
    hipStream_t side_stream();
    hipEvent_t sync_event(int k);
    uint64_t ceiling = 0;
    uint64_t held = 0;
    uint64_t peak = 0;   // the most bytes held at once since the context opened or the last reset_peak()
    std::mutex lock;
    std::unordered_map<void *, uint64_t> sizes;

    // Kernel timing (rt_ctx_set_timing): while on, the timed kernels (RT_TIMED_* slots of rooftop.h) record a pair of
    // HIP events around their launch on the stream they run on, with their algorithmic bytes; drain_timing() reads
    // the events and accumulates per slot.
    // bytes: the algorithmic bytes of the launch for the share of a batch's systems it served; full: for every system;
    // step: the Krylov step the launch belongs to (-1: none), by which the loop sets `bytes` exactly once the
    // systems' iteration counts are known (each system runs the kernels of the steps up to its last).
    struct TimedRecord {
        hipEvent_t a, b;
        int slot;
        double bytes;
        double full;
        int step;
    };
    bool timing = false;
    std::vector<TimedRecord> pending;
    std::vector<hipEvent_t> event_pool;
    double timed_scale = 1.0;   // share of a batch's systems still running (set by a Krylov loop at each poll)
    int timed_step = -1;        // the Krylov step whose kernels are being enqueued (-1: none)
    int64_t timed_launches[RT_TIMED_COUNT] = {};
    double timed_ms[RT_TIMED_COUNT] = {};
    double timed_bytes[RT_TIMED_COUNT] = {};
    hipEvent_t take_event();
    void drain_timing();

    // Solver buffers kept between solves (grown on demand, freed with the context): slot -> (pointer, bytes).
    std::vector<std::pair<void *, uint64_t>> pool;
    void *host_pinned = nullptr;   // pinned staging of the solver's device scalars and loop states
    size_t host_pinned_bytes = 0;
    void *scratch(int slot, uint64_t bytes, const char *what);

    // Allocates device memory after checking the ceiling like Budget.require (Budget.m:82-106).
    void *alloc(uint64_t bytes, const char *what);
    void release(void *ptr);
    // Raises rooftopMoM:resources:insufficientMemory if bytes more do not fit under the ceiling.
    void require(uint64_t bytes, const char *what);
    void activate() const { RT_HIP(hipSetDevice(device)); }
    // Binds the context to a device with its own stream and a byte ceiling (0: the device's free memory).
    void open(int dev, uint64_t budget);
    // Frees everything the context still holds (device buffers, events, pinned staging) and its stream.
    void close();
    // Frees the solver buffers kept between solves.
    void free_pool();
};

// A device buffer owned through a context.
// Timing of one kernel launch: a pair of events that the launch itself stamps at the kernel's start and end
// (hipExtLaunchKernelGGL), so the measured duration is that of the kernel alone. Inactive while timing is off.
struct KernelTimer {
    Context *ctx;
    int slot;
    double bytes;
    hipEvent_t a = nullptr, b = nullptr;
    KernelTimer(Context *c, int sl, double by) : ctx(c), slot(sl), bytes(by) {
        if (ctx->timing) {
            a = ctx->take_event();
            b = ctx->take_event();
        }
    }
    void done() {
        if (!a) return;
        ctx->pending.push_back(Context::TimedRecord{a, b, slot, bytes * ctx->timed_scale, bytes, ctx->timed_step});
        a = b = nullptr;
        if (ctx->pending.size() > 8192) ctx->drain_timing();
    }
};

// Launches a kernel, stamping the events of t (when timing is on) at its start and end.
template <typename F, typename... Args>
void launch_timed(const KernelTimer *t, F kernel, dim3 grid, dim3 block, uint32_t lds, hipStream_t s, Args... args) {
    hipExtLaunchKernelGGL(kernel, grid, block, lds, s, t ? t->a : nullptr, t ? t->b : nullptr, 0, args...);
    RT_LAUNCH_CHECK();
}

template <typename T>
struct DeviceArray {
    Context *ctx = nullptr;
    T *ptr = nullptr;
    size_t count = 0;
    DeviceArray() = default;
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    DeviceArray(DeviceArray &&o) noexcept : ctx(o.ctx), ptr(o.ptr), count(o.count) { o.ptr = nullptr; o.count = 0; }
    DeviceArray &operator=(DeviceArray &&o) noexcept {
        if (this != &o) { reset(); ctx = o.ctx; ptr = o.ptr; count = o.count; o.ptr = nullptr; o.count = 0; }
        return *this;
    }
    ~DeviceArray() { reset(); }
    void allocate(Context *c, size_t n, const char *what) {
        reset();
        ctx = c;
        count = n;
        ptr = n ? static_cast<T *>(c->alloc(n * sizeof(T), what)) : nullptr;
    }
    void reset() {
        if (ptr && ctx) ctx->release(ptr);
        ptr = nullptr;
        count = 0;
    }
    uint64_t bytes() const { return count * sizeof(T); }
    void zero(hipStream_t s) { if (count) RT_HIP(hipMemsetAsync(ptr, 0, bytes(), s)); }
};

inline unsigned blocks_for(size_t n, unsigned threads, unsigned cap = 65535u * 16u) {
    size_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return static_cast<unsigned>(b);
}

}  // namespace rt

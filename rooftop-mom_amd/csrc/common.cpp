// Error state and device-memory accounting of librooftop.
#include <algorithm>
#include <cstdarg>
#include <cstdio>

#include "common.hpp"

namespace rt {

namespace {
thread_local int t_status = RT_OK;
thread_local std::string t_id;
thread_local std::string t_message;

std::string format_bytes(uint64_t b) {
    char buf[64];
    const char *units[] = {"B", "KiB", "MiB", "GiB", "TiB"};
    double v = static_cast<double>(b);
    int u = 0;
    while (v >= 1024.0 && u < 4) { v /= 1024.0; ++u; }
    std::snprintf(buf, sizeof buf, "%.3g %s", v, units[u]);
    return buf;
}
}  // namespace

void fail(int status, const char *id, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Error(status, id, buf);
}

void set_last_error(int status, const std::string &id, const std::string &message) {
    t_status = status;
    t_id = id;
    t_message = message;
}

const char *last_error_id() { return t_id.c_str(); }
const char *last_error_message() { return t_message.c_str(); }

void Context::require(uint64_t bytes, const char *what) {
    std::lock_guard<std::mutex> g(lock);
    if (held + bytes > ceiling) {
        uint64_t left = ceiling > held ? ceiling - held : 0;
        fail(RT_ENOMEM, "rooftopMoM:resources:insufficientMemory",
             "%s needs %s, but only %s of the memory ceiling of %s are left.", what, format_bytes(bytes).c_str(),
             format_bytes(left).c_str(), format_bytes(ceiling).c_str());
    }
}

void *Context::alloc(uint64_t bytes, const char *what) {
    require(bytes, what);
    // Stream-ordered allocation on the context's stream: freeing does not wait for the whole device (hipFree does),
    // so slots that build and free their requests' buffers do not stall each other, and the device's memory pool
    // hands freed memory to the next request without a trip to the driver.
    void *p = nullptr;
    hipError_t e = stream ? hipMallocAsync(&p, bytes, stream) : hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(RT_ENOMEM, "rooftopMoM:resources:insufficientMemory", "%s needs %s, but the device refused it (%s).",
             what, format_bytes(bytes).c_str(), hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> g(lock);
    held += bytes;
    if (held > peak) peak = held;
    sizes[p] = bytes;
    return p;
}

void Context::release(void *ptr) {
    if (!ptr) return;
    uint64_t bytes = 0;
    {
        std::lock_guard<std::mutex> g(lock);
        auto it = sizes.find(ptr);
        if (it == sizes.end()) return;
        bytes = it->second;
        sizes.erase(it);
        held -= bytes;
    }
    if (stream) (void)hipFreeAsync(ptr, stream);   // after the work queued on the stream (every use of ptr) is done
    else (void)hipFree(ptr);
}

void *Context::scratch(int slot, uint64_t bytes, const char *what) {
    if (static_cast<int>(pool.size()) <= slot) pool.resize(slot + 1, {nullptr, 0});
    auto &e = pool[slot];
    if (e.second < bytes) {
        if (e.first) {
            RT_HIP(hipStreamSynchronize(stream));
            release(e.first);
        }
        e.first = nullptr;
        e.second = 0;
        e.first = alloc(bytes, what);
        e.second = bytes;
    }
    return e.first;
}

void Context::open(int dev, uint64_t budget) {
    int count = 0;
    RT_HIP(hipGetDeviceCount(&count));
    if (dev < 0 || dev >= count) fail(RT_EINVAL, "rooftop:ctx", "no such device");
    device = dev;
    activate();
    size_t free_b = 0, total_b = 0;
    RT_HIP(hipMemGetInfo(&free_b, &total_b));
    ceiling = budget ? budget : static_cast<uint64_t>(free_b);
    RT_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // The device's pool keeps up to 8 GiB of what the contexts free for their next allocations instead of returning
    // it to the driver at every synchronization (the buffers of served requests come and go); larger amounts go back,
    // so that the free memory later contexts measure their ceiling against is not held in the pool.
    hipMemPool_t mp = nullptr;
    if (hipDeviceGetDefaultMemPool(&mp, dev) == hipSuccess && mp) {
        uint64_t keep = std::min<uint64_t>(ceiling, 8ull << 30);
        (void)hipMemPoolSetAttribute(mp, hipMemPoolAttrReleaseThreshold, &keep);
    }
}

hipStream_t Context::side_stream() {
    if (!side) RT_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    return side;
}

hipEvent_t Context::sync_event(int k) {
    if (!sync_events[k]) RT_HIP(hipEventCreateWithFlags(&sync_events[k], hipEventDisableTiming));
    return sync_events[k];
}

void Context::close() {
    (void)hipSetDevice(device);
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
    if (side) {
        (void)hipStreamSynchronize(side);
        (void)hipStreamDestroy(side);
        side = nullptr;
    }
    for (hipEvent_t &e : sync_events)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    for (auto &kv : sizes) (void)hipFree(kv.first);   // the stream is gone: synchronous frees
    sizes.clear();
    pool.clear();
    held = 0;
    peak = 0;
    for (auto &r : pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    pending.clear();
    for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
    event_pool.clear();
    if (host_pinned) (void)hipHostFree(host_pinned);
    host_pinned = nullptr;
    host_pinned_bytes = 0;
}

void Context::free_pool() {
    if (stream) RT_HIP(hipStreamSynchronize(stream));
    for (auto &e : pool)
        if (e.first) release(e.first);
    pool.clear();
}

hipEvent_t Context::take_event() {
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    // Default events: measured on MI355X, events created with hipEventDisableSystemFence read 2-3 us longer per
    // bracketed launch than these (12.8 against 10.5 us for a 10.4 us kernel by rocprofv3's kernel trace).
    hipEvent_t e;
    RT_HIP(hipEventCreateWithFlags(&e, hipEventDefault));
    return e;
}

void Context::drain_timing() {
    for (TimedRecord &r : pending) {
        RT_HIP(hipEventSynchronize(r.b));
        float t = 0;
        RT_HIP(hipEventElapsedTime(&t, r.a, r.b));
        timed_ms[r.slot] += t;
        timed_bytes[r.slot] += r.bytes;
        timed_launches[r.slot] += 1;
        event_pool.push_back(r.a);
        event_pool.push_back(r.b);
    }
    pending.clear();
}

}  // namespace rt

// Measurement kernel (no counterpart in the reference): a plain device copy of known size, which gives the HBM
// bandwidth a streaming kernel reaches on this device and calibrates the rocprofv3 FETCH_SIZE / WRITE_SIZE counters.
#include "rooftop.hpp"

namespace rt {

// 16-byte elements, one read and one write per element, one element per thread (the grid covers the buffer);
// the stores bypass the caches' retention like any streaming output.
typedef int v4i __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) copy_probe(const v4i *__restrict__ src, v4i *__restrict__ dst, int64_t n) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(src[i], dst + i);
}

// Average milliseconds of one copy of `bytes` (rounded down to 16) over `reps` launches after one warm-up.
double bandwidth_probe(Context *ctx, uint64_t bytes, int reps) {
    const int64_t n = static_cast<int64_t>(bytes / 16);
    if (n < 1 || reps < 1) fail(RT_EINVAL, "rooftop:ctx", "the probe needs at least 16 bytes and one repetition");
    DeviceArray<c128> a, b;
    a.allocate(ctx, n, "the bandwidth probe");
    b.allocate(ctx, n, "the bandwidth probe");
    hipStream_t s = ctx->stream;
    a.zero(s);
    const int64_t grid = (n + 255) / 256;
    if (grid > 0x7fffffff) fail(RT_EINVAL, "rooftop:ctx", "the probe buffer is too large");
    const unsigned blocks = static_cast<unsigned>(grid);
    auto launch = [&] {
        hipLaunchKernelGGL(copy_probe, dim3(blocks), dim3(256), 0, s, reinterpret_cast<const v4i *>(a.ptr),
                           reinterpret_cast<v4i *>(b.ptr), n);
        RT_LAUNCH_CHECK();
    };
    launch();
    hipEvent_t e0 = ctx->take_event(), e1 = ctx->take_event();
    RT_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < reps; ++k) launch();
    RT_HIP(hipEventRecord(e1, s));
    RT_HIP(hipEventSynchronize(e1));
    float ms = 0;
    RT_HIP(hipEventElapsedTime(&ms, e0, e1));
    ctx->event_pool.push_back(e0);
    ctx->event_pool.push_back(e1);
    return ms / reps;
}

}  // namespace rt

// The allocation-free application of a product (rt_product_apply_device_noalloc): the matvec passes run with their
// working memory in a file-scope device array instead of the product's own, grown-on-demand work buffer, so that a
// caller that cannot allocate (code under HIP graph capture, a tight serving loop, a harness that forbids hipMalloc)
// can apply an existing product. The array is kScratchBytes per device, held from the library's first use of a device
// until the process ends, and not counted in any context's budget. Every call on one device shares it: calls that
// overlap on one device race (their passes write the same working memory). Nothing in it needs to start at zero:
// pass A writes every element that passes B and C read.
#include <map>
#include <mutex>

#include "rooftop.hpp"

namespace rt {

constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB: the largest working memory of one application

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex scratch_lock;
std::map<int, void *> scratch_ptrs;        // the array's device address on each device
std::map<int, std::string> scratch_errors;  // the HIP error of a device whose address could not be read
}  // namespace

// Reads the array's address on `device` (which must be the current device) once, and keeps it. Called when a context
// opens on the device (rt_ctx_create), so that the entry point itself only looks the address up.
void noalloc_setup(int device) {
    std::lock_guard<std::mutex> g(scratch_lock);
    if (scratch_ptrs.count(device) || scratch_errors.count(device)) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        scratch_errors[device] = hipGetErrorString(e);
        return;
    }
    scratch_ptrs[device] = ptr;
}

static void *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> g(scratch_lock);
    const auto it = scratch_ptrs.find(device);
    if (it != scratch_ptrs.end()) return it->second;
    const auto bad = scratch_errors.find(device);
    if (bad != scratch_errors.end())
        fail(RT_EHIP, "rooftop:hip", "hipGetSymbolAddress of the allocation-free working memory failed on device %d: %s",
             device, bad->second.c_str());
    fail(RT_EINVAL, "rooftop:product:noAlloc", "no context was created on device %d", device);
}

void product_apply_noalloc(Product &op, const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s) {
    if (op.type != kOperator && op.type != kCirculant)
        fail(RT_EINVAL, "rooftop:product:noAlloc",
             "only the FFT products (operators and circulant preconditioners) use working memory; apply the "
             "near-field and Jacobi preconditioners with rt_product_apply_device");
    operator_apply_noalloc(op, in, out, nrhs, prec, s, noalloc_scratch(op.problem->ctx->device), kScratchBytes);
}

}  // namespace rt

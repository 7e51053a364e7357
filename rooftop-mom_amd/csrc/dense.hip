// Operators that are not a Green's matrix of a grid, so that the device Krylov loops take what the reference's
// solver takes (src/+rooftopMoM/+solver/+utils/convertToHandle.m:10-16: a matrix, a function handle, or an object with
// apply): a dense matrix held on the device and applied by a batched matrix-vector product, or a host callback that
// the loop calls with host copies of the vectors.
#include <algorithm>
#include <cstring>

#include "rooftop.hpp"

namespace rt {

// y_r = A x_r for the R = gridDim.y vectors (system-major), A n x n column-major in TC (complex128 or complex64,
// accumulated in the same precision, the vectors complex128). A thread per row: the column loads of a wave are
// contiguous, x_r(j) is the same for the whole block.
template <typename TC>
__global__ void __launch_bounds__(256) dense_apply_k(const cplx<TC> *__restrict__ A, const c128 *__restrict__ x,
                                                     c128 *__restrict__ y, int64_t n, const int *__restrict__ skip,
                                                     int skip_stride) {
    const int r = blockIdx.y;
    if (skip && skip[static_cast<int64_t>(r) * skip_stride]) return;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const c128 *xr = x + r * n;
    cplx<TC> acc = mk(TC(0), TC(0));
    for (int64_t j = 0; j < n; ++j) acc = acc + A[i + j * n] * cvt<TC>(xr[j]);
    y[r * n + i] = cvt<double>(acc);
}

template <typename TC>
__global__ void narrow_k(const c128 *__restrict__ in, cplx<TC> *__restrict__ out, int64_t count) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        out[i] = cvt<TC>(in[i]);
}

// A problem without a grid: the context and the unknowns of a product that is not a Green's matrix.
static std::shared_ptr<Problem> bare_problem(Context *ctx, int64_t n) {
    auto p = std::make_shared<Problem>();
    p->ctx = ctx;
    p->g.n = n;
    p->assembled = true;
    return p;
}

static ProdDims vector_dims(int64_t n) {
    ProdDims d{};
    d.n = n;
    return d;
}

std::unique_ptr<Product> create_dense(Context *ctx, int64_t n, const c128 *host_matrix, int precision,
                                      std::shared_ptr<Problem> &owner) {
    owner = bare_problem(ctx, n);
    auto op = std::make_unique<Product>();
    op->type = kDense;
    op->problem = owner.get();
    op->precision = precision;
    op->dims = vector_dims(n);
    hipStream_t s = ctx->stream;
    const size_t count = static_cast<size_t>(n) * n;
    op->dense64.allocate(ctx, count, "The dense operator");
    RT_HIP(hipMemcpyAsync(op->dense64.ptr, host_matrix, count * sizeof(c128), hipMemcpyHostToDevice, s));
    if (precision != RT_C128) {
        op->dense32.allocate(ctx, count, "The dense operator");
        hipLaunchKernelGGL(narrow_k<float>, dim3(blocks_for(count, 256, 8192)), dim3(256), 0, s, op->dense64.ptr,
                           op->dense32.ptr, static_cast<int64_t>(count));
        RT_LAUNCH_CHECK();
    }
    if (precision == RT_C64) op->dense64.reset();
    RT_HIP(hipStreamSynchronize(s));
    return op;
}

std::unique_ptr<Product> create_callback(Context *ctx, int64_t n, rt_apply_fn fn, void *user,
                                         std::shared_ptr<Problem> &owner) {
    owner = bare_problem(ctx, n);
    auto op = std::make_unique<Product>();
    op->type = kCallback;
    op->problem = owner.get();
    op->precision = RT_MIXED;   // the callback computes in its own arithmetic, whatever the loop's precision
    op->dims = vector_dims(n);
    op->callback = fn;
    op->callback_user = user;
    return op;
}

void dense_apply(Product &op, const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s, const int *skip,
                 int skip_stride) {
    const int64_t n = op.dims.n;
    if (op.type == kCallback) {
        // The host function sees host arrays: the vectors go to pinned staging and back, in order on the stream.
        const size_t bytes = static_cast<size_t>(n) * nrhs * sizeof(c128);
        if (op.staging_bytes < 2 * bytes) {
            if (op.staging) RT_HIP(hipHostFree(op.staging));
            op.staging = nullptr;
            op.staging_bytes = 0;
            RT_HIP(hipHostMalloc(&op.staging, 2 * bytes, hipHostMallocDefault));
            op.staging_bytes = 2 * bytes;
        }
        c128 *hin = static_cast<c128 *>(op.staging), *hout = hin + static_cast<size_t>(n) * nrhs;
        RT_HIP(hipMemcpyAsync(hin, in, bytes, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        std::memset(hout, 0, bytes);
        op.callback(op.callback_user, hin, hout, n, nrhs);
        RT_HIP(hipMemcpyAsync(out, hout, bytes, hipMemcpyHostToDevice, s));
        RT_HIP(hipStreamSynchronize(s));   // the staging is reused by the next call
        return;
    }
    if (prec == RT_MIXED) prec = RT_C128;
    if (!op.has(prec)) fail(RT_EINVAL, "rooftop:product", "the dense operator does not hold the requested precision");
    if (nrhs > 65535) fail(RT_EINVAL, "rooftop:product", "at most 65535 right-hand sides per application");
    const dim3 grid(static_cast<unsigned>((n + 255) / 256), static_cast<unsigned>(nrhs));
    if (prec == RT_C128)
        hipLaunchKernelGGL(dense_apply_k<double>, grid, dim3(256), 0, s, op.dense64.ptr, in, out, n, skip,
                           skip_stride);
    else
        hipLaunchKernelGGL(dense_apply_k<float>, grid, dim3(256), 0, s, op.dense32.ptr, in, out, n, skip,
                           skip_stride);
    RT_LAUNCH_CHECK();
}

}  // namespace rt

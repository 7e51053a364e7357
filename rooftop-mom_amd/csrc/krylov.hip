// The preconditioned Krylov loops on gfx950: COCG and restarted GMRES(m) with CGS2, FP64 vectors on the device,
// fused vector kernels with deterministic two-stage reductions (wave64 shuffles, fixed block order), and the FP64
// iterative-refinement loop around an inner solve whose products run in complex64.
//
// Reference: src/+rooftopMoM/+solver/IterativeMethod.m:27-70, ConjugateGradient.m:32-59,
// PreconditionedConjugateGradient.m:11-23, @GeneralizedMinimumResidual/iterate_.m:1-85, orthogonalize_.m:15-23,
// generateRotation_.m:11-18. Only scalars cross to the host: one residual norm per COCG iteration, the new Hessenberg
// column (j + 2 values) per GMRES step; the small Hessenberg rotation and triangular solve run on the host.
#include <chrono>
#include <cmath>
#include <complex>
#include <vector>

#include "krylov.hpp"

namespace rt {

constexpr int kRedBlocks = 1024;   // first-stage blocks of every reduction (fixed: the order of the sums is fixed)
constexpr int kRedThreads = 256;
constexpr int kColGroup = 8;       // basis columns per multi-dot pass

// ------------------------------------------------------------------------------------------------------------------
// Reductions.
// ------------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum of one double over the block (all threads call it); the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (blockDim.x + 63) / 64; ++w) r += sh[w];
    return r;
}

// partial[b] = sum over the block's rows of x^T y (unconjugated) or x^H y (conj = true).
template <bool CONJ>
__global__ void __launch_bounds__(kRedThreads) dot_partial(const c128 *__restrict__ x, const c128 *__restrict__ y,
                                                           int64_t n, c128 *__restrict__ partial) {
    __shared__ double sh[8];
    double re = 0, im = 0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const c128 a = x[i], b = y[i];
        const c128 p = CONJ ? cmulc(a, b) : a * b;
        re += p.x;
        im += p.y;
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = mk(re, im);
}

__global__ void __launch_bounds__(kRedThreads) nrm2_partial(const c128 *__restrict__ x, int64_t n,
                                                            c128 *__restrict__ partial) {
    __shared__ double sh[8];
    double s = 0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        s += abs2(x[i]);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = mk(s, 0.0);
}

// out[k] = sum_b partial[b * stride + k] (one block per k, fixed order). sqrt_re takes the square root of the real
// part (norms).
__global__ void __launch_bounds__(kRedThreads) reduce_partials(const c128 *__restrict__ partial, int blocks,
                                                               int stride, c128 *__restrict__ out, int sqrt_re) {
    __shared__ double sh[8];
    const int k = blockIdx.x;
    double re = 0, im = 0;
    for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
        const c128 v = partial[static_cast<int64_t>(b) * stride + k];
        re += v.x;
        im += v.y;
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    if (threadIdx.x == 0) out[k] = sqrt_re ? mk(sqrt(re), 0.0) : mk(re, im);
}

// ------------------------------------------------------------------------------------------------------------------
// Vector kernels.
// ------------------------------------------------------------------------------------------------------------------
#define GRID_STRIDE(i, n)                                                                \
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < (n); \
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)

__global__ void k_copy(const c128 *__restrict__ x, c128 *__restrict__ y, int64_t n) {
    GRID_STRIDE(i, n) y[i] = x[i];
}
__global__ void k_scale(const c128 *__restrict__ x, c128 a, c128 *__restrict__ y, int64_t n) {
    GRID_STRIDE(i, n) y[i] = a * x[i];
}
__global__ void k_sub(const c128 *__restrict__ b, const c128 *__restrict__ ax, c128 *__restrict__ r, int64_t n) {
    GRID_STRIDE(i, n) r[i] = b[i] - ax[i];
}
__global__ void k_axpy(c128 a, const c128 *__restrict__ x, c128 *__restrict__ y, int64_t n) {
    GRID_STRIDE(i, n) y[i] = y[i] + a * x[i];
}
// d = z + (rho_new / rho_old) d, the ratio read from device scalars.
__global__ void k_cocg_direction(const c128 *__restrict__ z, c128 *__restrict__ d, const c128 *__restrict__ sc,
                                 int i_new, int i_old, int64_t n) {
    const c128 rn = sc[i_new], ro = sc[i_old];
    const double den = abs2(ro);
    const c128 beta = (1.0 / den) * cmulc(ro, rn);
    GRID_STRIDE(i, n) d[i] = z[i] + beta * d[i];
}
// alpha = rho / (d^T A d); x += alpha d; r -= alpha Ad; per-block partial of |r|^2.
__global__ void __launch_bounds__(kRedThreads) k_cocg_update(c128 *__restrict__ x, c128 *__restrict__ r,
                                                             const c128 *__restrict__ d, const c128 *__restrict__ ad,
                                                             const c128 *__restrict__ sc, int i_rho, int i_dad,
                                                             int64_t n, c128 *__restrict__ partial) {
    __shared__ double sh[8];
    const c128 rho = sc[i_rho], dad = sc[i_dad];
    const double den = abs2(dad);
    const c128 alpha = (1.0 / den) * cmulc(dad, rho);
    double s = 0;
    GRID_STRIDE(i, n) {
        x[i] = x[i] + alpha * d[i];
        const c128 v = r[i] - alpha * ad[i];
        r[i] = v;
        s += abs2(v);
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = mk(s, 0.0);
}

// CGS2 multi-dot: partial[b * ldp + k] = sum over the block's rows of conj(V[:, k]) w, k in [k0, k0 + cnt).
__global__ void __launch_bounds__(kRedThreads) k_multidot(const c128 *__restrict__ V, int64_t ldv, int k0, int cnt,
                                                          const c128 *__restrict__ w, int64_t n,
                                                          c128 *__restrict__ partial, int ldp) {
    __shared__ double sh[8];
    double re[kColGroup], im[kColGroup];
    for (int k = 0; k < kColGroup; ++k) re[k] = im[k] = 0;
    GRID_STRIDE(i, n) {
        const c128 wi = w[i];
#pragma unroll
        for (int k = 0; k < kColGroup; ++k)
            if (k < cnt) {
                const c128 p = cmulc(V[(k0 + k) * ldv + i], wi);
                re[k] += p.x;
                im[k] += p.y;
            }
    }
    for (int k = 0; k < cnt; ++k) {
        const double a = block_sum(re[k], sh);
        const double b = block_sum(im[k], sh);
        if (threadIdx.x == 0) partial[static_cast<int64_t>(blockIdx.x) * ldp + k0 + k] = mk(a, b);
    }
}

// w -= V[:, 0:cnt] h; with norm_partial != nullptr also the per-block partial of |w|^2.
__global__ void __launch_bounds__(kRedThreads) k_gemv_sub(const c128 *__restrict__ V, int64_t ldv, int cnt,
                                                          const c128 *__restrict__ h, c128 *__restrict__ w, int64_t n,
                                                          c128 *__restrict__ norm_partial) {
    __shared__ double sh[8];
    __shared__ c128 hs[64];
    for (int k = threadIdx.x; k < cnt && k < 64; k += blockDim.x) hs[k] = h[k];
    __syncthreads();
    double s = 0;
    GRID_STRIDE(i, n) {
        c128 v = w[i];
        for (int k = 0; k < cnt; ++k) v = v - (k < 64 ? hs[k] : h[k]) * V[k * ldv + i];
        w[i] = v;
        s += abs2(v);
    }
    if (norm_partial) {
        s = block_sum(s, sh);
        if (threadIdx.x == 0) norm_partial[blockIdx.x] = mk(s, 0.0);
    }
}

// x += V[:, 0:cnt] y.
__global__ void k_gemv_add(const c128 *__restrict__ V, int64_t ldv, int cnt, const c128 *__restrict__ y,
                           c128 *__restrict__ x, int64_t n) {
    __shared__ c128 ys[64];
    for (int k = threadIdx.x; k < cnt && k < 64; k += blockDim.x) ys[k] = y[k];
    __syncthreads();
    GRID_STRIDE(i, n) {
        c128 v = x[i];
        for (int k = 0; k < cnt; ++k) v = v + (k < 64 ? ys[k] : y[k]) * V[k * ldv + i];
        x[i] = v;
    }
}

__global__ void k_add_vec(c128 *__restrict__ a, const c128 *__restrict__ b, int cnt) {
    const int k = threadIdx.x;
    if (k < cnt) a[k] = a[k] + b[k];
}

// ------------------------------------------------------------------------------------------------------------------
// Host helpers.
// ------------------------------------------------------------------------------------------------------------------
static unsigned vec_blocks(int64_t n) { return blocks_for(n, 256, 4096); }

Workspace::Workspace(Context *c, int64_t n_) : ctx(c), s(c->stream), n(n_) {
    partial.allocate(ctx, static_cast<size_t>(kRedBlocks) * 64, "the solver's reduction buffers");
    scal.allocate(ctx, 256, "the solver's scalars");
    RT_HIP(hipHostMalloc(reinterpret_cast<void **>(&host), 256 * sizeof(c128), hipHostMallocDefault));
}

Workspace::~Workspace() {
    if (host) (void)hipHostFree(host);
}

int Workspace::red_blocks() const {
    int64_t b = (n + 2047) / 2048;
    if (b < 1) b = 1;
    if (b > kRedBlocks) b = kRedBlocks;
    return static_cast<int>(b);
}

void Workspace::dot(const c128 *x, const c128 *y, bool conj, int slot) {
    const int b = red_blocks();
    if (conj)
        hipLaunchKernelGGL(dot_partial<true>, dim3(b), dim3(kRedThreads), 0, s, x, y, n, partial.ptr);
    else
        hipLaunchKernelGGL(dot_partial<false>, dim3(b), dim3(kRedThreads), 0, s, x, y, n, partial.ptr);
    RT_LAUNCH_CHECK();
    hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(kRedThreads), 0, s, partial.ptr, b, 1, scal.ptr + slot, 0);
    RT_LAUNCH_CHECK();
}

void Workspace::nrm2(const c128 *x, int slot) {
    const int b = red_blocks();
    hipLaunchKernelGGL(nrm2_partial, dim3(b), dim3(kRedThreads), 0, s, x, n, partial.ptr);
    RT_LAUNCH_CHECK();
    hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(kRedThreads), 0, s, partial.ptr, b, 1, scal.ptr + slot, 1);
    RT_LAUNCH_CHECK();
}

void Workspace::fetch(int first, int count) {
    RT_HIP(hipMemcpyAsync(host + first, scal.ptr + first, count * sizeof(c128), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
}

void Workspace::copy(const c128 *x, c128 *y) {
    hipLaunchKernelGGL(k_copy, dim3(vec_blocks(n)), dim3(256), 0, s, x, y, n);
    RT_LAUNCH_CHECK();
}
void Workspace::scale(const c128 *x, c128 a, c128 *y) {
    hipLaunchKernelGGL(k_scale, dim3(vec_blocks(n)), dim3(256), 0, s, x, a, y, n);
    RT_LAUNCH_CHECK();
}
void Workspace::sub(const c128 *b, const c128 *ax, c128 *r) {
    hipLaunchKernelGGL(k_sub, dim3(vec_blocks(n)), dim3(256), 0, s, b, ax, r, n);
    RT_LAUNCH_CHECK();
}
void Workspace::axpy(c128 a, const c128 *x, c128 *y) {
    hipLaunchKernelGGL(k_axpy, dim3(vec_blocks(n)), dim3(256), 0, s, a, x, y, n);
    RT_LAUNCH_CHECK();
}
void Workspace::zero(c128 *x) { RT_HIP(hipMemsetAsync(x, 0, n * sizeof(c128), s)); }

// ------------------------------------------------------------------------------------------------------------------
// COCG (ConjugateGradient.m:32-59). Scalar slots: 0/1 rho (alternating), 2 d^T A d, 3 |r|.
// ------------------------------------------------------------------------------------------------------------------
static InnerResult cocg(Workspace &ws, const LinOp &A, const LinOp *P, const c128 *b, c128 *x, bool has_x0,
                        double tol, int max_iter, double bnorm) {
    InnerResult res;
    Context *ctx = ws.ctx;
    const int64_t n = ws.n;
    DeviceArray<c128> r, z, d, ad;
    r.allocate(ctx, n, "the solver's vectors");
    d.allocate(ctx, n, "the solver's vectors");
    ad.allocate(ctx, n, "the solver's vectors");
    if (P) z.allocate(ctx, n, "the solver's vectors");
    if (has_x0) {
        A(x, ad.ptr);
        ws.sub(b, ad.ptr, r.ptr);
    } else {
        ws.zero(x);
        ws.copy(b, r.ptr);
    }
    ws.nrm2(r.ptr, 3);
    ws.fetch(3, 1);
    res.norms.push_back(ws.host[3].x / bnorm);
    bool conv = res.norms.back() < tol;
    int it = 0;
    const int rb = ws.red_blocks();
    while (!conv && it < max_iter) {
        ++it;
        const c128 *zz = r.ptr;
        if (P) {
            (*P)(r.ptr, z.ptr);
            zz = z.ptr;
        }
        const int i_new = it & 1, i_old = 1 - i_new;
        ws.dot(r.ptr, zz, false, i_new);
        if (it == 1)
            ws.copy(zz, d.ptr);
        else {
            hipLaunchKernelGGL(k_cocg_direction, dim3(vec_blocks(n)), dim3(256), 0, ws.s, zz, d.ptr, ws.scal.ptr,
                               i_new, i_old, n);
            RT_LAUNCH_CHECK();
        }
        A(d.ptr, ad.ptr);
        ws.dot(d.ptr, ad.ptr, false, 2);
        hipLaunchKernelGGL(k_cocg_update, dim3(rb), dim3(kRedThreads), 0, ws.s, x, r.ptr, d.ptr, ad.ptr, ws.scal.ptr,
                           i_new, 2, n, ws.partial.ptr);
        RT_LAUNCH_CHECK();
        hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(kRedThreads), 0, ws.s, ws.partial.ptr, rb, 1,
                           ws.scal.ptr + 3, 1);
        RT_LAUNCH_CHECK();
        ws.fetch(3, 1);
        const double nr = ws.host[3].x / bnorm;
        res.norms.push_back(nr);
        conv = nr < tol;
        if (!std::isfinite(nr)) break;
    }
    res.converged = conv;
    res.iterations = it;
    return res;
}

// ------------------------------------------------------------------------------------------------------------------
// GMRES(m), left preconditioned, CGS2 (iterate_.m:1-85). Scalar slots: [0, m+1) Hessenberg column, 64 first-pass
// coefficients, 128 |w|.
// ------------------------------------------------------------------------------------------------------------------
static void multidot(Workspace &ws, const c128 *V, int64_t ldv, int cnt, const c128 *w, int slot) {
    const int rb = ws.red_blocks();
    for (int k0 = 0; k0 < cnt; k0 += kColGroup) {
        const int c = cnt - k0 < kColGroup ? cnt - k0 : kColGroup;
        hipLaunchKernelGGL(k_multidot, dim3(rb), dim3(kRedThreads), 0, ws.s, V, ldv, k0, c, w, ws.n, ws.partial.ptr,
                           64);
        RT_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(reduce_partials, dim3(cnt), dim3(kRedThreads), 0, ws.s, ws.partial.ptr, rb, 64,
                       ws.scal.ptr + slot, 0);
    RT_LAUNCH_CHECK();
}

static InnerResult gmres(Workspace &ws, const LinOp &A, const LinOp *P, const c128 *b, c128 *x, bool has_x0,
                         double tol, int max_iter, int restart) {
    InnerResult res;
    Context *ctx = ws.ctx;
    const int64_t n = ws.n;
    int64_t m64 = restart > 0 ? restart : n;
    if (m64 > n) m64 = n;
    if (m64 > max_iter) m64 = max_iter;
    const int m = static_cast<int>(m64);
    if (m > 62) fail(RT_EINVAL, "rooftopMoM:solver:invalidRestart", "restarts above 62 are not supported on the GPU");
    DeviceArray<c128> V, w, t;
    V.allocate(ctx, static_cast<size_t>(n) * (m + 1), "The GMRES basis");
    w.allocate(ctx, n, "the solver's vectors");
    t.allocate(ctx, n, "the solver's vectors");
    auto precondition = [&](const c128 *in, c128 *out) {
        if (P) (*P)(in, out);
        else if (in != out) ws.copy(in, out);
    };
    // reference = norm(P b)
    precondition(b, w.ptr);
    ws.nrm2(w.ptr, 128);
    ws.fetch(128, 1);
    const double reference = ws.host[128].x;
    if (has_x0) {
        A(x, t.ptr);
        ws.sub(b, t.ptr, t.ptr);
        precondition(t.ptr, w.ptr);
        ws.nrm2(w.ptr, 128);
        ws.fetch(128, 1);
    } else {
        ws.zero(x);
    }
    double rnorm = ws.host[128].x;
    res.norms.push_back(rnorm / reference);
    bool conv = res.norms.back() < tol;
    int it = 0;
    using cd = std::complex<double>;
    std::vector<cd> H(static_cast<size_t>(m + 1) * m), cs(m), sn(m), g(m + 1);
    auto Hat = [&](int i, int j) -> cd & { return H[i + static_cast<size_t>(m + 1) * j]; };
    DeviceArray<c128> ydev;
    ydev.allocate(ctx, 64, "the solver's scalars");
    while (!conv && it < max_iter) {
        ws.scale(w.ptr, mk(1.0 / rnorm, 0.0), V.ptr);   // v_0 = r / |r|
        std::fill(g.begin(), g.end(), cd(0));
        g[0] = rnorm;
        int step = 0;
        while (!conv && step < m && it < max_iter) {
            ++step;
            ++it;
            const c128 *vj = V.ptr + static_cast<int64_t>(step - 1) * n;
            A(vj, t.ptr);
            precondition(t.ptr, w.ptr);
            // CGS2 against columns 0..step-1.
            multidot(ws, V.ptr, n, step, w.ptr, 64);
            hipLaunchKernelGGL(k_gemv_sub, dim3(vec_blocks(n)), dim3(256), 0, ws.s, V.ptr, n, step, ws.scal.ptr + 64,
                               w.ptr, n, nullptr);
            RT_LAUNCH_CHECK();
            multidot(ws, V.ptr, n, step, w.ptr, 0);
            const int rb = ws.red_blocks();
            hipLaunchKernelGGL(k_gemv_sub, dim3(rb), dim3(kRedThreads), 0, ws.s, V.ptr, n, step, ws.scal.ptr, w.ptr, n,
                               ws.partial.ptr);
            RT_LAUNCH_CHECK();
            hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(kRedThreads), 0, ws.s, ws.partial.ptr, rb, 1,
                               ws.scal.ptr + 128, 1);
            RT_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_add_vec, dim3(1), dim3(64), 0, ws.s, ws.scal.ptr, ws.scal.ptr + 64, step);
            RT_LAUNCH_CHECK();
            RT_HIP(hipMemcpyAsync(ws.host, ws.scal.ptr, step * sizeof(c128), hipMemcpyDeviceToHost, ws.s));
            ws.fetch(128, 1);
            const int j = step - 1;
            for (int k = 0; k < step; ++k) Hat(k, j) = cd(ws.host[k].x, ws.host[k].y);
            Hat(step, j) = ws.host[128].x;
            for (int k = 0; k < j; ++k) {
                const cd h0 = Hat(k, j), h1 = Hat(k + 1, j);
                Hat(k, j) = std::conj(cs[k]) * h0 + std::conj(sn[k]) * h1;
                Hat(k + 1, j) = -sn[k] * h0 + cs[k] * h1;
            }
            const cd a = Hat(j, j), bb = Hat(step, j);
            const double radius = std::sqrt(std::norm(a) + std::norm(bb));
            if (radius > 0) { cs[j] = a / radius; sn[j] = bb / radius; }
            else { cs[j] = 1.0; sn[j] = 0.0; }
            const cd sub = Hat(step, j);
            Hat(j, j) = std::conj(cs[j]) * Hat(j, j) + std::conj(sn[j]) * sub;
            Hat(step, j) = 0;
            g[step] = -sn[j] * g[j];
            g[j] = std::conj(cs[j]) * g[j];
            res.norms.push_back(std::abs(g[step]) / reference);
            conv = res.norms.back() < tol;
            if (!conv && sub != cd(0)) ws.scale(w.ptr, mk(1.0 / sub.real(), 0.0), V.ptr + static_cast<int64_t>(step) * n);
        }
        // y = H(1:step, 1:step) \ g(1:step), upper triangular.
        std::vector<cd> y(step);
        for (int k = step - 1; k >= 0; --k) {
            cd sacc = g[k];
            for (int l = k + 1; l < step; ++l) sacc -= Hat(k, l) * y[l];
            y[k] = sacc / Hat(k, k);
        }
        std::vector<c128> yh(step);
        for (int k = 0; k < step; ++k) yh[k] = mk(y[k].real(), y[k].imag());
        RT_HIP(hipMemcpyAsync(ydev.ptr, yh.data(), step * sizeof(c128), hipMemcpyHostToDevice, ws.s));
        hipLaunchKernelGGL(k_gemv_add, dim3(vec_blocks(n)), dim3(256), 0, ws.s, V.ptr, n, step, ydev.ptr, x, n);
        RT_LAUNCH_CHECK();
        RT_HIP(hipStreamSynchronize(ws.s));
        if (!conv && it < max_iter) {
            A(x, t.ptr);
            ws.sub(b, t.ptr, t.ptr);
            precondition(t.ptr, w.ptr);
            ws.nrm2(w.ptr, 128);
            ws.fetch(128, 1);
            rnorm = ws.host[128].x;
        }
    }
    res.converged = conv;
    res.iterations = it;
    return res;
}

// ------------------------------------------------------------------------------------------------------------------
// IterativeMethod.solve (IterativeMethod.m:27-70) with the optional FP64 refinement loop on the true residual.
// ------------------------------------------------------------------------------------------------------------------
SolveResult solve_one(Context *ctx, Product &op, Product *pc, const SolveOptions &o, const c128 *b, const c128 *x0,
                      c128 *x) {
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t n = op.dims.n;
    Workspace ws(ctx, n);
    hipStream_t s = ctx->stream;
    SolveResult out;
    ws.nrm2(b, 200);
    ws.fetch(200, 1);
    const double bnorm = ws.host[200].x;
    DeviceArray<c128> t;
    t.allocate(ctx, n, "the solver's vectors");
    // The outer residual runs in FP64 whenever the operator holds FP64 spectra.
    const int outer_prec = op.has(RT_C128) ? RT_C128 : RT_C64;
    auto true_residual = [&](const c128 *xx) {
        op.apply(xx, t.ptr, 1, outer_prec, s);
        ws.sub(b, t.ptr, t.ptr);
        ws.nrm2(t.ptr, 201);
        ws.fetch(201, 1);
        return ws.host[201].x / bnorm;
    };
    if (!(bnorm > 0)) {
        ws.zero(x);
        RT_HIP(hipStreamSynchronize(s));
        out.converged = true;
        out.norms.push_back(0.0);
        out.residual = 0;
    } else {
        const int inner_prec = o.precision;
        LinOp A = [&](const c128 *in, c128 *res) { op.apply(in, res, 1, inner_prec, s); };
        LinOp Pop = [&](const c128 *in, c128 *res) { pc->apply(in, res, 1, inner_prec, s); };
        const LinOp *P = pc ? &Pop : nullptr;
        auto run = [&](const c128 *rhs, c128 *sol, bool has_x0, double tol, int max_it) {
            if (o.method == RT_METHOD_CG) {
                ws.nrm2(rhs, 202);
                ws.fetch(202, 1);
                return cocg(ws, A, P, rhs, sol, has_x0, tol, max_it, ws.host[202].x);
            }
            return gmres(ws, A, P, rhs, sol, has_x0, tol, max_it, o.restart);
        };
        if (!o.refine) {
            InnerResult r = run(b, x, x0 != nullptr, o.tolerance, o.max_iterations);
            out.converged = r.converged;
            out.iterations = r.iterations;
            out.norms = r.norms;
            out.residual = true_residual(x);
        } else {
            // x_{k+1} = x_k + d_k, A d_k = b - A x_k solved by the inner method in its precision.
            if (x0) ws.copy(x0, x);
            else ws.zero(x);
            DeviceArray<c128> r, d;
            r.allocate(ctx, n, "the solver's vectors");
            d.allocate(ctx, n, "the solver's vectors");
            double rel = true_residual(x);
            out.norms.push_back(rel);
            const double inner_tol = o.inner_tolerance > 0 ? o.inner_tolerance
                                                           : (o.precision == RT_C64 ? 1e-5 : o.tolerance);
            int total = 0;
            while (!(rel < o.tolerance) && out.outer < o.max_outer && total < o.max_iterations) {
                ++out.outer;
                ws.copy(t.ptr, r.ptr);   // t holds b - A x from true_residual
                // Ask the inner solve for what is left to reach the tolerance, but no less than its floor.
                double want = 0.5 * o.tolerance / rel;
                if (want < inner_tol) want = inner_tol;
                if (want > 0.5) want = 0.5;
                InnerResult ir = run(r.ptr, d.ptr, false, want, o.max_iterations - total);
                total += ir.iterations;
                for (size_t k = 1; k < ir.norms.size(); ++k) out.norms.push_back(ir.norms[k] * rel);
                ws.axpy(mk(1.0, 0.0), d.ptr, x);
                const double next = true_residual(x);
                if (!std::isfinite(next)) { rel = next; break; }
                rel = next;
            }
            out.iterations = total;
            out.residual = rel;
            out.converged = rel < o.tolerance;
        }
    }
    RT_HIP(hipStreamSynchronize(s));
    out.time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

}  // namespace rt

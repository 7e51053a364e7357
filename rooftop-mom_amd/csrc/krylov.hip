// The preconditioned Krylov loops on gfx950: COCG and restarted GMRES(m) with CGS2, FP64 vectors on the device,
// and the FP64 iterative-refinement loop around an inner solve whose products run in complex64.
//
// Reference: src/+rooftopMoM/+solver/IterativeMethod.m:27-70, ConjugateGradient.m:32-59,
// PreconditionedConjugateGradient.m:11-23, @GeneralizedMinimumResidual/iterate_.m:1-85, orthogonalize_.m:15-23,
// generateRotation_.m:11-18.
//
// Everything a step needs stays on the device: the scalars (rho, d^T A d, the Hessenberg column), the Givens rotations,
// the residual history and the stopping test, written by one-block "finalize" kernels. The host enqueues a window of
// steps and polls the loop state once per window; when a finalize kernel stops the loop (convergence, breakdown, the
// iteration limit) it raises a flag that makes every later kernel of the window -- products, preconditioner, vector
// kernels -- return at once, so the result is exactly that of stopping at that step. Reductions are deterministic:
// fixed row blocks, wave64 shuffles, and a fixed-order second stage.
//
// CGS2 (orthogonalize_.m:15-23) runs as three row-partitioned passes over the live columns only (the reference takes
// the whole basis with zeroed coefficients): h1 = V^H w; w -= V h1 fused with h2 = V^H w; w -= V h2 fused with |w|^2.
#include <chrono>
#include <cmath>
#include <complex>
#include <vector>

#include "krylov.hpp"

namespace rt {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 2048;
constexpr int kLdp = 64;        // partial columns per block
constexpr int kPoll = 8;        // steps enqueued between two polls of the loop state

// Pool slots of the context's solver buffers.
enum Slot {
    kSlotPartial, kSlotScal, kSlotState, kSlotHistory, kSlotV, kSlotW, kSlotT, kSlotR, kSlotZ, kSlotD, kSlotAd,
    kSlotH, kSlotG, kSlotOuterT, kSlotOuterR, kSlotOuterD
};

// ------------------------------------------------------------------------------------------------------------------
// Reductions.
// ------------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum over the block (all threads call it); valid in thread 0.
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

#define SKIP_IF_DONE() \
    if (skip && *skip) return

// partial[b * kLdp] = sum over block b's rows of x^T y (or x^H y).
template <bool CONJ>
__global__ void __launch_bounds__(kThreads) k_dot(const c128 *__restrict__ x, const c128 *__restrict__ y, int64_t n,
                                                  int64_t chunk, c128 *__restrict__ partial,
                                                  const int *__restrict__ skip) {
    SKIP_IF_DONE();
    __shared__ double sh[8];
    const int64_t base = blockIdx.x * chunk, end = min(n, base + chunk);
    double re = 0, im = 0;
    for (int64_t i = base + threadIdx.x; i < end; i += kThreads) {
        const c128 p = CONJ ? cmulc(x[i], y[i]) : x[i] * y[i];
        re += p.x;
        im += p.y;
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    if (threadIdx.x == 0) partial[blockIdx.x * kLdp] = mk(re, im);
}

__global__ void __launch_bounds__(kThreads) k_nrm2(const c128 *__restrict__ x, int64_t n, int64_t chunk,
                                                   c128 *__restrict__ partial, const int *__restrict__ skip) {
    SKIP_IF_DONE();
    __shared__ double sh[8];
    const int64_t base = blockIdx.x * chunk, end = min(n, base + chunk);
    double s = 0;
    for (int64_t i = base + threadIdx.x; i < end; i += kThreads) s += abs2(x[i]);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x * kLdp] = mk(s, 0.0);
}

// out[k] = sum_b partial[b * kLdp + k], one block per column k; sqrt_re: the square root of the real part (a norm).
__global__ void __launch_bounds__(kThreads) k_reduce(const c128 *__restrict__ partial, int blocks,
                                                     c128 *__restrict__ out, int sqrt_re,
                                                     const int *__restrict__ skip) {
    SKIP_IF_DONE();
    __shared__ double sh[8];
    const int k = blockIdx.x;
    double re = 0, im = 0;
    for (int b = threadIdx.x; b < blocks; b += kThreads) {
        const c128 v = partial[b * kLdp + k];
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
#define GRID_STRIDE(i, n)                                                                  \
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < (n); \
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)

__global__ void k_copy(const c128 *__restrict__ x, c128 *__restrict__ y, int64_t n, const int *__restrict__ skip) {
    SKIP_IF_DONE();
    GRID_STRIDE(i, n) y[i] = x[i];
}
__global__ void k_scale(const c128 *__restrict__ x, c128 a, c128 *__restrict__ y, int64_t n,
                        const int *__restrict__ skip) {
    SKIP_IF_DONE();
    GRID_STRIDE(i, n) y[i] = a * x[i];
}
// y = a * x with a read from the device.
__global__ void k_scale_dev(const c128 *__restrict__ x, const c128 *__restrict__ a, c128 *__restrict__ y, int64_t n,
                            const int *__restrict__ skip) {
    SKIP_IF_DONE();
    const c128 s = *a;
    GRID_STRIDE(i, n) y[i] = s * x[i];
}
__global__ void k_sub(const c128 *__restrict__ b, const c128 *__restrict__ ax, c128 *__restrict__ r, int64_t n,
                      const int *__restrict__ skip) {
    SKIP_IF_DONE();
    GRID_STRIDE(i, n) r[i] = b[i] - ax[i];
}
__global__ void k_axpy(c128 a, const c128 *__restrict__ x, c128 *__restrict__ y, int64_t n,
                       const int *__restrict__ skip) {
    SKIP_IF_DONE();
    GRID_STRIDE(i, n) y[i] = y[i] + a * x[i];
}
__global__ void k_zero_state(LoopState *st) {
    if (threadIdx.x == 0) *st = LoopState{0, 0, 0, 0, {0, 0, 0, 0}};
}

// ------------------------------------------------------------------------------------------------------------------
// COCG kernels.
// ------------------------------------------------------------------------------------------------------------------
// d = z + (rho_new / rho_old) d.
__global__ void k_cocg_direction(const c128 *__restrict__ z, c128 *__restrict__ d, const c128 *__restrict__ sc,
                                 int i_new, int i_old, int64_t n, const int *__restrict__ skip) {
    SKIP_IF_DONE();
    const c128 rn = sc[i_new], ro = sc[i_old];
    const c128 beta = (1.0 / abs2(ro)) * cmulc(ro, rn);
    GRID_STRIDE(i, n) d[i] = z[i] + beta * d[i];
}

// alpha = rho / (d^T A d); x += alpha d; r -= alpha Ad; per-block partial of |r|^2.
__global__ void __launch_bounds__(kThreads) k_cocg_update(c128 *__restrict__ x, c128 *__restrict__ r,
                                                          const c128 *__restrict__ d, const c128 *__restrict__ ad,
                                                          const c128 *__restrict__ sc, int i_rho, int i_dad,
                                                          int64_t n, int64_t chunk, c128 *__restrict__ partial,
                                                          const int *__restrict__ skip) {
    SKIP_IF_DONE();
    __shared__ double sh[8];
    const c128 rho = sc[i_rho], dad = sc[i_dad];
    const c128 alpha = (1.0 / abs2(dad)) * cmulc(dad, rho);
    const int64_t base = blockIdx.x * chunk, end = min(n, base + chunk);
    double s = 0;
    for (int64_t i = base + threadIdx.x; i < end; i += kThreads) {
        x[i] = x[i] + alpha * d[i];
        const c128 v = r[i] - alpha * ad[i];
        r[i] = v;
        s += abs2(v);
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x * kLdp] = mk(s, 0.0);
}

// The residual norm of iteration `it`, the stopping test (ConjugateGradient.m:126-136).
__global__ void __launch_bounds__(kThreads) k_cocg_finalize(const c128 *__restrict__ partial, int blocks,
                                                            double bnorm, double tol, int it, int max_iter,
                                                            LoopState *st, double *history,
                                                            const int *__restrict__ skip) {
    SKIP_IF_DONE();
    __shared__ double sh[8];
    double s = 0;
    for (int b = threadIdx.x; b < blocks; b += kThreads) s += partial[b * kLdp].x;
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
        const double res = sqrt(s) / bnorm;
        history[it] = res;
        st->iterations = it;
        if (res < tol) {
            st->converged = 1;
            st->done = 1;
        } else if (!isfinite(res) || it >= max_iter) {
            st->done = 1;
        }
    }
}

// ------------------------------------------------------------------------------------------------------------------
// GMRES kernels.
// ------------------------------------------------------------------------------------------------------------------
// partial[b * kLdp + k] = sum over block b's rows of conj(V[:, k]) w, k < cnt <= 64. Wave q takes the columns
// q, q + 4, ...; its lanes run over the block's rows (coalesced column reads), w is read once per wave.
__global__ void __launch_bounds__(kThreads) k_cgs_dot(const c128 *__restrict__ V, int64_t ldv, int cnt,
                                                      const c128 *__restrict__ w, int64_t n, int64_t chunk,
                                                      c128 *__restrict__ partial, const int *__restrict__ skip) {
    SKIP_IF_DONE();
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int64_t base = blockIdx.x * chunk, end = min(n, base + chunk);
    double re[16], im[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) re[c] = im[c] = 0;
    for (int64_t r = base + lane; r < end; r += 64) {
        const c128 wr = w[r];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int k = q + 4 * c;
            if (k < cnt) {
                const c128 v = V[k * ldv + r];
                re[c] += v.x * wr.x + v.y * wr.y;
                im[c] += v.x * wr.y - v.y * wr.x;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int k = q + 4 * c;
        if (k < cnt) {
            const double a = wave_sum(re[c]), b = wave_sum(im[c]);
            if (lane == 0) partial[blockIdx.x * kLdp + k] = mk(a, b);
        }
    }
}

// w -= V[:, 0:cnt] h, then partial[b * kLdp + k] = conj(V[:, k]) w over the block's rows (w staged in LDS per tile of
// 256 rows).
__global__ void __launch_bounds__(kThreads) k_cgs_update_dot(const c128 *__restrict__ V, int64_t ldv, int cnt,
                                                             const c128 *__restrict__ h, c128 *__restrict__ w,
                                                             int64_t n, int64_t chunk, c128 *__restrict__ partial,
                                                             const int *__restrict__ skip) {
    SKIP_IF_DONE();
    __shared__ c128 hs[kLdp];
    __shared__ c128 wt[kThreads];
    for (int k = threadIdx.x; k < cnt; k += kThreads) hs[k] = h[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int64_t base = blockIdx.x * chunk, end = min(n, base + chunk);
    double re[16], im[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) re[c] = im[c] = 0;
    for (int64_t t0 = base; t0 < end; t0 += kThreads) {
        const int64_t r = t0 + threadIdx.x;
        c128 v = mk(0.0, 0.0);
        if (r < end) {
            v = w[r];
            for (int k = 0; k < cnt; ++k) v = v - hs[k] * V[k * ldv + r];
            w[r] = v;
        }
        wt[threadIdx.x] = v;
        __syncthreads();
        for (int u = 0; u < 4; ++u) {
            const int row = lane + 64 * u;
            const int64_t rr = t0 + row;
            if (rr < end) {
                const c128 wr = wt[row];
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    const int k = q + 4 * c;
                    if (k < cnt) {
                        const c128 vv = V[k * ldv + rr];
                        re[c] += vv.x * wr.x + vv.y * wr.y;
                        im[c] += vv.x * wr.y - vv.y * wr.x;
                    }
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int k = q + 4 * c;
        if (k < cnt) {
            const double a = wave_sum(re[c]), b = wave_sum(im[c]);
            if (lane == 0) partial[blockIdx.x * kLdp + k] = mk(a, b);
        }
    }
}

// w -= V[:, 0:cnt] h; partial[b * kLdp] = |w|^2 over the block's rows.
__global__ void __launch_bounds__(kThreads) k_cgs_update_norm(const c128 *__restrict__ V, int64_t ldv, int cnt,
                                                              const c128 *__restrict__ h, c128 *__restrict__ w,
                                                              int64_t n, int64_t chunk, c128 *__restrict__ partial,
                                                              const int *__restrict__ skip) {
    SKIP_IF_DONE();
    __shared__ c128 hs[kLdp];
    __shared__ double sh[8];
    for (int k = threadIdx.x; k < cnt; k += kThreads) hs[k] = h[k];
    __syncthreads();
    const int64_t base = blockIdx.x * chunk, end = min(n, base + chunk);
    double s = 0;
    for (int64_t r = base + threadIdx.x; r < end; r += kThreads) {
        c128 v = w[r];
        for (int k = 0; k < cnt; ++k) v = v - hs[k] * V[k * ldv + r];
        w[r] = v;
        s += abs2(v);
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x * kLdp] = mk(s, 0.0);
}

// Step j (0-based) of the cycle, iteration `it` overall: the new Hessenberg column h1 + h2 and |w|, the previous
// rotations, the new rotation, the rotated right-hand side g, the monitored residual |g(j+1)| / reference and the
// stopping test (iterate_.m:252-278); inv_sub = 1 / H(j+1, j) for the next basis vector.
__global__ void __launch_bounds__(kThreads) k_gmres_finalize(const c128 *__restrict__ partial, int blocks,
                                                             const c128 *__restrict__ h1,
                                                             const c128 *__restrict__ h2, c128 *__restrict__ H,
                                                             int ldh, c128 *__restrict__ g, c128 *__restrict__ cs,
                                                             c128 *__restrict__ sn, int j, int it, double reference,
                                                             double tol, int max_iter, LoopState *st,
                                                             double *history, c128 *inv_sub,
                                                             const int *__restrict__ skip) {
    SKIP_IF_DONE();
    __shared__ double sh[8];
    double s = 0;
    for (int b = threadIdx.x; b < blocks; b += kThreads) s += partial[b * kLdp].x;
    s = block_sum(s, sh);
    if (threadIdx.x != 0) return;
    c128 *col = H + static_cast<int64_t>(ldh) * j;
    for (int k = 0; k <= j; ++k) col[k] = h1[k] + h2[k];
    const double sub = sqrt(s);
    col[j + 1] = mk(sub, 0.0);
    for (int k = 0; k < j; ++k) {
        const c128 a0 = col[k], a1 = col[k + 1];
        col[k] = cmulc(cs[k], a0) + cmulc(sn[k], a1);
        col[k + 1] = cs[k] * a1 - sn[k] * a0;
    }
    const c128 a = col[j];
    const double radius = sqrt(abs2(a) + sub * sub);
    c128 c = mk(1.0, 0.0), sv = mk(0.0, 0.0);
    if (radius > 0) {
        c = (1.0 / radius) * a;
        sv = mk(sub / radius, 0.0);
    }
    cs[j] = c;
    sn[j] = sv;
    col[j] = cmulc(c, col[j]) + cmulc(sv, mk(sub, 0.0));
    col[j + 1] = mk(0.0, 0.0);
    g[j + 1] = mk(0.0, 0.0) - sv * g[j];
    g[j] = cmulc(c, g[j]);
    const double res = sqrt(abs2(g[j + 1])) / reference;
    history[it] = res;
    st->steps = j + 1;
    st->iterations = it;
    const bool conv = res < tol;
    if (conv) {
        st->converged = 1;
        st->done = 1;
    } else if (it >= max_iter || !isfinite(res)) {
        st->done = 1;
    }
    *inv_sub = (!conv && sub != 0) ? mk(1.0 / sub, 0.0) : mk(0.0, 0.0);
}

// g = (rnorm, 0, ..., 0) at the start of a cycle.
__global__ void k_gmres_init(c128 *__restrict__ g, int m, double rnorm) {
    for (int k = threadIdx.x; k <= m; k += blockDim.x) g[k] = mk(k == 0 ? rnorm : 0.0, 0.0);
}

// y = H(0:steps, 0:steps) \ g(0:steps) (upper triangular after the rotations), steps read from the loop state.
__global__ void k_gmres_solve(const c128 *__restrict__ H, int ldh, const c128 *__restrict__ g,
                              const LoopState *st, c128 *__restrict__ y) {
    if (threadIdx.x != 0) return;
    const int steps = st->steps;
    for (int k = steps - 1; k >= 0; --k) {
        c128 acc = g[k];
        for (int l = k + 1; l < steps; ++l) acc = acc - H[k + static_cast<int64_t>(ldh) * l] * y[l];
        const c128 d = H[k + static_cast<int64_t>(ldh) * k];
        y[k] = (1.0 / abs2(d)) * cmulc(d, acc);
    }
}

// x += V[:, 0:steps] y, steps read from the loop state.
__global__ void k_gemv_add(const c128 *__restrict__ V, int64_t ldv, const LoopState *st,
                           const c128 *__restrict__ y, c128 *__restrict__ x, int64_t n) {
    __shared__ c128 ys[kLdp];
    const int cnt = st->steps;
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) ys[k] = y[k];
    __syncthreads();
    GRID_STRIDE(i, n) {
        c128 v = x[i];
        for (int k = 0; k < cnt; ++k) v = v + ys[k] * V[k * ldv + i];
        x[i] = v;
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Host side.
// ------------------------------------------------------------------------------------------------------------------
static unsigned vec_blocks(int64_t n) { return blocks_for(n, 256, 4096); }

Workspace::Workspace(Context *c, int64_t n_, int max_iter) : ctx(c), s(c->stream), n(n_) {
    int64_t b = (n + kThreads - 1) / kThreads;
    if (b > kMaxBlocks) b = kMaxBlocks;
    if (b < 1) b = 1;
    int64_t per = (n + b - 1) / b;
    chunk = (per + kThreads - 1) / kThreads * kThreads;
    blocks = static_cast<int>((n + chunk - 1) / chunk);
    if (blocks < 1) blocks = 1;
    partial = static_cast<c128 *>(c->scratch(kSlotPartial, sizeof(c128) * kLdp * kMaxBlocks, "the solver's scratch"));
    scal = static_cast<c128 *>(c->scratch(kSlotScal, sizeof(c128) * 256, "the solver's scalars"));
    state = static_cast<LoopState *>(c->scratch(kSlotState, sizeof(LoopState), "the solver's scalars"));
    history_cap = max_iter + 2;
    history = static_cast<double *>(c->scratch(kSlotHistory, sizeof(double) * history_cap, "the solver's history"));
    if (!c->host_pinned) RT_HIP(hipHostMalloc(&c->host_pinned, 4096, hipHostMallocDefault));
    host = static_cast<c128 *>(c->host_pinned);
}

void Workspace::dot(const c128 *x, const c128 *y, bool conj, int slot) {
    if (conj)
        hipLaunchKernelGGL(k_dot<true>, dim3(blocks), dim3(kThreads), 0, s, x, y, n, chunk, partial, ctx->skip);
    else
        hipLaunchKernelGGL(k_dot<false>, dim3(blocks), dim3(kThreads), 0, s, x, y, n, chunk, partial, ctx->skip);
    RT_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kThreads), 0, s, partial, blocks, scal + slot, 0, ctx->skip);
    RT_LAUNCH_CHECK();
}

void Workspace::nrm2(const c128 *x, int slot) {
    hipLaunchKernelGGL(k_nrm2, dim3(blocks), dim3(kThreads), 0, s, x, n, chunk, partial, ctx->skip);
    RT_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kThreads), 0, s, partial, blocks, scal + slot, 1, ctx->skip);
    RT_LAUNCH_CHECK();
}

void Workspace::fetch(int first, int count) {
    RT_HIP(hipMemcpyAsync(host + first, scal + first, count * sizeof(c128), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
}

LoopState Workspace::read_state() {
    LoopState *h = reinterpret_cast<LoopState *>(reinterpret_cast<char *>(host) + 256 * sizeof(c128) - 64);
    RT_HIP(hipMemcpyAsync(h, state, sizeof(LoopState), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    return *h;
}

void Workspace::read_history(int first, int count, std::vector<double> &out) {
    if (count <= 0) return;
    std::vector<double> tmp(count);
    RT_HIP(hipMemcpy(tmp.data(), history + first, count * sizeof(double), hipMemcpyDeviceToHost));
    out.insert(out.end(), tmp.begin(), tmp.end());
}

void Workspace::copy(const c128 *x, c128 *y) {
    hipLaunchKernelGGL(k_copy, dim3(vec_blocks(n)), dim3(256), 0, s, x, y, n, ctx->skip);
    RT_LAUNCH_CHECK();
}
void Workspace::scale(const c128 *x, c128 a, c128 *y) {
    hipLaunchKernelGGL(k_scale, dim3(vec_blocks(n)), dim3(256), 0, s, x, a, y, n, ctx->skip);
    RT_LAUNCH_CHECK();
}
void Workspace::sub(const c128 *b, const c128 *ax, c128 *r) {
    hipLaunchKernelGGL(k_sub, dim3(vec_blocks(n)), dim3(256), 0, s, b, ax, r, n, ctx->skip);
    RT_LAUNCH_CHECK();
}
void Workspace::axpy(c128 a, const c128 *x, c128 *y) {
    hipLaunchKernelGGL(k_axpy, dim3(vec_blocks(n)), dim3(256), 0, s, a, x, y, n, ctx->skip);
    RT_LAUNCH_CHECK();
}
void Workspace::zero(c128 *x) { RT_HIP(hipMemsetAsync(x, 0, n * sizeof(c128), s)); }

// Points the products' kernels at the loop's stop flag while a loop runs.
struct SkipGuard {
    Context *ctx;
    SkipGuard(Context *c, const LoopState *st) : ctx(c) { ctx->skip = &st->done; }
    ~SkipGuard() { ctx->skip = nullptr; }
};

static c128 *vec(Context *ctx, int slot, int64_t n, int64_t count = 1) {
    return static_cast<c128 *>(ctx->scratch(slot, sizeof(c128) * n * count, "the solver's vectors"));
}

// ------------------------------------------------------------------------------------------------------------------
// COCG (ConjugateGradient.m:32-59). Scalar slots: 0/1 rho (alternating), 2 d^T A d, 3 |r|.
// ------------------------------------------------------------------------------------------------------------------
static InnerResult cocg(Workspace &ws, const LinOp &A, const LinOp *P, const c128 *b, c128 *x, bool has_x0,
                        double tol, int max_iter, double bnorm) {
    InnerResult res;
    Context *ctx = ws.ctx;
    const int64_t n = ws.n;
    c128 *r = vec(ctx, kSlotR, n), *d = vec(ctx, kSlotD, n), *ad = vec(ctx, kSlotAd, n);
    c128 *z = P ? vec(ctx, kSlotZ, n) : nullptr;
    if (has_x0) {
        A(x, ad);
        ws.sub(b, ad, r);
    } else {
        ws.zero(x);
        ws.copy(b, r);
    }
    ws.nrm2(r, 3);
    ws.fetch(3, 1);
    res.norms.push_back(ws.host[3].x / bnorm);
    if (res.norms.back() < tol) {
        res.converged = true;
        return res;
    }
    hipLaunchKernelGGL(k_zero_state, dim3(1), dim3(64), 0, ws.s, ws.state);
    RT_LAUNCH_CHECK();
    SkipGuard guard(ctx, ws.state);
    int it = 0;
    LoopState st{};
    while (it < max_iter) {
        const int window = std::min(kPoll, max_iter - it);
        for (int w = 0; w < window; ++w) {
            ++it;
            const c128 *zz = r;
            if (P) {
                (*P)(r, z);
                zz = z;
            }
            const int i_new = it & 1, i_old = 1 - i_new;
            ws.dot(r, zz, false, i_new);
            if (it == 1) {
                ws.copy(zz, d);
            } else {
                hipLaunchKernelGGL(k_cocg_direction, dim3(vec_blocks(n)), dim3(256), 0, ws.s, zz, d, ws.scal, i_new,
                                   i_old, n, ctx->skip);
                RT_LAUNCH_CHECK();
            }
            A(d, ad);
            ws.dot(d, ad, false, 2);
            hipLaunchKernelGGL(k_cocg_update, dim3(ws.blocks), dim3(kThreads), 0, ws.s, x, r, d, ad, ws.scal, i_new, 2,
                               n, ws.chunk, ws.partial, ctx->skip);
            RT_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_cocg_finalize, dim3(1), dim3(kThreads), 0, ws.s, ws.partial, ws.blocks, bnorm, tol,
                               it, max_iter, ws.state, ws.history, ctx->skip);
            RT_LAUNCH_CHECK();
        }
        st = ws.read_state();
        if (st.done) break;
    }
    ws.read_history(1, st.iterations, res.norms);
    res.converged = st.converged;
    res.iterations = st.iterations;
    return res;
}

// ------------------------------------------------------------------------------------------------------------------
// GMRES(m), left preconditioned, CGS2 (iterate_.m:1-85). Scalar slots: [0, 64) second-pass coefficients, [64, 128)
// first-pass coefficients, 128 |r|, 130 1 / H(j+1, j), [160, 224) the solution coefficients y.
// ------------------------------------------------------------------------------------------------------------------
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
    c128 *V = vec(ctx, kSlotV, n, m + 1), *w = vec(ctx, kSlotW, n), *t = vec(ctx, kSlotT, n);
    c128 *H = vec(ctx, kSlotH, m + 1, m);
    c128 *gcs = vec(ctx, kSlotG, m + 1, 3);
    c128 *g = gcs, *cs = gcs + (m + 1), *sn = gcs + 2 * (m + 1);
    auto precondition = [&](const c128 *in, c128 *out) {
        if (P) (*P)(in, out);
        else if (in != out) ws.copy(in, out);
    };
    precondition(b, w);   // reference = norm(P b)
    ws.nrm2(w, 128);
    ws.fetch(128, 1);
    const double reference = ws.host[128].x;
    if (has_x0) {
        A(x, t);
        ws.sub(b, t, t);
        precondition(t, w);
        ws.nrm2(w, 128);
        ws.fetch(128, 1);
    } else {
        ws.zero(x);
    }
    double rnorm = ws.host[128].x;
    res.norms.push_back(rnorm / reference);
    bool conv = res.norms.back() < tol;
    int it = 0;
    hipLaunchKernelGGL(k_zero_state, dim3(1), dim3(64), 0, ws.s, ws.state);
    RT_LAUNCH_CHECK();
    while (!conv && it < max_iter) {
        ws.scale(w, mk(1.0 / rnorm, 0.0), V);
        hipLaunchKernelGGL(k_gmres_init, dim3(1), dim3(64), 0, ws.s, g, m, rnorm);
        RT_LAUNCH_CHECK();
        const int it0 = it;
        LoopState st{};
        {
            SkipGuard guard(ctx, ws.state);
            int step = 0;
            while (step < m && it < max_iter) {
                const int window = std::min({kPoll, m - step, max_iter - it});
                for (int k = 0; k < window; ++k) {
                    const int j = step;
                    ++step;
                    ++it;
                    const c128 *vj = V + static_cast<int64_t>(j) * n;
                    A(vj, t);
                    precondition(t, w);
                    hipLaunchKernelGGL(k_cgs_dot, dim3(ws.blocks), dim3(kThreads), 0, ws.s, V, n, step, w, n,
                                       ws.chunk, ws.partial, ctx->skip);
                    RT_LAUNCH_CHECK();
                    hipLaunchKernelGGL(k_reduce, dim3(step), dim3(kThreads), 0, ws.s, ws.partial, ws.blocks,
                                       ws.scal + 64, 0, ctx->skip);
                    RT_LAUNCH_CHECK();
                    hipLaunchKernelGGL(k_cgs_update_dot, dim3(ws.blocks), dim3(kThreads), 0, ws.s, V, n, step,
                                       ws.scal + 64, w, n, ws.chunk, ws.partial, ctx->skip);
                    RT_LAUNCH_CHECK();
                    hipLaunchKernelGGL(k_reduce, dim3(step), dim3(kThreads), 0, ws.s, ws.partial, ws.blocks, ws.scal,
                                       0, ctx->skip);
                    RT_LAUNCH_CHECK();
                    hipLaunchKernelGGL(k_cgs_update_norm, dim3(ws.blocks), dim3(kThreads), 0, ws.s, V, n, step,
                                       ws.scal, w, n, ws.chunk, ws.partial, ctx->skip);
                    RT_LAUNCH_CHECK();
                    hipLaunchKernelGGL(k_gmres_finalize, dim3(1), dim3(kThreads), 0, ws.s, ws.partial, ws.blocks,
                                       ws.scal + 64, ws.scal, H, m + 1, g, cs, sn, j, it, reference, tol, max_iter,
                                       ws.state, ws.history, ws.scal + 130, ctx->skip);
                    RT_LAUNCH_CHECK();
                    if (step < m + 1) {
                        hipLaunchKernelGGL(k_scale_dev, dim3(vec_blocks(n)), dim3(256), 0, ws.s, w, ws.scal + 130,
                                           V + static_cast<int64_t>(step) * n, n, ctx->skip);
                        RT_LAUNCH_CHECK();
                    }
                }
                st = ws.read_state();
                if (st.done) break;
            }
        }
        it = st.iterations;
        ws.read_history(it0 + 1, it - it0, res.norms);
        conv = st.converged;
        // x += V y with y = H \ g over the steps of this cycle.
        hipLaunchKernelGGL(k_gmres_solve, dim3(1), dim3(64), 0, ws.s, H, m + 1, g, ws.state, ws.scal + 160);
        RT_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_gemv_add, dim3(vec_blocks(n)), dim3(256), 0, ws.s, V, n, ws.state, ws.scal + 160, x, n);
        RT_LAUNCH_CHECK();
        if (!conv && it < max_iter && !(st.done && !std::isfinite(res.norms.back()))) {
            // The next cycle starts from the true preconditioned residual.
            A(x, t);
            ws.sub(b, t, t);
            precondition(t, w);
            ws.nrm2(w, 128);
            ws.fetch(128, 1);
            rnorm = ws.host[128].x;
            hipLaunchKernelGGL(k_zero_state, dim3(1), dim3(64), 0, ws.s, ws.state);
            RT_LAUNCH_CHECK();
            // Keep the global iteration count in the state for the next cycle's finalize kernels.
        } else {
            break;
        }
    }
    RT_HIP(hipStreamSynchronize(ws.s));
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
    Workspace ws(ctx, n, o.max_iterations);
    hipStream_t s = ctx->stream;
    SolveResult out;
    ws.nrm2(b, 200);
    ws.fetch(200, 1);
    const double bnorm = ws.host[200].x;
    c128 *t = vec(ctx, kSlotOuterT, n);
    // The outer residual runs in FP64 whenever the operator holds FP64 spectra.
    const int outer_prec = op.has(RT_C128) ? RT_C128 : RT_C64;
    auto true_residual = [&](const c128 *xx) {
        op.apply(xx, t, 1, outer_prec, s);
        ws.sub(b, t, t);
        ws.nrm2(t, 201);
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
        LinOp A = [&](const c128 *in, c128 *res) { op.apply(in, res, 1, inner_prec, s, ctx->skip); };
        LinOp Pop = [&](const c128 *in, c128 *res) { pc->apply(in, res, 1, inner_prec, s, ctx->skip); };
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
            if (x0 && x0 != x) ws.copy(x0, x);
            InnerResult r = run(b, x, x0 != nullptr, o.tolerance, o.max_iterations);
            out.converged = r.converged;
            out.iterations = r.iterations;
            out.norms = r.norms;
            out.residual = true_residual(x);
        } else {
            // x_{k+1} = x_k + d_k, A d_k = b - A x_k solved by the inner method in its precision.
            if (x0) ws.copy(x0, x);
            else ws.zero(x);
            c128 *r = vec(ctx, kSlotOuterR, n), *d = vec(ctx, kSlotOuterD, n);
            double rel = true_residual(x);
            out.norms.push_back(rel);
            // Inner tolerance: complex64 GMRES reaches about 1e-6 before the rounding of its products shows; COCG's
            // short recurrence loses the complex symmetry of complex64 products near 1e-4.
            double inner_tol = o.tolerance;
            if (o.inner_tolerance > 0) inner_tol = o.inner_tolerance;
            else if (o.precision == RT_C64) inner_tol = std::max(o.tolerance, o.method == RT_METHOD_CG ? 1e-4 : 1e-6);
            int total = 0;
            while (!(rel < o.tolerance) && out.outer < o.max_outer && total < o.max_iterations) {
                ++out.outer;
                ws.copy(t, r);   // t holds b - A x from true_residual
                // Ask the inner solve for what is left to reach the tolerance, but no less than its floor.
                double want = 0.5 * o.tolerance / rel;
                if (want < inner_tol) want = inner_tol;
                if (want > 0.5) want = 0.5;
                InnerResult ir = run(r, d, false, want, o.max_iterations - total);
                total += ir.iterations;
                for (size_t k = 1; k < ir.norms.size(); ++k) out.norms.push_back(ir.norms[k] * rel);
                ws.axpy(mk(1.0, 0.0), d, x);
                const double next = true_residual(x);
                rel = next;
                if (!std::isfinite(next)) break;
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

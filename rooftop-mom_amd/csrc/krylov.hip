// The preconditioned Krylov loops on gfx950: COCG and restarted GMRES(m) on R right-hand sides at once, FP64
// vectors on the device, and the FP64 iterative-refinement loop around an inner solve whose products run in
// complex64.
//
// Reference: src/+rooftopMoM/+solver/IterativeMethod.m:27-70, ConjugateGradient.m:32-59,
// PreconditionedConjugateGradient.m:11-23, @GeneralizedMinimumResidual/iterate_.m:1-85, orthogonalize_.m:15-23,
// generateRotation_.m:11-18. The reference solves one right-hand side per call; here R systems with the same operator
// advance in lockstep, every kernel covering all of them (blockIdx.y = system), and every product applied to the R
// vectors in one launch per pass, so the spectra and the near-field coefficients are read once per step for all R.
//
// Everything a step needs stays on the device: the scalars (rho, d^T A d, the Hessenberg column), the Givens rotations,
// the residual histories and the stopping tests, written by one block per system (a "finalize" kernel, or in GMRES
// the update pass's first block). The host enqueues a window of steps and polls the R loop states once per window. A
// system that stops (convergence, breakdown, the iteration limit) raises its flag, and every later kernel of that
// system -- products, preconditioner, vector kernels -- returns at once, so each system's result is exactly that of
// stopping at its own step. Reductions are
// deterministic: fixed row blocks, wave64 shuffles and a fixed-order second stage.
//
// The Arnoldi step orthogonalizes twice, as CGS2 does (orthogonalize_.m:15-23), but with the second pass delayed by
// one step (DCGS2, below): two row-partitioned passes over the live columns only per step instead of three. The bases
// are stored column by column across the systems, V[k][r][n]; the vector the product is applied to is the pending
// FP64 vector p of every system, one batch of R vectors.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "krylov.hpp"

namespace rt {

constexpr int kThreads = 256;
constexpr int kLdp = 64;        // partial columns per block
constexpr int kPoll = 8;        // steps enqueued between two polls of the loop states
constexpr int kPollMax = 32;    // GMRES: the most steps enqueued between two polls
constexpr int kStall = 40;      // COCG steps without a new best norm that end an inner solve of the refinement
// Block COCG inner solve: a norm this far above the best ends the system. The block recurrence's first steps swing
// far above the starting norm (a Galerkin step on an indefinite projection), a lost recurrence grows without bound.
constexpr double kBlockGrowth = 1e4;
constexpr double kBlockInner = 1e-2;   // the reduction one inner block COCG cycle of the refinement aims at
// Mixed COCG: the first inner solve in complex64, to kCocgInner or a stall, and the later ones in FP64. Complex64
// COCG loses the complex symmetry of its products on systems whose residual swings high (on the 512 x 512 plate
// under grazing plane waves it stalls between 1e-2 and 1e-4); the FP64 inner solve restarted from the refined
// residual converges where a single FP64 COCG from zero may not within its iteration limit.
constexpr bool kCocgFallback = true;
constexpr double kCocgInner = 1e-2;

// Scalar slots of a system.
enum ScalSlot {
    kRho0 = 0, kRho1 = 1, kDad = 2, kRnormCg = 3,   // COCG
    kH2 = 0, kH1 = 64,                              // GMRES: second- and first-pass coefficients
    kRnorm = 128, kInvSub = 130, kReference = 132, kTol = 134, kBnorm = 136,
    kMaxIt = 138, kBest = 140,                      // per-system iteration limit; COCG best monitored norm
    kAlpha = 142, kT = 143, kOmega = 144,           // GMRES: p^H p, p^H w, w^H w
    kB2 = 145, kBeta = 146, kNu = 147, kZj = 148,   // GMRES: the step's beta^2, beta, nu, z_j
    kInvBeta = 149, kInvBetaNu = 150,               // GMRES: the update pass's coefficients
    kY = 256,                                       // GMRES: solution coefficients of a cycle
    kOuterB = 400, kOuterRes = 401
};

// Pool slots of the context's solver buffers.
enum PoolSlot {
    kSlotPartial, kSlotScal, kSlotState, kSlotHistory, kSlotV, kSlotW, kSlotT, kSlotR, kSlotZ, kSlotD, kSlotAd,
    kSlotH, kSlotG, kSlotOuterT, kSlotOuterR, kSlotOuterD, kSlotFlags, kSlotP, kSlotHraw,
    kSlotPartialL, kSlotHc, kSlotY, kSlotBlockCoef, kSlotBlockPartial, kSlotBlockBest, kSlotOuterSolo
};

// Restarts up to this length run the delayed-reorthogonalization kernels, whose Gram-Schmidt coefficients, rotations
// and triangular solve fit one wave; longer ones (up to the iteration limit, restart = Inf) the CGS2 kernels below.
constexpr int kFastRestart = 62;

// ------------------------------------------------------------------------------------------------------------------
// Reductions.
// ------------------------------------------------------------------------------------------------------------------
// Cross-lane moves of FP64 values as two 32-bit halves. dpp_f64: a DPP move (CTRL) into the rows of ROW_MASK; a
// lane without a source, or outside the rows, gets `old`. readlane_f64: the value of the uniform lane k.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_f64(double old, double v) {
    const unsigned long long o = __builtin_bit_cast(unsigned long long, old);
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo =
        __builtin_amdgcn_update_dpp(static_cast<int>(o), static_cast<int>(b), CTRL, ROW_MASK, 0xf, false);
    const unsigned hi =
        __builtin_amdgcn_update_dpp(static_cast<int>(o >> 32), static_cast<int>(b >> 32), CTRL, ROW_MASK, 0xf, false);
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
__device__ __forceinline__ double readlane_f64(double v, int k) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane(static_cast<int>(b), k);
    const unsigned hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), k);
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// Sum over the wave, in every lane (the GMRES Gram-Schmidt kernels): an inclusive scan by DPP moves (row_shr 1, 2, 4, 8 within the rows of 16 lanes,
// then row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3) leaves the total in lane 63.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0x111>(0.0, v);
    v += dpp_f64<0x112>(0.0, v);
    v += dpp_f64<0x114>(0.0, v);
    v += dpp_f64<0x118>(0.0, v);
    v += dpp_f64<0x142, 0xa>(0.0, v);
    v += dpp_f64<0x143, 0xc>(0.0, v);
    return readlane_f64(v, 63);
}

// Sum over the wave by a shuffle tree, valid in lane 0: the vector reductions and COCG keep this order (COCG's
// short recurrence on these systems is sensitive to the rounding of its inner products).
__device__ __forceinline__ double wave_sum_tree(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum over the block (all threads call it); valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *sh) {
    v = wave_sum_tree(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (blockDim.x + 63) / 64; ++w) r += sh[w];
    return r;
}

// The system of this block; returns from the kernel when that system is stopped (uniform over the block).
#define SYSTEM(r)             \
    const int r = blockIdx.y; \
    if (st && st[r].done) return

// a where c holds, else b, component by component (a select of whole structures would go through memory).
__device__ __forceinline__ c128 pick(bool c, c128 a, c128 b) { return mk(c ? a.x : b.x, c ? a.y : b.y); }

// partial[(r * blocks + b) * kLdp] = sum over block b's rows of system r of x^T y (or x^H y).
template <bool CONJ>
__global__ void __launch_bounds__(kThreads) k_dot(const c128 *__restrict__ x, const c128 *__restrict__ y, int64_t n,
                                                  int64_t chunk, c128 *__restrict__ partial,
                                                  const LoopState *__restrict__ st) {
    SYSTEM(r);
    __shared__ double sh[8];
    const int64_t base = blockIdx.x * chunk, end = min(n, base + chunk);
    const c128 *xr = x + r * n, *yr = y + r * n;
    double re = 0, im = 0;
    for (int64_t i = base + threadIdx.x; i < end; i += kThreads) {
        const c128 p = CONJ ? cmulc(xr[i], yr[i]) : xr[i] * yr[i];
        re += p.x;
        im += p.y;
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    if (threadIdx.x == 0) partial[(static_cast<int64_t>(r) * gridDim.x + blockIdx.x) * kLdp] = mk(re, im);
}

__global__ void __launch_bounds__(kThreads) k_nrm2(const c128 *__restrict__ x, int64_t n, int64_t chunk,
                                                   c128 *__restrict__ partial, const LoopState *__restrict__ st) {
    SYSTEM(r);
    __shared__ double sh[8];
    const int64_t base = blockIdx.x * chunk, end = min(n, base + chunk);
    const c128 *xr = x + r * n;
    double s = 0;
    for (int64_t i = base + threadIdx.x; i < end; i += kThreads) s += abs2(xr[i]);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[(static_cast<int64_t>(r) * gridDim.x + blockIdx.x) * kLdp] = mk(s, 0.0);
}

// scal[r][slot + k] = sum_b partial[(r * blocks + b) * kLdp + k], one block per (k, r); sqrt_re: a norm.
__global__ void __launch_bounds__(kThreads) k_reduce(const c128 *__restrict__ partial, int blocks,
                                                     c128 *__restrict__ scal, int slot, int sqrt_re,
                                                     const LoopState *__restrict__ st) {
    SYSTEM(r);
    __shared__ double sh[8];
    const int k = blockIdx.x;
    // Rounds of eight loads per thread from addresses clamped into range (a load under a branch is waited for inside
    // it), summed in the order of the blocks.
    const c128 *part = partial + static_cast<int64_t>(r) * blocks * kLdp + k;
    double re = 0, im = 0;
    for (int b0 = 0; b0 < blocks; b0 += 8 * kThreads) {
        c128 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int b = b0 + static_cast<int>(threadIdx.x) + u * kThreads;
            v[u] = pick(b < blocks, part[static_cast<int64_t>(min(b, blocks - 1)) * kLdp], mk(0.0, 0.0));
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            re += v[u].x;
            im += v[u].y;
        }
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    if (threadIdx.x == 0) scal[r * kScal + slot + k] = sqrt_re ? mk(sqrt(re), 0.0) : mk(re, im);
}

// ------------------------------------------------------------------------------------------------------------------
// Vector kernels (grid.y = system).
// ------------------------------------------------------------------------------------------------------------------
#define ROWS(i, n)                                                                         \
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < (n); \
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)

__global__ void k_copy(const c128 *__restrict__ x, c128 *__restrict__ y, int64_t n, const LoopState *__restrict__ st) {
    SYSTEM(r);
    const int64_t o = r * n;
    ROWS(i, n) y[o + i] = x[o + i];
}
__global__ void k_sub(const c128 *__restrict__ b, const c128 *__restrict__ ax, c128 *__restrict__ res, int64_t n,
                      const LoopState *__restrict__ st) {
    SYSTEM(r);
    const int64_t o = r * n;
    ROWS(i, n) res[o + i] = b[o + i] - ax[o + i];
}
__global__ void k_add(const c128 *__restrict__ x, c128 *__restrict__ y, int64_t n, const LoopState *__restrict__ st) {
    SYSTEM(r);
    const int64_t o = r * n;
    ROWS(i, n) y[o + i] = y[o + i] + x[o + i];
}
// Per-system state reset: stopped = finished[r] (systems left out of this loop), everything else zero.
__global__ void k_reset_state(LoopState *st, const int *__restrict__ finished, int R) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int f = finished[r];
    st[r] = LoopState{f, 0, f, 0, 0, 0, 0, -1, 0.f};
}
// Marks the systems whose starting residual is already below their tolerance as converged at iteration 0.
__global__ void k_initial_check(LoopState *st, const c128 *__restrict__ scal, int num_slot, int den_slot, int R) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R || st[r].done) return;
    const double res = scal[r * kScal + num_slot].x / scal[r * kScal + den_slot].x;
    st[r].monitored = static_cast<float>(res);
    if (res < scal[r * kScal + kTol].x) st[r].done = st[r].finished = st[r].converged = 1;
}
__global__ void k_copy_slot(c128 *scal, int from, int to, int R) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) scal[r * kScal + to] = scal[r * kScal + from];
}

// ------------------------------------------------------------------------------------------------------------------
// COCG kernels.
// ------------------------------------------------------------------------------------------------------------------
// d = z + (rho_new / rho_old) d.
__global__ void k_cocg_direction(const c128 *__restrict__ z, c128 *__restrict__ d, const c128 *__restrict__ scal,
                                 int i_new, int i_old, int64_t n, const LoopState *__restrict__ st) {
    SYSTEM(r);
    const c128 rn = scal[r * kScal + i_new], ro = scal[r * kScal + i_old];
    const c128 beta = (1.0 / abs2(ro)) * cmulc(ro, rn);
    const int64_t o = r * n;
    ROWS(i, n) d[o + i] = z[o + i] + beta * d[o + i];
}

// alpha = rho / (d^T A d); x += alpha d; r -= alpha Ad; per-block partial of |r|^2.
__global__ void __launch_bounds__(kThreads) k_cocg_update(c128 *__restrict__ x, c128 *__restrict__ res,
                                                          const c128 *__restrict__ d, const c128 *__restrict__ ad,
                                                          const c128 *__restrict__ scal, int i_rho, int64_t n,
                                                          int64_t chunk, c128 *__restrict__ partial,
                                                          const LoopState *__restrict__ st) {
    SYSTEM(r);
    __shared__ double sh[8];
    const c128 rho = scal[r * kScal + i_rho], dad = scal[r * kScal + kDad];
    const c128 alpha = (1.0 / abs2(dad)) * cmulc(dad, rho);
    const int64_t base = blockIdx.x * chunk, end = min(n, base + chunk), o = r * n;
    double s = 0;
    for (int64_t i = base + threadIdx.x; i < end; i += kThreads) {
        x[o + i] = x[o + i] + alpha * d[o + i];
        const c128 v = res[o + i] - alpha * ad[o + i];
        res[o + i] = v;
        s += abs2(v);
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[(static_cast<int64_t>(r) * gridDim.x + blockIdx.x) * kLdp] = mk(s, 0.0);
}

// The residual norm of iteration `it`, the stopping test (ConjugateGradient.m:126-136) with the system's own
// iteration limit. stall > 0 (inner solves of the refinement loop only) also stops a system whose monitored norm has
// not improved on its best for `stall` iterations: a complex64 short recurrence that has lost its complex symmetry to
// rounding, which the outer FP64 loop then restarts from the true residual. Grid (1, R). In a block loop (block = 1)
// a system that stops is marked finished only: its residual stays in the block's search space, and it is skipped here
// from then on; with stall > 0 a block system also stops once its norm exceeds kBlockGrowth times its best (a block
// recurrence that has lost its conjugacy diverges, and its solution is kept before it strays far).
__global__ void __launch_bounds__(kThreads) k_cocg_finalize(const c128 *__restrict__ partial, int blocks,
                                                            c128 *__restrict__ scal, int it, int stall,
                                                            LoopState *st, double *history, int cap, int block) {
    SYSTEM(r);
    if (block && st[r].finished) return;   // uniform over the block, before its first barrier
    __shared__ double sh[8];
    // Every load is issued before any is used, from addresses clamped into range (a load under a branch is waited
    // for inside it): the scalars, then the partials in rounds of eight per thread, summed in the order of the blocks.
    c128 *sc = scal + r * kScal;
    const double bnorm = sc[kBnorm].x, best = sc[kBest].x, tol = sc[kTol].x, maxit = sc[kMaxIt].x;
    const c128 *part = partial + static_cast<int64_t>(r) * blocks * kLdp;
    double s = 0;
    for (int b0 = 0; b0 < blocks; b0 += 8 * kThreads) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int b = b0 + static_cast<int>(threadIdx.x) + u * kThreads;
            const double x = part[static_cast<int64_t>(min(b, blocks - 1)) * kLdp].x;
            v[u] = b < blocks ? x : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    s = block_sum(s, sh);
    if (threadIdx.x != 0) return;
    const double res = sqrt(s) / bnorm;
    history[static_cast<int64_t>(r) * cap + it] = res;
    st[r].iterations = it;
    if (res < best) {
        sc[kBest] = mk(res, 0.0);
        st[r].best = it;
    }
    const int stop = block ? 0 : 1;
    if (res < tol) {
        st[r].converged = st[r].finished = 1;
        st[r].done = stop;
    } else if (!isfinite(res)) {
        st[r].finished = 1;
        st[r].done = stop;
    } else if (it >= static_cast<int>(maxit) || (stall > 0 && it - st[r].best >= stall) ||
               (block && stall > 0 && res > kBlockGrowth * best)) {
        if (block) st[r].finished = 1;
        else st[r].done = 1;
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Block COCG kernels: the R <= 64 right-hand sides of one complex symmetric system share their search space. The
// block vectors are the R systems' vectors (system-major, [r][n]); the R x R coefficient matrices are column-major
// with leading dimension kBlockMax, element (i, j) at [j * kBlockMax + i]. keep[j].x = 1 marks the directions kept
// by the last orthonormalization, 0 the dropped ones (zero columns of the direction block).
// ------------------------------------------------------------------------------------------------------------------
constexpr int kBlockMax = 64;   // right-hand sides of one block solve
constexpr int kBtRows = 16;     // rows of one staged tile of the transposed product
constexpr int kBlockLd = kBlockMax + 1;
constexpr size_t kBlockLds = 2 * kBlockMax * kBlockLd * sizeof(c128);   // two padded s x s matrices

// partial[blk] (element (i, j)) = sum over block blk's rows of U_i[k] V_j[k] (CONJ: conj(U_i[k]) V_j[k]), i, j < s:
// a tile of 16 rows of all s columns of U and V staged in LDS (coalesced runs of 16 rows per column), each thread
// accumulating 4 x 4 entries.
template <bool CONJ>
__global__ void __launch_bounds__(kThreads) k_block_tn(const c128 *__restrict__ U, const c128 *__restrict__ V,
                                                       int64_t n, int s, int64_t chunk, c128 *__restrict__ partial) {
    __shared__ c128 us[kBlockMax][kBtRows + 1], vs[kBlockMax][kBtRows + 1];
    const int t = threadIdx.x;
    const int ti = t >> 4, tj = t & 15;   // entries (4 ti + a, 4 tj + b)
    c128 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = mk(0.0, 0.0);
    const int64_t base = blockIdx.x * chunk, end = min(n, base + chunk);
    for (int64_t k0 = base; k0 < end; k0 += kBtRows) {
        __syncthreads();   // the previous tile is consumed
#pragma unroll
        for (int u = 0; u < kBlockMax * kBtRows / kThreads; ++u) {
            const int e = t + u * kThreads, i = e / kBtRows, kk = e % kBtRows;
            const int64_t k = k0 + kk;
            const bool in = i < s && k < end;
            const int64_t at = static_cast<int64_t>(min(i, s - 1)) * n + min(k, end - 1);
            const c128 uv = U[at], vv = V[at];
            us[i][kk] = pick(in, uv, mk(0.0, 0.0));
            vs[i][kk] = pick(in, vv, mk(0.0, 0.0));
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < kBtRows; ++kk) {
            c128 a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = us[4 * ti + q][kk];
                b[q] = vs[4 * tj + q][kk];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] = acc[p][q] + (CONJ ? cmulc(a[p], b[q]) : a[p] * b[q]);
        }
    }
    c128 *out = partial + static_cast<int64_t>(blockIdx.x) * kBlockMax * kBlockMax;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) out[(4 * tj + q) * kBlockMax + 4 * ti + p] = acc[p][q];
}

// C = sum over the blocks of partial, in the blocks' order.
__global__ void k_block_sum(const c128 *__restrict__ partial, int blocks, c128 *__restrict__ C) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kBlockMax * kBlockMax) return;
    c128 acc = mk(0.0, 0.0);
    for (int b = 0; b < blocks; ++b) acc = acc + partial[static_cast<int64_t>(b) * kBlockMax * kBlockMax + e];
    C[e] = acc;
}

// Every system still running ends (finished, and done for every system): with it >= 0 iteration it counts with a NaN
// monitored norm, as COCG's Inf or NaN residual ends its loop (ConjugateGradient.m:56-59). Thread-strided, s <= R.
__device__ void block_stop_all(LoopState *st, int s, int it, double *history, int cap) {
    for (int r = threadIdx.x; r < s; r += blockDim.x) {
        if (!st[r].finished) {
            if (it >= 0) {
                history[static_cast<int64_t>(r) * cap + it] = __builtin_nan("");
                st[r].iterations = it;
            }
            st[r].finished = 1;
        }
        st[r].done = 1;
    }
}

// X = M^-1 B on the kept directions: M's rows and columns of dropped directions are the identity's and B's rows there
// zero, so X is zero in those rows; every column of B (one per system) is solved. Gauss-Jordan with partial pivoting
// in LDS (kBlockLds dynamic), one block. A zero or non-finite pivot is a block breakdown: block_stop_all (it >= 0
// for the solve for alpha of iteration it).
__global__ void __launch_bounds__(kThreads) k_block_solve(const c128 *__restrict__ M, const c128 *__restrict__ B,
                                                          c128 *__restrict__ X, const c128 *__restrict__ keep, int s,
                                                          LoopState *st, int it, double *history, int cap) {
    extern __shared__ unsigned char solve_lds[];
    constexpr int ld = kBlockLd;
    c128 *m = reinterpret_cast<c128 *>(solve_lds), *b = m + kBlockMax * ld;   // row-major [row][col]
    __shared__ int piv;
    __shared__ int broken;
    const int t = threadIdx.x;
    for (int e = t; e < s * s; e += kThreads) {
        const int i = e % s, j = e / s;   // row i, column j (column-major input)
        const bool ki = keep[i].x > 0.5, kj = keep[j].x > 0.5;
        m[i * ld + j] = ki && kj ? M[j * kBlockMax + i] : mk(i == j ? 1.0 : 0.0, 0.0);
        b[i * ld + j] = ki ? B[j * kBlockMax + i] : mk(0.0, 0.0);
    }
    if (t == 0) broken = 0;
    __syncthreads();
    for (int k = 0; k < s; ++k) {
        if (t == 0) {
            int p = k;
            double best = abs2(m[k * ld + k]);
            for (int i = k + 1; i < s; ++i) {
                const double v = abs2(m[i * ld + k]);
                if (v > best) {
                    best = v;
                    p = i;
                }
            }
            piv = p;
            if (!(best > 0) || !isfinite(best)) broken = 1;
        }
        __syncthreads();
        if (broken) break;   // uniform: read after the barrier
        const int p = piv;
        if (p != k)
            for (int j = t; j < s; j += kThreads) {
                const c128 a = m[k * ld + j], c = b[k * ld + j];
                m[k * ld + j] = m[p * ld + j];
                b[k * ld + j] = b[p * ld + j];
                m[p * ld + j] = a;
                b[p * ld + j] = c;
            }
        __syncthreads();
        const c128 d = m[k * ld + k];
        const c128 inv = (1.0 / abs2(d)) * conjg(d);
        __syncthreads();
        for (int j = t; j < s; j += kThreads) {
            m[k * ld + j] = inv * m[k * ld + j];
            b[k * ld + j] = inv * b[k * ld + j];
        }
        __syncthreads();
        for (int e = t; e < s * s; e += kThreads) {   // every other row minus its multiple of row k
            const int i = e / s, j = e % s;
            if (i == k) continue;
            const c128 f = m[i * ld + k];
            if (j != k) m[i * ld + j] = m[i * ld + j] - f * m[k * ld + j];
            b[i * ld + j] = b[i * ld + j] - f * b[k * ld + j];
        }
        __syncthreads();
        for (int i = t; i < s; i += kThreads)
            if (i != k) m[i * ld + k] = mk(0.0, 0.0);
        __syncthreads();
    }
    if (broken) {
        block_stop_all(st, s, it, history, cap);
        return;
    }
    for (int e = t; e < s * s; e += kThreads) {
        const int i = e % s, j = e / s;
        X[j * kBlockMax + i] = b[i * ld + j];
    }
}

// The rank-revealing orthonormalization of a block W from its Gram matrix G = W^H W: the Cholesky factor L (G = L
// L^H), column by column, keeps column j when what is left of it after the earlier kept columns, d = G_jj - sum
// |L_jk|^2, exceeds tau^2 G_jj. Out: C = L^-H on the kept columns (upper triangular, zero elsewhere), so that Q = W C
// is orthonormal there and zero in the dropped columns; T (upper triangular) with W = Q T up to the dropped parts:
// T(i, j) = conj(L(j, i)) for kept i and j >= i, zero in the rows of dropped i; keep. With no column kept and
// stop_empty, every running system ends (nothing is left to search with). One block; L and C in LDS (kBlockLds).
__global__ void __launch_bounds__(kThreads) k_block_orth(const c128 *__restrict__ G, c128 *__restrict__ C,
                                                         c128 *__restrict__ T, c128 *__restrict__ keep, int s,
                                                         double tau, int stop_empty, LoopState *st, double *history,
                                                         int cap) {
    extern __shared__ unsigned char orth_lds[];
    constexpr int ld = kBlockLd;
    c128 *L = reinterpret_cast<c128 *>(orth_lds), *Cs = L + kBlockMax * ld;   // row-major [i][j]
    __shared__ double diag[kBlockMax];
    __shared__ int kept[kBlockMax];
    __shared__ int count;
    const int t = threadIdx.x;
    for (int e = t; e < s * s; e += kThreads) {
        const int i = e % s, j = e / s;
        L[i * ld + j] = G[j * kBlockMax + i];
    }
    if (t == 0) count = 0;
    __syncthreads();
    for (int j = 0; j < s; ++j) {
        if (t == 0) {
            const double g = L[j * ld + j].x;
            double d = g;
            for (int k = 0; k < j; ++k)
                if (kept[k]) d -= abs2(L[j * ld + k]);
            const int kj = isfinite(d) && d > tau * tau * g;
            kept[j] = kj;
            diag[j] = kj ? sqrt(d) : 0.0;
            count += kj;
        }
        __syncthreads();
        if (kept[j]) {
            const double inv = 1.0 / diag[j];
            for (int i = j + 1 + t; i < s; i += kThreads) {
                c128 v = L[i * ld + j];
                for (int k = 0; k < j; ++k)
                    if (kept[k]) v = v - L[i * ld + k] * conjg(L[j * ld + k]);
                L[i * ld + j] = inv * v;
            }
        }
        __syncthreads();
    }
    if (count == 0 && stop_empty) {
        block_stop_all(st, s, -1, history, cap);
        return;
    }
    // Column j of C solves L^H c = e_j (back substitution over the kept rows), one thread per column.
    if (t < s) {
        const int j = t;
        for (int i = 0; i < s; ++i) Cs[i * ld + j] = mk(0.0, 0.0);
        if (kept[j]) {
            Cs[j * ld + j] = mk(1.0 / diag[j], 0.0);
            for (int i = j - 1; i >= 0; --i) {
                if (!kept[i]) continue;
                c128 v = mk(0.0, 0.0);
                for (int k = i + 1; k <= j; ++k)
                    if (kept[k]) v = v + conjg(L[k * ld + i]) * Cs[k * ld + j];
                Cs[i * ld + j] = (-1.0 / diag[i]) * v;
            }
        }
    }
    __syncthreads();
    for (int e = t; e < s * s; e += kThreads) {
        const int i = e % s, j = e / s;
        C[j * kBlockMax + i] = Cs[i * ld + j];
        c128 v = mk(0.0, 0.0);
        if (kept[i] && j >= i) v = j == i ? mk(diag[i], 0.0) : conjg(L[j * ld + i]);
        T[j * kBlockMax + i] = v;
    }
    for (int j = t; j < s; j += kThreads) keep[j] = mk(kept[j] ? 1.0 : 0.0, 0.0);
}

// Out = A B for s x s matrices (column-major, leading dimension kBlockMax), one block.
__global__ void __launch_bounds__(kThreads) k_block_mm(const c128 *__restrict__ A, const c128 *__restrict__ B,
                                                       c128 *__restrict__ Out, int s) {
    for (int e = threadIdx.x; e < s * s; e += kThreads) {
        const int i = e % s, j = e / s;
        c128 v = mk(0.0, 0.0);
        for (int k = 0; k < s; ++k) v = v + A[k * kBlockMax + i] * B[j * kBlockMax + k];
        Out[j * kBlockMax + i] = v;
    }
}

// The squared residual norms of the block W S (column j: S_j^H G S_j with G = W^H W), written where the COCG
// finalize reads one block's partial sum: partial[j * kLdp]. One block, a thread per column.
__global__ void __launch_bounds__(kThreads) k_block_norms(const c128 *__restrict__ S, const c128 *__restrict__ G,
                                                          c128 *__restrict__ partial, int s) {
    const int j = threadIdx.x;
    if (j >= s) return;
    double v = 0.0;
    for (int i = 0; i < s; ++i) {
        c128 gs = mk(0.0, 0.0);
        for (int k = 0; k < s; ++k) gs = gs + G[k * kBlockMax + i] * S[j * kBlockMax + k];
        v += cmulc(S[j * kBlockMax + i], gs).x;
    }
    partial[static_cast<int64_t>(j) * kLdp] = mk(isnan(v) ? v : fmax(v, 0.0), 0.0);   // a NaN stays visible
}

// xbest_r = x_r for the systems whose monitored norm reached a new best at iteration it (grid (blocks, R)).
__global__ void __launch_bounds__(kThreads) k_block_keep_best(const c128 *__restrict__ x, c128 *__restrict__ xbest,
                                                              int64_t n, int it, const LoopState *__restrict__ st) {
    SYSTEM(r);
    if (st[r].best != it) return;
    const int64_t o = r * n;
    ROWS(i, n) xbest[o + i] = x[o + i];
}

// Y_j = [Y_j +] sign sum_i W_i C(i, j) for the columns j the mode selects (kBlockAll: every column; kBlockRunning:
// the systems not done; kBlockSearching: not done and not finished), a thread per row, C in LDS a quarter at a time.
enum BlockColumns { kBlockAll = 0, kBlockRunning = 1, kBlockSearching = 2 };
__global__ void __launch_bounds__(kThreads) k_block_nn(c128 *__restrict__ Y, const c128 *__restrict__ W,
                                                       const c128 *__restrict__ C, double sign, int64_t n, int s,
                                                       const LoopState *__restrict__ st, int mode, int accumulate) {
    __shared__ c128 cs[kBlockMax * kBlockMax / 4];   // 16 columns j of C
    const int t = threadIdx.x;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * kThreads + t;
    const int64_t rr = min(row, n - 1);
    for (int j0 = 0; j0 < s; j0 += 16) {
        __syncthreads();
        for (int e = t; e < 16 * kBlockMax; e += kThreads) {
            const int jj = e / kBlockMax, i = e % kBlockMax;
            cs[e] = j0 + jj < s ? C[(j0 + jj) * kBlockMax + i] : mk(0.0, 0.0);
        }
        __syncthreads();
        c128 acc[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = mk(0.0, 0.0);
        for (int i = 0; i < s; ++i) {
            const c128 w = W[static_cast<int64_t>(i) * n + rr];
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[q] = acc[q] + w * cs[q * kBlockMax + i];
        }
        if (row < n)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int j = j0 + q;
                if (j >= s) continue;
                if (mode != kBlockAll && (st[j].done || (mode == kBlockSearching && st[j].finished))) continue;
                c128 *y = Y + static_cast<int64_t>(j) * n + row;
                *y = accumulate ? *y + sign * acc[q] : sign * acc[q];
            }
    }
}

// ------------------------------------------------------------------------------------------------------------------
// GMRES kernels. Column k of the basis of system r is V + (k * R + r) * n (R = gridDim.y). The basis is held in
// complex128, or in complex64 for a complex64 inner solve (VT): its columns are the rounded Arnoldi vectors, every
// coefficient and update is computed in FP64 from them, and the operator is applied to the FP64 vector q.
// ------------------------------------------------------------------------------------------------------------------
template <typename VT>
__device__ __forceinline__ const VT *basis_col(const VT *V, int k, int r, int R, int64_t n) {
    return V + (static_cast<int64_t>(k) * R + r) * n;
}
__device__ __forceinline__ c128 widen(c128 v) { return v; }
__device__ __forceinline__ c128 widen(c64 v) { return mk(static_cast<double>(v.x), static_cast<double>(v.y)); }
template <typename VT>
__device__ __forceinline__ VT narrow(c128 v) {
    if constexpr (sizeof(VT) == sizeof(c128)) return v;
    else return mk(static_cast<float>(v.x), static_cast<float>(v.y));
}

// ------------------------------------------------------------------------------------------------------------------
// Arnoldi with delayed reorthogonalization (DCGS2: Swirydowicz, Langou, Ananthan, Yang and Thomas, Numer. Linear
// Algebra Appl. 2021; Bielich, Langou, Thomas, Swirydowicz, Yamazaki and Boman, Parallel Computing 2022). At step j
// the basis holds the final orthonormal columns q_0 .. q_{j-1}, and the FP64 vector p the once-orthogonalized,
// approximately normalized next one. One pass over the basis forms the inner products of p and of w = M A p,
//   s = V^H p, z = V^H w, alpha = p^H p, t = p^H w, omega = w^H w,
// from which (V = [q_0 .. q_{j-1}], H the Hessenberg matrix of the final columns)
//   beta = sqrt(alpha - |s|^2), q_j = (p - V s) / beta        the reorthogonalization of p, one step late,
//   z_j = (t - s^H z) / beta, h_j = (z - H s) / beta          column j of H (z extended by z_j),
//   nu = sqrt(omega - |z|^2) / beta, p' = (w - V z - z_j q_j) / (beta nu),
// and column j - 1 of H becomes final: + nu_{j-1} s above the diagonal, subdiagonal nu_{j-1} beta. A second pass
// writes q_j and p'. Two passes over the basis and one reduction per step, against the three passes and two
// reductions of CGS2, with the same orthogonality.
// ------------------------------------------------------------------------------------------------------------------
constexpr int kTileRows = 64;
constexpr int kTileStride = 65;
constexpr int kLdp2 = 128;                           // partial slots per block of the dot pass
constexpr int kPartZ = 64, kPartScalars = 125;       // s at k, z at kPartZ + k, then alpha, t, omega
constexpr size_t kDotsRedBytes = 2 * 4 * kTileRows * sizeof(c128);

// partial[(r * blocks + b) * kLdp2 + ...] = the inner products above over block b's rows: `per` tiles of 64 rows,
// tiles b * per .. b * per + per - 1, or, strided, b, b + blocks, b + 2 blocks ... (with the blocks a multiple of 8,
// every tile is then read on the XCD that runs block `tile` of a one-tile-per-block pass: the update pass finds the
// basis rows in its own L2). The live columns of a tile are staged in LDS with p and w (row stride 65 keeps the
// column-wise reads off shared banks); the columns' sums are split over row groups (below), wave 0 sums alpha, t and
// omega. Dynamic LDS: cnt * 65 basis elements.
template <typename VT>
__global__ void __launch_bounds__(kThreads) k_dcgs_dots(const VT *__restrict__ V, int cnt, const c128 *__restrict__ p,
                                                        const c128 *__restrict__ w, int64_t n, int per, int strided,
                                                        c128 *__restrict__ partial, const LoopState *__restrict__ st) {
    SYSTEM(r);
    const int R = gridDim.y;
    // Dynamic LDS: the tile [cnt][kTileStride], then the same bytes for the four waves' sums (kDotsRedBytes at least).
    extern __shared__ unsigned char tile_raw[];
    VT *tile = reinterpret_cast<VT *>(tile_raw);
    c128(*red)[4 * kTileRows] = reinterpret_cast<c128(*)[4 * kTileRows]>(tile_raw);
    __shared__ c128 pt[kTileRows], wt[kTileRows];
    // The wave index is uniform: the columns' base addresses live in scalar registers and the 16 loads of a thread
    // share one 32-bit row offset.
    const int lane = threadIdx.x & 63, q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const c128 *pr = p + r * n, *wr = w + r * n;
    const int64_t tiles = (n + kTileRows - 1) / kTileRows;
    // The dot phase: thread (k, g) sums column k over the rows i = g mod G, with G = 256 / cnt row groups, so that
    // every thread of the block works however few the columns.
    const int groups = cnt > 0 ? min(kTileRows, kThreads / cnt) : 1;
    const int dk = cnt > 0 ? static_cast<int>(threadIdx.x) % cnt : 0, dg = cnt > 0 ? static_cast<int>(threadIdx.x) / cnt : groups;
    double sre = 0, sim = 0, zre = 0, zim = 0;   // column dk, rows dg mod groups
    double al = 0, tre = 0, tim = 0, om = 0;     // wave 0: row `lane` of every tile
    constexpr int kPer = kLdp * kTileRows / kThreads;
    for (int tk = 0; tk < per; ++tk) {
        const int64_t tb = strided ? blockIdx.x + static_cast<int64_t>(tk) * gridDim.x
                                   : static_cast<int64_t>(blockIdx.x) * per + tk;
        if (tb >= tiles) break;
        const int64_t t0 = tb * kTileRows;
        const int rows = static_cast<int>(min(static_cast<int64_t>(kTileRows), n - t0));
        // All loads of the tile are issued before any is stored: the staging costs one memory latency.
        // Thread (q, lane) loads the columns k = q + 4 u (uniform over the wave) at the row `lane`.
        VT vals[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int k = q + 4 * u;
            vals[u] = (k < cnt && lane < rows) ? basis_col(V, k, r, R, n)[t0 + lane] : VT{0, 0};
        }
        c128 pv = mk(0.0, 0.0), wv = mk(0.0, 0.0);
        if (q == 0 && lane < rows) {
            pv = pr[t0 + lane];
            wv = wr[t0 + lane];
        }
        __syncthreads();   // the previous tile is consumed
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int k = q + 4 * u;
            if (k < cnt) tile[k * kTileStride + lane] = vals[u];
        }
        if (q == 0) {
            pt[lane] = pv;
            wt[lane] = wv;
            const c128 pw = cmulc(pv, wv);
            al += abs2(pv);
            tre += pw.x;
            tim += pw.y;
            om += abs2(wv);
        }
        __syncthreads();
        if (dg < groups) {
            const VT *col = tile + dk * kTileStride;
            for (int i = dg; i < rows; i += groups) {
                const c128 v = widen(col[i]), pi = pt[i], wi = wt[i];
                sre += v.x * pi.x + v.y * pi.y;
                sim += v.x * pi.y - v.y * pi.x;
                zre += v.x * wi.x + v.y * wi.y;
                zim += v.x * wi.y - v.y * wi.x;
            }
        }
    }
    __syncthreads();   // the tile is consumed: its bytes take the sums
    if (dg < groups) {
        red[0][dg * cnt + dk] = mk(sre, sim);
        red[1][dg * cnt + dk] = mk(zre, zim);
    }
    if (q == 0) {
        al = wave_sum(al);
        tre = wave_sum(tre);
        tim = wave_sum(tim);
        om = wave_sum(om);
    }
    __syncthreads();
    // Slot-major partials, [R][slot][block]: the reduction reads each slot's blocks contiguously.
    const int nb = gridDim.x;
    c128 *out = partial + static_cast<int64_t>(r) * kLdp2 * nb + blockIdx.x;
    if (q < 2 && lane < cnt) {   // the groups' sums in a fixed order
        const c128 *v = red[q];
        c128 sum = v[lane];
        for (int gi = 1; gi < groups; ++gi) sum = sum + v[gi * cnt + lane];
        out[static_cast<int64_t>(q * kPartZ + lane) * nb] = sum;
    }
    if (threadIdx.x == 0) {
        out[static_cast<int64_t>(kPartScalars) * nb] = mk(al, 0.0);
        out[static_cast<int64_t>(kPartScalars + 1) * nb] = mk(tre, tim);
        out[static_cast<int64_t>(kPartScalars + 2) * nb] = mk(om, 0.0);
    }
}

// One slot of the dot pass summed over its blocks by one wave: slot x gives s_x (x < cnt) into kH2 + x, z into kH1,
// then alpha, t and omega into kAlpha, kT, kOmega.
__device__ __forceinline__ void dcgs_slot_sum(const c128 *__restrict__ partial, int r, int blocks, int cnt, int x,
                                              int lane, c128 *__restrict__ scal) {
    int src, dst;
    if (x < cnt) {
        src = x;
        dst = kH2 + x;
    } else if (x < 2 * cnt) {
        src = kPartZ + x - cnt;
        dst = kH1 + x - cnt;
    } else {
        src = kPartScalars + x - 2 * cnt;
        dst = kAlpha + x - 2 * cnt;
    }
    // Up to 768 blocks in one round of loads (twelve per lane), the rest after. The loads past the last block read
    // the last one again and are dropped by a select: a load under a branch would be waited for inside it.
    constexpr int kRound = 12;
    double re[kRound], im[kRound];
    const c128 *row = partial + (static_cast<int64_t>(r) * kLdp2 + src) * blocks;
#pragma unroll
    for (int u = 0; u < kRound; ++u) {
        const int b = lane + 64 * u;
        const c128 v = row[min(b, blocks - 1)];
        re[u] = b < blocks ? v.x : 0.0;
        im[u] = b < blocks ? v.y : 0.0;
    }
    for (int b = lane + 64 * kRound; b < blocks; b += 64) {
        const c128 v = row[b];
        re[0] += v.x;
        im[0] += v.y;
    }
    static_assert(kRound == 12, "the pairwise sum below is written for twelve");
#pragma unroll
    for (int u = 0; u < 6; ++u) {
        re[u] += re[u + 6];
        im[u] += im[u + 6];
    }
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        re[u] += re[u + 3];
        im[u] += im[u + 3];
    }
    const double sr = wave_sum((re[0] + re[1]) + re[2]);
    const double si = wave_sum((im[0] + im[1]) + im[2]);
    if (lane == 0) scal[r * kScal + dst] = mk(sr, si);
}

// One step of a wave's inclusive scan of the affine maps x -> M x + B (two B at once, the later map applied after
// the earlier): every lane composes its map after the one the DPP move brings (the identity where none).
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ void affine_scan_step(double &M, c128 &bf, c128 &bn) {
    const double Mp = dpp_f64<CTRL, ROW_MASK>(1.0, M);
    const c128 fp = mk(dpp_f64<CTRL, ROW_MASK>(0.0, bf.x), dpp_f64<CTRL, ROW_MASK>(0.0, bf.y));
    const c128 np = mk(dpp_f64<CTRL, ROW_MASK>(0.0, bn.x), dpp_f64<CTRL, ROW_MASK>(0.0, bn.y));
    bf = M * fp + bf;
    bn = M * np + bn;
    M = M * Mp;
}


// Givens rotation of (a, sub) with sub real: c = a / radius, s = sub / radius; returns the rotated diagonal.
__device__ __forceinline__ c128 givens(c128 a, double sub, c128 &c, c128 &sv) {
    const double radius = sqrt(abs2(a) + sub * sub);
    c = mk(1.0, 0.0);
    sv = mk(0.0, 0.0);
    if (radius > 0) {
        const double inv = 1.0 / radius;
        c = inv * a;
        sv = mk(sub * inv, 0.0);
    }
    return cmulc(c, a) + cmulc(sv, mk(sub, 0.0));
}

// beta, z_j and nu of step j by one wave, lane k holding s_k and z_k (zero past j), from the dot pass's p^H p, p^H w
// and w^H w: beta = sqrt(b2), b2 = alpha - |s|^2, z_j = (t - s^H z) / beta, nu = sqrt(omega - |z|^2 - |z_j|^2) / beta.
// beta = 0 where b2 is not positive (breakdown, or p not finite); nu = 0 where w lies in the span.
__device__ __forceinline__ void step_coefficients(c128 sk, c128 zk, double alpha, double omega, c128 t, double &b2,
                                                  double &beta, c128 &zj, double &nu) {
    const c128 v = cmulc(sk, zk);
    const double ss = wave_sum(abs2(sk)), szr = wave_sum(v.x), szi = wave_sum(v.y), zz = wave_sum(abs2(zk));
    b2 = alpha - ss;
    beta = b2 > 0 && isfinite(b2) ? sqrt(b2) : 0.0;
    zj = beta > 0 ? (1.0 / beta) * (t - mk(szr, szi)) : mk(0.0, 0.0);
    const double n2 = omega - (zz + abs2(zj));
    nu = beta > 0 && n2 > 0 ? sqrt(n2) / beta : 0.0;
}

// Step j (0-based) of the cycle, iteration `it` overall, in two launches. First the block sums of the dot pass
// (k_dgmres_sums: a wave per slot, s, z, alpha, t, omega into scal), whose block that finishes last forms beta, z_j and
// nu and the update pass's coefficients 1 / beta and 1 / (beta nu). Then the update pass (k_dcgs_update), whose first
// block per system finalizes the step meanwhile (finalize_step): column j - 1 of H in its final form, its rotation
// formed again and the monitored residual of iteration it - 1 refined; column j with the subdiagonal nu (final one step
// later), its rotation, the rotated right-hand side g, the monitored residual |g(j+1)| / reference and the stopping
// test (iterate_.m:51-66). The serial finalize thus runs beside the update instead of before it. H holds the columns
// as formed (Hraw, for H s) and rotated (Hall, for the triangular solve); gall holds g, the cosines, the sines and g
// before each rotation. Grid (ceil((2 j + 3) / 4), R); LoopState.live = it marks the systems whose step runs.
__global__ void __launch_bounds__(kThreads) k_dgmres_sums(const c128 *__restrict__ partial, int blocks,
                                                          c128 *__restrict__ scal, int j, int it, LoopState *st) {
    SYSTEM(r);
    __shared__ int last;
    const int x = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (x < 2 * j + 3) dcgs_slot_sum(partial, r, blocks, j, x, threadIdx.x & 63, scal);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");   // the sums are visible to every XCD ...
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(&st[r].arrivals, 1) == static_cast<int>(gridDim.x) - 1;
    __syncthreads();
    if (!last || threadIdx.x >= 64) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // ... before the last block reads them
    c128 *sc = scal + r * kScal;
    const int lane = threadIdx.x, kk = min(lane, max(j - 1, 0));
    const c128 sk = sc[kH2 + kk], zk = sc[kH1 + kk];
    double b2, beta, nu;
    c128 zj;
    step_coefficients(pick(lane < j, sk, mk(0.0, 0.0)), pick(lane < j, zk, mk(0.0, 0.0)), sc[kAlpha].x,
                      sc[kOmega].x, sc[kT], b2, beta, zj, nu);
    if (lane != 0) return;
    sc[kB2] = mk(b2, 0.0);
    sc[kBeta] = mk(beta, 0.0);
    sc[kNu] = mk(nu, 0.0);
    sc[kZj] = zj;
    sc[kInvBeta] = mk(beta > 0 ? 1.0 / beta : 0.0, 0.0);   // 0: no update (no new column)
    sc[kInvBetaNu] = mk(beta > 0 && nu > 0 ? 1.0 / (beta * nu) : 0.0, 0.0);
    st[r].arrivals = 0;
    st[r].live = it;
}

// The finalize of step j for system r by one block (the update pass's first block of the system).
__device__ __forceinline__ void finalize_step(int r, c128 *__restrict__ scal, c128 *__restrict__ Hraw, c128 *__restrict__ Hall,
                              c128 *__restrict__ gall, int m, int j, int it, LoopState *st, double *history, int cap) {
    __shared__ c128 s[kLdp], z[kLdp + 1], part[4][kLdp + 1], colp[kLdp], colf[kLdp + 1], coln[kLdp + 1];
    __shared__ c128 csv[kLdp], snv[kLdp];
    c128 *sc = scal + r * kScal;
    const int ldh = m + 1;
    c128 *Hr = Hraw + static_cast<int64_t>(r) * ldh * m, *Hc = Hall + static_cast<int64_t>(r) * ldh * m;
    c128 *g = gall + static_cast<int64_t>(r) * 4 * ldh, *cs = g + ldh, *sn = g + 2 * ldh, *gs = g + 3 * ldh;
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    // Everything the step reads is loaded before any of it is stored, from addresses clamped into range (a load under
    // a branch would be waited for inside it): entry k = thread of s, z, the rotations and column j - 1 as formed
    // (j <= 62 entries), the scalars, and H(k, l) of the columns l < j - 1, l = q mod 4, for (H s)_k with k = lane, in
    // two halves of eight columns per lane (the second one, where j needs it, once the first is summed), so that the
    // block stays within the update pass's registers; beta, nu and z_j come from the dot pass's block sums.
    // s, z, the cosines, the sines and column j - 1 as formed: vector q by wave q and vector 4 by wave 0 too.
    const int kk = min(lane, max(j - 1, 0)), jm = max(j - 1, 0);
    auto vec_src = [&](int v) -> const c128 * {
        return v == 0 ? sc + kH2 : v == 1 ? sc + kH1 : v == 2 ? cs : v == 3 ? sn : Hr + static_cast<int64_t>(jm) * ldh;
    };
    auto vec_dst = [&](int v) { return v == 0 ? s : v == 1 ? z : v == 2 ? csv : v == 3 ? snv : colp; };
    const int v0 = __builtin_amdgcn_readfirstlane(q), v1 = q == 0 ? 4 : v0;   // v1 = v0: the same load again
    const c128 a0 = vec_src(v0)[kk], a1 = vec_src(v1)[kk];
    const int k0 = max(0, lane - 1), l0 = k0 + ((q - k0) & 3);
    constexpr int kHalf = kLdp / 8;   // columns per half: l = l0 + 4 u, u in [u0, u0 + kHalf)
    c128 hk[kHalf];
    auto load_h = [&](int u0) {
#pragma unroll
        for (int u = 0; u < kHalf; ++u) {
            const int l = l0 + 4 * (u0 + u);
            const c128 v = Hr[static_cast<int64_t>(min(l, max(j - 2, 0))) * ldh + min(lane, j)];
            hk[u] = pick(lane <= j && l < j - 1, v, mk(0.0, 0.0));
        }
    };
    load_h(0);
    const double hsub = Hr[static_cast<int64_t>(jm) * ldh + j].x;
    const double ref = sc[kReference].x, tol = sc[kTol].x, maxit = sc[kMaxIt].x;
    const double b2 = sc[kB2].x, beta = sc[kBeta].x, nu = sc[kNu].x;
    const c128 zj = sc[kZj], gprev = (j >= 1 ? gs + j - 1 : g)[0];
    if (lane < j) {
        vec_dst(v0)[lane] = a0;
        vec_dst(v1)[lane] = a1;
    }
    if (threadIdx.x == 0) z[j] = zj;
    __syncthreads();
    // (H s)_k over the columns l < j - 1 as formed, rows k <= j (H(k, l) = 0 below l + 1); wave q sums l = q mod 4 in
    // increasing order, lane k; column j - 1, final only below, is added last.
    c128 acc = mk(0.0, 0.0);
    auto add_h = [&](int u0) {
#pragma unroll
        for (int u = 0; u < kHalf; ++u) {
            const int l = l0 + 4 * (u0 + u);
            if (l < j - 1) acc = acc + hk[u] * s[l];
        }
    };
    if (lane <= j) add_h(0);
    const double nup = j >= 1 ? hsub : 0.0;
    // A finite p that vanishes after its reorthogonalization lies in the span of the basis: the subdiagonal of column
    // j - 1 is zero, so the residual of iteration it - 1 is zero and the system has converged there (a zero
    // subdiagonal in iterate_.m:62-66). Column j - 1 is finalized and rotated below; column j is never formed.
    const bool brk = j >= 1 && isfinite(b2) && !(b2 > 0);
    if (!(beta > 0) && !brk) {   // p is not finite: the system stops without a new column
        if (threadIdx.x == 0) st[r].finished = st[r].done = 1;
        return;
    }
    // Column j - 1 final: + nu_{j-1} s above the subdiagonal, subdiagonal nu_{j-1} beta (stored to H below, after the
    // last loads).
    if (j >= 1)
        for (int k = threadIdx.x; k <= j; k += kThreads) colf[k] = k < j ? colp[k] + nup * s[k] : mk(nup * beta, 0.0);
    const bool hs_on = lane <= j && !brk;
    if (j > 33) {   // uniform over the block: the second half's columns are l >= 32
        load_h(kHalf);
        if (hs_on) add_h(kHalf);
    }
    if (hs_on && l0 <= j - 1 && ((j - 1 - l0) & 3) == 0) {
        const int k = lane;
        acc = acc + (k < j ? colp[k] + nup * s[k] : mk(nup * beta, 0.0)) * s[j - 1];
    }
    if (hs_on) part[q][lane] = acc;
    __syncthreads();
    for (int k = threadIdx.x; k <= j; k += kThreads) {
        if (j >= 1) Hr[static_cast<int64_t>(j - 1) * ldh + k] = colf[k];
        if (brk) {
            coln[k] = mk(0.0, 0.0);
            continue;
        }
        const c128 hs = (part[0][k] + part[1][k]) + (part[2][k] + part[3][k]);
        coln[k] = (1.0 / beta) * (z[k] - hs);
        Hr[static_cast<int64_t>(j) * ldh + k] = coln[k];
    }
    __syncthreads();
    if (q != 0) return;
    // The rotations 0 .. j-2 on both columns. With run_0 = a_0, rotation k gives the final entry
    // conj(c_k) run_k + s_k a_{k+1} and the running entry run_{k+1} = -s_k run_k + c_k a_{k+1} (s_k real): an affine
    // recurrence, whose prefix compositions (M, B) (run_{k+1} = M a_0 + B) one wave scans, lane k holding step k.
    const int K = j >= 1 ? j - 1 : 0;
    c128 ck = mk(0.0, 0.0), af = mk(0.0, 0.0), an = mk(0.0, 0.0), bf = mk(0.0, 0.0), bn = mk(0.0, 0.0);
    double sk = 0, M = 1;
    if (lane < K) {
        ck = csv[lane];
        sk = snv[lane].x;
        af = colf[lane + 1];
        an = coln[lane + 1];
        M = -sk;
        bf = ck * af;
        bn = ck * an;
    }
    // The scan by DPP moves: row_shr 1, 2, 4, 8 within the rows of 16 lanes, then lane 15 of rows 0 and 2 into rows 1
    // and 3 (row_bcast:15) and lane 31 into rows 2 and 3 (row_bcast:31).
    affine_scan_step<0x111>(M, bf, bn);
    affine_scan_step<0x112>(M, bf, bn);
    affine_scan_step<0x114>(M, bf, bn);
    affine_scan_step<0x118>(M, bf, bn);
    affine_scan_step<0x142, 0xa>(M, bf, bn);
    affine_scan_step<0x143, 0xc>(M, bf, bn);
    // run_k on lane k from the composition of the steps before it (wave_shr:1; lane 0: the identity).
    const double Me = dpp_f64<0x138>(1.0, M);
    const c128 fe = mk(dpp_f64<0x138>(0.0, bf.x), dpp_f64<0x138>(0.0, bf.y));
    const c128 ne = mk(dpp_f64<0x138>(0.0, bn.x), dpp_f64<0x138>(0.0, bn.y));
    const c128 runf = Me * colf[0] + fe, runn = Me * coln[0] + ne;
    const c128 outf = cmulc(ck, runf) + sk * af, outn = cmulc(ck, runn) + sk * an;   // final entries k < K
    c128 cb = mk(readlane_f64(runn.x, K), readlane_f64(runn.y, K));
    c128 gj = gprev;   // the values thread 0 loaded
    c128 colnj1 = mk(0.0, 0.0);
    if (j >= 1) {
        const c128 ca = mk(readlane_f64(runf.x, K), readlane_f64(runf.y, K));
        c128 c, sv;
        const c128 diag = givens(ca, nup * beta, c, sv);   // rotation j - 1, formed again on the final column
        const c128 b1 = coln[j];
        colnj1 = cmulc(c, cb) + cmulc(sv, b1);
        cb = c * b1 - sv * cb;
        if (lane < j) Hc[static_cast<int64_t>(j - 1) * ldh + lane] = pick(lane < K, outf, diag);
        if (lane == 0) {
            cs[j - 1] = c;
            sn[j - 1] = sv;
            g[j - 1] = cmulc(c, gprev);
            gj = mk(0.0, 0.0) - sv * gprev;
            history[static_cast<int64_t>(r) * cap + it - 1] = sqrt(abs2(gj)) / ref;
        }
    }
    if (brk) {   // converged at iteration it - 1 with the j columns formed; no update pass for iteration it
        if (lane == 0) {
            st[r].monitored = 0.f;
            st[r].steps = j;
            st[r].iterations = it - 1;
            st[r].converged = st[r].finished = st[r].done = 1;
        }
        return;
    }
    c128 c, sv;
    const c128 diag = givens(cb, nu, c, sv);
    if (lane <= j) Hc[static_cast<int64_t>(j) * ldh + lane] = pick(lane < K, outn, pick(lane < j, colnj1, diag));
    if (lane != 0) return;
    sc[kH1 + j] = z[j];
    Hr[static_cast<int64_t>(j) * ldh + j + 1] = mk(nu, 0.0);
    cs[j] = c;
    sn[j] = sv;
    gs[j] = gj;
    const c128 gnext = mk(0.0, 0.0) - sv * gj;
    g[j + 1] = gnext;
    g[j] = cmulc(c, gj);
    const double res = sqrt(abs2(gnext)) / ref;
    history[static_cast<int64_t>(r) * cap + it] = res;
    st[r].monitored = static_cast<float>(res);
    st[r].steps = j + 1;
    st[r].iterations = it;
    const bool conv = res < tol;
    if (conv) {
        st[r].converged = st[r].finished = st[r].done = 1;
    } else if (!isfinite(res)) {
        st[r].finished = st[r].done = 1;
    } else if (it >= static_cast<int>(maxit)) {
        st[r].done = 1;
    }
}

// The update pass: q_j = (p - V s) / beta into column j of the basis, p = (w - V z - z_j q_j) / (beta nu). Runs for
// the systems whose step j runs (LoopState.live = it), the ones that stop in it included (q_j enters their solution),
// unless beta is not positive (no new column). The
// first block of each system (blockIdx.x = 0) finalizes the step instead (finalize_step), beside the update; the rows
// start at block kLeadBlocks, so that block b of the rows runs on XCD b mod 8 as the dot pass's tiles expect.
// S lanes share a row (S = 1 for batches, whose many rows fill the chip; S = 4 for a few systems): lane group s sums
// the columns k = s mod S, so that each lane's chain of column loads is S times shorter, and the S partial sums meet
// by lane shuffles in a fixed order.
constexpr int kLeadBlocks = 8;   // the update pass's blocks before its rows: the finalize and seven idle ones

template <typename VT, int S>
__global__ void __launch_bounds__(kThreads)
k_dcgs_update(VT *__restrict__ V, int cnt, c128 *__restrict__ p,
                                                          const c128 *__restrict__ w, int64_t n, int64_t chunk,
                                                          c128 *__restrict__ scal, int it, LoopState *st,
                                                          c128 *__restrict__ Hraw, c128 *__restrict__ Hall,
                                                          c128 *__restrict__ gall, int m, double *history, int cap) {
    static_assert(S == 1 || S == 2 || S == 4, "one, two or four lanes per row");
    const int r = blockIdx.y;
    if (st[r].live != it) return;
    if (blockIdx.x < kLeadBlocks) {   // dispatched first, so that the finalize starts with the update
        if (blockIdx.x == 0) finalize_step(r, scal, Hraw, Hall, gall, m, cnt, it, st, history, cap);
        return;
    }
    const int64_t blk = blockIdx.x - kLeadBlocks;
    const int R = gridDim.y;
    __shared__ c128 hs[kLdp], hz[kLdp];
    const c128 *sc = scal + r * kScal;
    VT *vj = V + (static_cast<int64_t>(cnt) * R + r) * n;
    c128 *pr = p + r * n;
    const c128 *wr = w + r * n;
    auto stage = [&]() {   // s and z of this step into LDS
        for (int k = threadIdx.x; k < cnt; k += kThreads) {
            hs[k] = sc[kH2 + k];
            hz[k] = sc[kH1 + k];
        }
        __syncthreads();
    };
    const double ib = sc[kInvBeta].x, isub = sc[kInvBetaNu].x;
    const c128 zj = sc[kZj];
    if (!(ib > 0)) return;   // no new column: p is not updated
    if constexpr (S == 1) {
        stage();
        const int64_t base = blk * chunk, end = min(n, base + chunk);
        for (int64_t i = base + threadIdx.x; i < end; i += kThreads) {
            c128 as = mk(0.0, 0.0), az = mk(0.0, 0.0);
            for (int k0 = 0; k0 < cnt; k0 += 8) {   // eight columns' loads in flight at once
                // Loads past the last column read it again and are not summed: a load under a branch would be
                // waited for inside it.
                VT vv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) vv[u] = basis_col(V, min(k0 + u, cnt - 1), r, R, n)[i];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (k0 + u < cnt) {
                        const c128 v = widen(vv[u]);
                        as = as + hs[k0 + u] * v;
                        az = az + hz[k0 + u] * v;
                    }
            }
            const c128 qj = ib * (pr[i] - as);
            vj[i] = narrow<VT>(qj);
            pr[i] = isub * (wr[i] - az - zj * qj);
        }
        return;
    }
    constexpr int kRowsPerWave = 64 / S;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sub = lane / kRowsPerWave;                                  // this lane's column group
    const int row = wave * kRowsPerWave + (lane & (kRowsPerWave - 1));   // row within the block's step
    const int64_t base = blk * chunk, end = min(n, base + chunk);
    const int last = max(cnt - 1, 0);   // column `cnt` is allocated: a valid address when cnt = 0
    for (int64_t i0 = base; i0 < end; i0 += kThreads / S) {
        const int64_t i = i0 + row;
        const bool in = i < end;
        // Loads past the last column or row read the last one again and are not summed or stored (a load under a
        // branch would be waited for inside it). The row's p and w and its first eight columns are requested before
        // s and z are staged, so that the step costs one memory latency.
        const int64_t ic = in ? i : end - 1;
        const c128 pv = pr[ic], wv = wr[ic];
        VT vv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) vv[u] = basis_col(V, min(sub + u * S, last), r, R, n)[ic];
        if (i0 == base) stage();
        c128 as = mk(0.0, 0.0), az = mk(0.0, 0.0);
        for (int k0 = sub;; ) {   // eight columns' loads in flight at once
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = k0 + u * S;
                if (k < cnt) {
                    const c128 v = widen(vv[u]);
                    as = as + hs[k] * v;
                    az = az + hz[k] * v;
                }
            }
            k0 += 8 * S;
            if (k0 >= cnt) break;
#pragma unroll
            for (int u = 0; u < 8; ++u) vv[u] = basis_col(V, min(k0 + u * S, last), r, R, n)[ic];
        }
        if constexpr (S > 1) {   // the column groups' sums, lane l gathering l ^ 32 (and l ^ 16)
#pragma unroll
            for (int o = 32; o >= kRowsPerWave; o >>= 1) {
                as.x += __shfl_xor(as.x, o, 64);
                as.y += __shfl_xor(as.y, o, 64);
                az.x += __shfl_xor(az.x, o, 64);
                az.y += __shfl_xor(az.y, o, 64);
            }
        }
        if (in && sub == 0) {
            const c128 qj = ib * (pv - as);
            vj[i] = narrow<VT>(qj);
            pr[i] = isub * (wv - az - zj * qj);
        }
    }
}

// Start of a cycle: p = w / |w|, g = (|w|, 0, ..., 0).
__global__ void k_gmres_start(const c128 *__restrict__ w, c128 *__restrict__ p, int m1, int64_t n,
                              const c128 *__restrict__ scal, c128 *__restrict__ gall, const LoopState *__restrict__ st) {
    SYSTEM(r);
    const double rn = scal[r * kScal + kRnorm].x;
    const c128 *wr = w + r * n;
    c128 *pr = p + r * n;
    ROWS(i, n) pr[i] = (1.0 / rn) * wr[i];
    if (blockIdx.x == 0) {
        c128 *g = gall + static_cast<int64_t>(r) * 4 * m1;
        for (int k = threadIdx.x; k < m1; k += blockDim.x) g[k] = mk(k == 0 ? rn : 0.0, 0.0);
    }
}

// Start of a cycle: the stop flag is the permanent one, the step count is zero.
__global__ void k_cycle_reset(LoopState *st, int R) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    st[r].done = st[r].finished;
    st[r].steps = 0;
}

// y = H(0:steps, 0:steps) \ g(0:steps) (upper triangular after the rotations). Grid (1, R); no-op at steps = 0.
__global__ __launch_bounds__(64) void k_gmres_solve(const c128 *__restrict__ Hall, const c128 *__restrict__ gall,
                                                    int m, const LoopState *st, c128 *__restrict__ scal) {
    // Column-oriented back substitution, one wave per right-hand side, lane l owning row l (steps <= m <= 62): the
    // rows a lane reads are staged in LDS first (all loads in flight at once), then each step finishes y_k on lane
    // k, broadcasts it and subtracts column k from the rows above. Every lane reads only the LDS it wrote itself.
    __shared__ c128 hs[62 * 64];
    const int r = blockIdx.y;
    const int lane = threadIdx.x;
    const int steps = st[r].steps;
    const int ldh = m + 1;
    const c128 *H = Hall + static_cast<int64_t>(r) * ldh * m;
    const c128 *g = gall + static_cast<int64_t>(r) * 4 * ldh;
    const bool row = lane < steps;
    for (int k = 0; k < steps; ++k)
        if (row && k >= lane) hs[k * 64 + lane] = H[lane + static_cast<int64_t>(ldh) * k];
    c128 acc = row ? g[lane] : mk(0.0, 0.0);
    c128 mine = mk(0.0, 0.0);
    for (int k = steps - 1; k >= 0; --k) {
        c128 yk = mk(0.0, 0.0);
        if (lane == k) {
            const c128 d = hs[k * 64 + k];
            yk = (1.0 / abs2(d)) * cmulc(d, acc);
            mine = yk;
        }
        yk = mk(readlane_f64(yk.x, k), readlane_f64(yk.y, k));   // k is uniform: a scalar read of lane k
        if (lane < k) acc = acc - hs[k * 64 + lane] * yk;
    }
    if (row) scal[r * kScal + kY + lane] = mine;
}

// x_r += V_r[:, 0:steps] y_r. Grid (blocks, R); no-op at steps = 0.
template <typename VT>
__global__ void k_gemv_add(const VT *__restrict__ V, const LoopState *st, const c128 *__restrict__ scal,
                           c128 *__restrict__ x, int64_t n) {
    __shared__ c128 ys[kLdp];
    const int r = blockIdx.y, R = gridDim.y;
    const int cnt = st[r].steps;
    if (cnt == 0) return;
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) ys[k] = scal[r * kScal + kY + k];
    __syncthreads();
    c128 *xr = x + r * n;
    ROWS(i, n) {
        c128 v = xr[i];
        for (int k = 0; k < cnt; ++k) v = v + ys[k] * widen(basis_col(V, k, r, R, n)[i]);
        xr[i] = v;
    }
}

// ------------------------------------------------------------------------------------------------------------------
// GMRES with restarts longer than kFastRestart (up to the iteration limit, restart = Inf): the reference's classical
// Gram-Schmidt run twice (orthogonalize_.m:15-23) on an FP64 basis V[k][r][n] of normalized columns. A step forms
// w = M A v_j, then twice h = V^H w (a dot pass over the j + 1 live columns and its reduction) and w -= V h, the norm
// of w as the subdiagonal, the Hessenberg column with all earlier rotations (one thread, in order), its own rotation
// and the stopping test, and v_{j+1} = w / norm. Slower per step than the delayed-reorthogonalization kernels (three
// passes over the basis and three reductions), but with no bound on the restart but memory.
// ------------------------------------------------------------------------------------------------------------------
constexpr int kCgsCols = 16;   // columns per wave of the dot pass (a block: 64 columns, blockIdx.z the column group)

// partial[(r * ldp + c) * gridDim.x + b] = sum over block b's rows of conj(V_c) w, c < cnt. Every wave runs over all
// rows of the block for its 16 columns; the sums are one wave's DPP reduction per column, in a fixed order.
__global__ void __launch_bounds__(kThreads) k_cgs_dots(const c128 *__restrict__ V, int cnt, const c128 *__restrict__ w,
                                                       int64_t n, int64_t chunk, c128 *__restrict__ partial, int ldp,
                                                       const LoopState *__restrict__ st) {
    SYSTEM(r);
    const int R = gridDim.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = blockIdx.z * 64 + wave * kCgsCols;
    if (c0 >= cnt) return;   // uniform over the wave; no barrier follows
    double re[kCgsCols], im[kCgsCols];
#pragma unroll
    for (int u = 0; u < kCgsCols; ++u) re[u] = im[u] = 0;
    const int64_t base = blockIdx.x * chunk, end = min(n, base + chunk);
    const c128 *wr = w + r * n;
    for (int64_t i = base + lane; i < end; i += 64) {
        const c128 wv = wr[i];
        c128 v[kCgsCols];
#pragma unroll
        for (int u = 0; u < kCgsCols; ++u) v[u] = basis_col(V, min(c0 + u, cnt - 1), r, R, n)[i];
#pragma unroll
        for (int u = 0; u < kCgsCols; ++u) {
            re[u] += v[u].x * wv.x + v[u].y * wv.y;
            im[u] += v[u].x * wv.y - v[u].y * wv.x;
        }
    }
#pragma unroll
    for (int u = 0; u < kCgsCols; ++u) {
        const double sr = wave_sum(re[u]), si = wave_sum(im[u]);
        if (lane == 0 && c0 + u < cnt)
            partial[(static_cast<int64_t>(r) * ldp + c0 + u) * gridDim.x + blockIdx.x] = mk(sr, si);
    }
}

// hc[r][c] = sum over the blocks of partial slot c (a wave per slot); the first pass stores it as column j of the
// Hessenberg matrix as formed (Hraw), the second adds it. Grid (ceil(cnt / 4), R).
__global__ void __launch_bounds__(kThreads) k_cgs_reduce(const c128 *__restrict__ partial, int blocks, int ldp, int cnt,
                                                         c128 *__restrict__ hc, c128 *__restrict__ Hraw, int m, int j,
                                                         int first, const LoopState *__restrict__ st) {
    SYSTEM(r);
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (c >= cnt) return;
    const c128 *row = partial + (static_cast<int64_t>(r) * ldp + c) * blocks;
    double re = 0, im = 0;
    for (int b = lane; b < blocks; b += 64) {
        re += row[b].x;
        im += row[b].y;
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane != 0) return;
    const c128 h = mk(re, im);
    hc[static_cast<int64_t>(r) * (m + 1) + c] = h;
    c128 *col = Hraw + (static_cast<int64_t>(r) * m + j) * (m + 1);
    col[c] = first ? h : col[c] + h;
}

// y[i] += sign * sum_k V_k[i] h[k] over k < cnt (cnt < 0: the system's steps), h staged in LDS 256 at a time.
// skip_done: return for stopped systems (the Gram-Schmidt updates; the solution update at a cycle's end runs for all).
__global__ void __launch_bounds__(kThreads) k_cgs_gemv(const c128 *__restrict__ V, int cnt, const c128 *__restrict__ h,
                                                       int ldh, double sign, c128 *__restrict__ y, int64_t n,
                                                       int skip_done, const LoopState *__restrict__ st) {
    const int r = blockIdx.y, R = gridDim.y;
    if (skip_done && st[r].done) return;
    if (cnt < 0) cnt = st[r].steps;
    if (cnt == 0) return;
    __shared__ c128 hs[kThreads];
    const c128 *hr = h + static_cast<int64_t>(r) * ldh;
    c128 *yr = y + r * n;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
    for (int64_t i0 = static_cast<int64_t>(blockIdx.x) * kThreads; i0 < n; i0 += stride) {
        const int64_t i = i0 + threadIdx.x;
        c128 acc = mk(0.0, 0.0);
        for (int k0 = 0; k0 < cnt; k0 += kThreads) {
            __syncthreads();
            if (k0 + static_cast<int>(threadIdx.x) < cnt) hs[threadIdx.x] = hr[k0 + threadIdx.x];
            __syncthreads();
            const int kn = min(kThreads, cnt - k0);
            if (i < n)
                for (int k = 0; k < kn; ++k) acc = acc + hs[k] * basis_col(V, k0 + k, r, R, n)[i];
        }
        if (i < n) yr[i] = yr[i] + sign * acc;
    }
}

// Step j of the cycle, iteration `it` overall: the Hessenberg column j as formed (Hraw, rows <= j) and the subdiagonal
// norm (scal kRnorm2 slot) rotated by the earlier rotations in order, its own rotation (generateRotation_.m:11-18), the
// rotated right-hand side g, the monitored residual |g(j+1)| / reference and the stopping test (iterate_.m:51-66);
// 1 / norm for the next basis vector. One thread per system. Grid (1, R).
constexpr int kSubSlot = 146;   // scal slot of the subdiagonal norm of the step
__global__ void k_cgs_finalize(const c128 *__restrict__ Hraw, c128 *__restrict__ Hall, c128 *__restrict__ gall,
                               c128 *__restrict__ scal, int m, int j, int it, LoopState *st, double *history,
                               int cap) {
    SYSTEM(r);
    if (threadIdx.x != 0) return;
    const int ldh = m + 1;
    c128 *sc = scal + r * kScal;
    const c128 *raw = Hraw + (static_cast<int64_t>(r) * m + j) * ldh;
    c128 *col = Hall + (static_cast<int64_t>(r) * m + j) * ldh;
    c128 *g = gall + static_cast<int64_t>(r) * 4 * ldh, *cs = g + ldh, *sn = g + 2 * ldh;
    const double sub = sc[kSubSlot].x;
    for (int k = 0; k <= j; ++k) col[k] = raw[k];
    for (int k = 0; k < j; ++k) {   // [conj(c) conj(s); -s c] on rows k, k + 1
        const c128 a = col[k], b = col[k + 1], c = cs[k], s = sn[k];
        col[k] = cmulc(c, a) + cmulc(s, b);
        col[k + 1] = c * b - s * a;
    }
    c128 c, sv;
    col[j] = givens(col[j], sub, c, sv);
    col[j + 1] = mk(0.0, 0.0);
    cs[j] = c;
    sn[j] = sv;
    const c128 gj = g[j];
    g[j + 1] = mk(0.0, 0.0) - sv * gj;
    g[j] = cmulc(c, gj);
    const double res = sqrt(abs2(g[j + 1])) / sc[kReference].x;
    history[static_cast<int64_t>(r) * cap + it] = res;
    st[r].monitored = static_cast<float>(res);
    st[r].steps = j + 1;
    st[r].iterations = it;
    const bool conv = res < sc[kTol].x;
    if (conv) {
        st[r].converged = st[r].finished = st[r].done = 1;
    } else if (!isfinite(res)) {
        st[r].finished = st[r].done = 1;
    } else if (it >= static_cast<int>(sc[kMaxIt].x)) {
        st[r].done = 1;
    }
    // A zero subdiagonal means the basis spans the solution: the residual above is zero and the system converged.
    sc[kInvSub] = (!conv && sub > 0) ? mk(1.0 / sub, 0.0) : mk(0.0, 0.0);
}

// v_{j+1} = w / norm for the systems still running after step j.
__global__ void k_cgs_next(const c128 *__restrict__ w, c128 *__restrict__ V, int j1, int64_t n,
                           const c128 *__restrict__ scal, const LoopState *__restrict__ st) {
    SYSTEM(r);
    const int R = gridDim.y;
    const double inv = scal[r * kScal + kInvSub].x;
    const c128 *wr = w + r * n;
    c128 *vr = V + (static_cast<int64_t>(j1) * R + r) * n;
    ROWS(i, n) vr[i] = inv * wr[i];
}

// y = H(0:steps, 0:steps) \ g(0:steps), any number of steps: back substitution, one block per system, the right-hand
// side in LDS (dynamic, steps entries), y into Y[r][m]. Grid (1, R).
__global__ void __launch_bounds__(kThreads) k_cgs_solve(const c128 *__restrict__ Hall, const c128 *__restrict__ gall,
                                                        int m, const LoopState *st, c128 *__restrict__ Y) {
    extern __shared__ unsigned char solve_raw[];
    c128 *acc = reinterpret_cast<c128 *>(solve_raw);
    __shared__ c128 yk_s;
    const int r = blockIdx.y;
    const int steps = st[r].steps;
    const int ldh = m + 1;
    const c128 *H = Hall + static_cast<int64_t>(r) * ldh * m;
    const c128 *g = gall + static_cast<int64_t>(r) * 4 * ldh;
    c128 *y = Y + static_cast<int64_t>(r) * m;
    for (int k = threadIdx.x; k < steps; k += kThreads) acc[k] = g[k];
    __syncthreads();
    for (int k = steps - 1; k >= 0; --k) {
        if (threadIdx.x == 0) {
            const c128 d = H[static_cast<int64_t>(k) * ldh + k];
            const c128 yk = (1.0 / abs2(d)) * cmulc(d, acc[k]);
            yk_s = yk;
            y[k] = yk;
        }
        __syncthreads();
        const c128 yk = yk_s;
        for (int i = threadIdx.x; i < k; i += kThreads) acc[i] = acc[i] - H[static_cast<int64_t>(k) * ldh + i] * yk;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Host side.
// ------------------------------------------------------------------------------------------------------------------
static unsigned vec_blocks(int64_t n) { return blocks_for(n, 256, 1024); }

// Raises a kernel's dynamic LDS limit (the static LDS of the kernel counts against the same 64 KiB default).
static void raise_lds(const void *kernel, size_t dynamic_bytes) {
    static std::mutex lock;
    static std::map<const void *, size_t> raised;
    std::lock_guard<std::mutex> g(lock);
    size_t &have = raised[kernel];
    if (dynamic_bytes <= have) return;
    RT_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(dynamic_bytes)));
    have = dynamic_bytes;
}

static c128 *vec(Context *ctx, int slot, int64_t count) {
    return static_cast<c128 *>(ctx->scratch(slot, sizeof(c128) * count, "the solver's vectors"));
}

Workspace::Workspace(Context *c, int64_t n_, int R_, int max_iter) : ctx(c), s(c->stream), n(n_), R(R_) {
    int cap_blocks = 16384 / R;
    cap_blocks = std::max(16, std::min(2048, cap_blocks));
    int64_t b = (n + kThreads - 1) / kThreads;
    b = std::max<int64_t>(1, std::min<int64_t>(b, cap_blocks));
    const int64_t per = (n + b - 1) / b;
    chunk = (per + kThreads - 1) / kThreads * kThreads;
    blocks = static_cast<int>((n + chunk - 1) / chunk);
    // The Gram-Schmidt passes run on 64-row tiles. With few systems (under 2048 reduction blocks in all) the update
    // pass takes one tile per block while the blocks stay within 16384, so that small systems are not left with a few
    // blocks walking many tiles one memory latency at a time; otherwise it runs on the reduction blocks' rows. The dot
    // pass takes at least three tiles per block: its blocks' partial sums are summed again on the critical path of
    // every step, and with one system a third of the blocks is faster overall (256 x 256 plate: 680 instead of 2040
    // blocks, 4 % more solves per second mixed, 7 % in FP64). Beside the one-tile update its tiles are strided over a
    // multiple of 8 blocks (each tile on the XCD of the update block that reads it next); beside the row-block update
    // it takes the same rows as the update's blocks (runs of at least four tiles).
    const int64_t tiles = (n + kTileRows - 1) / kTileRows;
    gs_tile_update = static_cast<int64_t>(blocks) * R < 2048;
    gs_blocks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(tiles, std::max(16, 16384 / R))));
    gs_chunk = (tiles + gs_blocks - 1) / gs_blocks * kTileRows;
    gs_blocks = static_cast<int>((n + gs_chunk - 1) / gs_chunk);
    if (gs_tile_update) {
        dot_strided = 1;
        dot_blocks = static_cast<int>(std::min<int64_t>(tiles, ((tiles + 2) / 3 + 7) / 8 * 8));
        dot_per = static_cast<int>((tiles + dot_blocks - 1) / dot_blocks);
    } else {
        dot_strided = 0;
        dot_per = static_cast<int>(chunk / kTileRows);   // chunk: a multiple of 256 rows
        dot_blocks = static_cast<int>((tiles + dot_per - 1) / dot_per);
    }
    partial = vec(c, kSlotPartial, static_cast<int64_t>(R) * std::max(blocks, dot_blocks) * kLdp2);
    scal = vec(c, kSlotScal, static_cast<int64_t>(R) * kScal);
    state = static_cast<LoopState *>(c->scratch(kSlotState, sizeof(LoopState) * R, "the solver's state"));
    cap = max_iter + 2;
    history = static_cast<double *>(c->scratch(kSlotHistory, sizeof(double) * cap * R, "the solver's history"));
    // Reads: up to three slots and the loop states of every system. Staging: room for many small writes between
    // two synchronizations, so that setting a slot or resetting the states does not wait for the device.
    read_bytes = std::max<size_t>(4096, static_cast<size_t>(R) * 128);
    host_bytes = read_bytes + 16 * std::max<size_t>(4096, static_cast<size_t>(R) * sizeof(c128));
    if (!c->host_pinned || c->host_pinned_bytes < host_bytes) {
        if (c->host_pinned) RT_HIP(hipHostFree(c->host_pinned));
        c->host_pinned = nullptr;
        RT_HIP(hipHostMalloc(&c->host_pinned, host_bytes, hipHostMallocDefault));
        c->host_pinned_bytes = host_bytes;
    }
    host = static_cast<unsigned char *>(c->host_pinned);
}

void Workspace::dot(const c128 *x, const c128 *y, bool conj, int slot, const LoopState *st) {
    if (conj)
        hipLaunchKernelGGL(k_dot<true>, dim3(blocks, R), dim3(kThreads), 0, s, x, y, n, chunk, partial, st);
    else
        hipLaunchKernelGGL(k_dot<false>, dim3(blocks, R), dim3(kThreads), 0, s, x, y, n, chunk, partial, st);
    RT_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce, dim3(1, R), dim3(kThreads), 0, s, partial, blocks, scal, slot, 0, st);
    RT_LAUNCH_CHECK();
}

void Workspace::nrm2(const c128 *x, int slot, const LoopState *st) {
    hipLaunchKernelGGL(k_nrm2, dim3(blocks, R), dim3(kThreads), 0, s, x, n, chunk, partial, st);
    RT_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce, dim3(1, R), dim3(kThreads), 0, s, partial, blocks, scal, slot, 1, st);
    RT_LAUNCH_CHECK();
}

std::vector<double> Workspace::fetch_re(int slot) {
    return fetch_slots({slot}, nullptr)[0];
}

std::vector<std::vector<double>> Workspace::fetch_slots(const std::vector<int> &slots, std::vector<LoopState> *states) {
    const size_t per = sizeof(c128) * R;
    for (size_t k = 0; k < slots.size(); ++k)
        RT_HIP(hipMemcpy2DAsync(host + k * per, sizeof(c128), scal + slots[k], sizeof(c128) * kScal, sizeof(c128), R,
                                hipMemcpyDeviceToHost, s));
    const size_t at = slots.size() * per;
    if (states) RT_HIP(hipMemcpyAsync(host + at, state, sizeof(LoopState) * R, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    stage_off = 0;   // every staged write before the synchronization has been copied
    std::vector<std::vector<double>> out(slots.size(), std::vector<double>(R));
    for (size_t k = 0; k < slots.size(); ++k) {
        const c128 *h = reinterpret_cast<const c128 *>(host + k * per);
        for (int r = 0; r < R; ++r) out[k][r] = h[r].x;
    }
    if (states) {
        const LoopState *h = reinterpret_cast<const LoopState *>(host + at);
        states->assign(h, h + R);
    }
    return out;
}

unsigned char *Workspace::stage(size_t bytes) {
    bytes = (bytes + 255) / 256 * 256;
    if (stage_off + bytes > host_bytes - read_bytes) {   // full: wait for the copies in flight to finish
        RT_HIP(hipStreamSynchronize(s));
        stage_off = 0;
    }
    unsigned char *out = host + read_bytes + stage_off;
    stage_off += bytes;
    return out;
}

void Workspace::put_re(int slot, const std::vector<double> &v) {
    c128 *h = reinterpret_cast<c128 *>(stage(sizeof(c128) * R));
    for (int r = 0; r < R; ++r) h[r] = mk(v[r], 0.0);
    RT_HIP(hipMemcpy2DAsync(scal + slot, sizeof(c128) * kScal, h, sizeof(c128), sizeof(c128), R,
                            hipMemcpyHostToDevice, s));
}

void Workspace::reset_state(const std::vector<char> &finished) {
    int *flags = static_cast<int *>(ctx->scratch(kSlotFlags, sizeof(int) * R, "the solver's state"));
    int *h = reinterpret_cast<int *>(stage(sizeof(int) * R));
    for (int r = 0; r < R; ++r) h[r] = finished[r] ? 1 : 0;
    RT_HIP(hipMemcpyAsync(flags, h, sizeof(int) * R, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_reset_state, dim3((R + 63) / 64), dim3(64), 0, s, state, flags, R);
    RT_LAUNCH_CHECK();
}

void Workspace::set_scale(const std::vector<LoopState> &st) {
    int active = 0;
    for (const auto &q : st) active += q.done ? 0 : 1;
    ctx->timed_scale = static_cast<double>(active) / R;
}

// The timed launches enqueued since `mark` (pending records of this loop) that belong to a step s get their exact
// algorithmic bytes: the kernels of step s served the systems whose iteration count reached s, which the loop states
// read after the step tell. Returns the new mark.
static size_t settle_timed(Context *ctx, const std::vector<LoopState> &state, size_t mark) {
    if (!ctx->timing) return mark;
    if (ctx->pending.size() < mark) mark = 0;   // drained meanwhile: what is pending now is this loop's
    const int R = static_cast<int>(state.size());
    for (size_t k = mark; k < ctx->pending.size(); ++k) {
        Context::TimedRecord &rec = ctx->pending[k];
        if (rec.step < 0) continue;
        int active = 0;
        for (const auto &q : state) active += q.iterations >= rec.step ? 1 : 0;
        rec.bytes = rec.full * active / R;
    }
    return ctx->pending.size();
}

std::vector<LoopState> Workspace::read_state() {
    RT_HIP(hipMemcpyAsync(host, state, sizeof(LoopState) * R, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    stage_off = 0;
    const LoopState *h = reinterpret_cast<const LoopState *>(host);
    return std::vector<LoopState>(h, h + R);
}

std::vector<double> Workspace::read_history(int r, int first, int count) {
    std::vector<double> out(count > 0 ? count : 0);
    if (count > 0)
        RT_HIP(hipMemcpy(out.data(), history + static_cast<int64_t>(r) * cap + first, count * sizeof(double),
                         hipMemcpyDeviceToHost));
    return out;
}

void Workspace::copy(const c128 *x, c128 *y, const LoopState *st) {
    hipLaunchKernelGGL(k_copy, dim3(vec_blocks(n), R), dim3(256), 0, s, x, y, n, st);
    RT_LAUNCH_CHECK();
}
void Workspace::sub(const c128 *b, const c128 *ax, c128 *r, const LoopState *st) {
    hipLaunchKernelGGL(k_sub, dim3(vec_blocks(n), R), dim3(256), 0, s, b, ax, r, n, st);
    RT_LAUNCH_CHECK();
}
void Workspace::add(const c128 *x, c128 *y, const LoopState *st) {
    hipLaunchKernelGGL(k_add, dim3(vec_blocks(n), R), dim3(256), 0, s, x, y, n, st);
    RT_LAUNCH_CHECK();
}
void Workspace::zero(c128 *x) { RT_HIP(hipMemsetAsync(x, 0, static_cast<size_t>(R) * n * sizeof(c128), s)); }

struct InnerResult {
    std::vector<char> converged;
    std::vector<int> iterations;
    std::vector<std::vector<double>> norms;   // monitored norms from iteration 0
};

// GMRES: the steps to enqueue before the next poll of the loop states. A poll drains the stream (about as long as
// a step), and steps enqueued past a system's convergence run as no-ops (as long again): the window is the
// iterations the slowest running system still needs at the rate its monitored residual fell since the previous
// poll (log-linear), between 1 and kPollMax; kPoll for a system without such a rate, and the previous estimate for
// one that made no step since (a poll at the end of a cycle).
struct PollPlan {
    std::vector<int> it0, est;
    std::vector<float> res0;
    int next(const std::vector<LoopState> &st, const std::vector<double> &tol) {
        const size_t R = st.size();
        if (it0.size() != R) {
            it0.assign(R, -1);
            est.assign(R, kPoll);
            res0.assign(R, 0.f);
        }
        int need = 1;
        for (size_t r = 0; r < R; ++r) {
            const LoopState &q = st[r];
            if (q.done) continue;
            const double res = q.monitored, prev = res0[r];
            if (q.iterations == it0[r]) {
                need = std::max(need, est[r]);
                continue;
            }
            int k = kPoll;
            if (it0[r] >= 0 && q.iterations > it0[r] && res > 0 && res < prev && tol[r] > 0) {
                const double slope = std::log(res / prev) / (q.iterations - it0[r]);
                const double left = std::ceil(std::log(tol[r] / res) / slope);
                k = static_cast<int>(std::min<double>(kPollMax, std::max(1.0, left)));
            }
            it0[r] = q.iterations;
            res0[r] = q.monitored;
            est[r] = k;
            need = std::max(need, k);
        }
        return std::min(kPollMax, need);
    }
};

static bool all_done(const std::vector<LoopState> &st) {
    for (const auto &x : st)
        if (!x.done) return false;
    return true;
}

static InnerResult collect(Workspace &ws, const std::vector<LoopState> &state, const std::vector<char> &skip_sys,
                           const std::vector<double> &num, const std::vector<double> &den) {
    const int R = ws.R;
    InnerResult res;
    res.converged.assign(R, 0);
    res.iterations.assign(R, 0);
    res.norms.resize(R);
    for (int q = 0; q < R; ++q) {
        if (skip_sys[q]) continue;
        res.converged[q] = state[q].converged;
        res.iterations[q] = state[q].iterations;
        res.norms[q].push_back(num[q] / den[q]);
        std::vector<double> h = ws.read_history(q, 1, state[q].iterations);
        res.norms[q].insert(res.norms[q].end(), h.begin(), h.end());
    }
    return res;
}

// ------------------------------------------------------------------------------------------------------------------
// COCG (ConjugateGradient.m:32-59) on the systems not marked `skip_sys`. Tolerances per system.
// ------------------------------------------------------------------------------------------------------------------
// Iteration limits per system; returns the largest over the systems that run.
static int put_limits(Workspace &ws, const std::vector<int> &limits, const std::vector<char> &skip_sys) {
    std::vector<double> v(ws.R);
    int top = 0;
    for (int r = 0; r < ws.R; ++r) {
        v[r] = limits[r];
        if (!skip_sys[r]) top = std::max(top, limits[r]);
    }
    ws.put_re(kMaxIt, v);
    return top;
}

static InnerResult cocg(Workspace &ws, const LinOp &A, const LinOp *P, const c128 *b, c128 *x, bool has_x0,
                        const std::vector<double> &tol, const std::vector<char> &skip_sys,
                        const std::vector<int> &limits, int stall) {
    Context *ctx = ws.ctx;
    const int64_t n = ws.n;
    const int R = ws.R;
    const LoopState *st = ws.state;
    c128 *r = vec(ctx, kSlotR, R * n), *d = vec(ctx, kSlotD, R * n), *ad = vec(ctx, kSlotAd, R * n);
    c128 *z = P ? vec(ctx, kSlotZ, R * n) : nullptr;
    ws.reset_state(skip_sys);
    ws.put_re(kTol, tol);
    ws.put_re(kBest, std::vector<double>(R, HUGE_VAL));
    const int max_iter = put_limits(ws, limits, skip_sys);
    if (has_x0) {
        A(x, ad, ws.skip());
        ws.sub(b, ad, r, st);
    } else {
        ws.zero(x);
        ws.copy(b, r, st);
    }
    ws.nrm2(r, kRnormCg, st);
    ws.nrm2(b, kBnorm, st);
    hipLaunchKernelGGL(k_initial_check, dim3((R + 63) / 64), dim3(64), 0, ws.s, ws.state, ws.scal, kRnormCg, kBnorm, R);
    RT_LAUNCH_CHECK();
    std::vector<LoopState> state;
    const auto first = ws.fetch_slots({kRnormCg, kBnorm}, &state);
    const std::vector<double> &r0 = first[0], &b0 = first[1];
    size_t timed_mark = ctx->pending.size();
    int it = 0;
    while (it < max_iter && !all_done(state)) {
        const int window = std::min(kPoll, max_iter - it);
        for (int w = 0; w < window; ++w) {
            ++it;
            ctx->timed_step = it;
            const c128 *zz = r;
            if (P) {
                (*P)(r, z, ws.skip());
                zz = z;
            }
            const int i_new = (it & 1) ? kRho1 : kRho0, i_old = (it & 1) ? kRho0 : kRho1;
            ws.dot(r, zz, false, i_new, st);
            if (it == 1) {
                ws.copy(zz, d, st);
            } else {
                hipLaunchKernelGGL(k_cocg_direction, dim3(vec_blocks(n), R), dim3(256), 0, ws.s, zz, d, ws.scal, i_new,
                                   i_old, n, st);
                RT_LAUNCH_CHECK();
            }
            A(d, ad, ws.skip());
            ws.dot(d, ad, false, kDad, st);
            hipLaunchKernelGGL(k_cocg_update, dim3(ws.blocks, R), dim3(kThreads), 0, ws.s, x, r, d, ad, ws.scal,
                               i_new, n, ws.chunk, ws.partial, st);
            RT_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_cocg_finalize, dim3(1, R), dim3(kThreads), 0, ws.s, ws.partial, ws.blocks, ws.scal,
                               it, stall, ws.state, ws.history, ws.cap, 0);
            RT_LAUNCH_CHECK();
        }
        ctx->timed_step = -1;
        state = ws.read_state();
        timed_mark = settle_timed(ctx, state, timed_mark);
        ws.set_scale(state);
    }
    ctx->timed_scale = 1.0;
    return collect(ws, state, skip_sys, r0, b0);
}

// ------------------------------------------------------------------------------------------------------------------
// Block COCG: the R right-hand sides of one complex symmetric system (A^T = A, a symmetric preconditioner M) share
// one block Krylov space: block conjugate gradients with the unconjugated products of ConjugateGradient.m:32-59. The
// residual block is carried as an orthonormal basis and its coordinates, R = Q S (the residual orthonormalization of
// Dubrulle's retooled block CG), and the direction block P is an orthonormal basis of Z + P beta (the breakdown-free
// block CG of Ji and Li, 2017):
//   A P, D = P^T A P, a = D^-1 P^T Q, X += P a S, W = Q - A P a, W = Q' T (rank-revealing), S' = T S,
//   Z = M Q', beta = -D^-1 (A P)^T Z, P' = orth(Z + P beta).
// A block Krylov iteration depends on its blocks only through their spans, so these are the iterates of block COCG,
// with every coefficient matrix formed from orthonormal blocks: residuals that converge at different rates, or right-
// hand sides that are (nearly) dependent, drop directions instead of making a coefficient matrix singular. The
// monitored norm of system j is norm(R_j) / norm(b_j), norm(R_j)^2 = S_j^H W^H W S_j. A system that falls below its
// tolerance, or reaches its iteration limit, stops there (its solution and record frozen). As an inner solve of the
// refinement (stall > 0) a system also stops when its norm has not improved for `stall` iterations or exceeds
// kBlockGrowth times its best, and hands back the iterate of its best norm: the outer loop then restarts a fresh block
// recurrence from the true residual.
// ------------------------------------------------------------------------------------------------------------------
constexpr double kBlockDrop = 1e-6;   // rank-revealing threshold of the residual and direction blocks

static bool all_finished(const std::vector<LoopState> &st) {
    for (const auto &x : st)
        if (!x.finished && !x.done) return false;
    return true;
}

static InnerResult block_cocg(Workspace &ws, const LinOp &A, const LinOp *P, const c128 *b, c128 *x, bool has_x0,
                              const std::vector<double> &tol, const std::vector<char> &skip_sys,
                              const std::vector<int> &limits, int stall) {
    Context *ctx = ws.ctx;
    const int64_t n = ws.n;
    const int R = ws.R;
    if (R > kBlockMax)
        fail(RT_EINVAL, "rooftopMoM:solver:invalidRightHandSide",
             "Block COCG solves at most %d right-hand sides together.", kBlockMax);
    LoopState *st = ws.state;
    // The columns of Q, Z, P and A P are basis vectors, not systems: every product runs on all of them.
    c128 *qb = vec(ctx, kSlotR, R * n), *zb = vec(ctx, kSlotZ, R * n), *pd = vec(ctx, kSlotD, R * n);
    c128 *ap = vec(ctx, kSlotAd, R * n);
    constexpr int kSq = kBlockMax * kBlockMax;
    c128 *coef = vec(ctx, kSlotBlockCoef, 14 * kSq);
    c128 *D = coef, *F = coef + kSq, *a = coef + 2 * kSq, *ax = coef + 3 * kSq, *E = coef + 4 * kSq;
    c128 *beta = coef + 5 * kSq, *G = coef + 6 * kSq, *Cm = coef + 7 * kSq, *Tm = coef + 8 * kSq;
    c128 *S = coef + 9 * kSq, *S2 = coef + 10 * kSq, *keep_p = coef + 11 * kSq, *keep_r = coef + 12 * kSq;
    c128 *Tp = coef + 13 * kSq;
    const int tn_blocks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(512, (n + kBtRows - 1) / kBtRows)));
    const int64_t tn_chunk = ((n + tn_blocks - 1) / tn_blocks + kBtRows - 1) / kBtRows * kBtRows;
    c128 *tn_partial = vec(ctx, kSlotBlockPartial, static_cast<int64_t>(tn_blocks) * kSq);
    raise_lds(reinterpret_cast<const void *>(&k_block_solve), kBlockLds);
    raise_lds(reinterpret_cast<const void *>(&k_block_orth), kBlockLds);
    auto transposed = [&](const c128 *U, const c128 *V, c128 *Out, bool conj) {   // Out = U^T V (conj: U^H V)
        if (conj)
            hipLaunchKernelGGL(k_block_tn<true>, dim3(tn_blocks), dim3(kThreads), 0, ws.s, U, V, n, R, tn_chunk,
                               tn_partial);
        else
            hipLaunchKernelGGL(k_block_tn<false>, dim3(tn_blocks), dim3(kThreads), 0, ws.s, U, V, n, R, tn_chunk,
                               tn_partial);
        RT_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_block_sum, dim3(kSq / kThreads), dim3(kThreads), 0, ws.s, tn_partial, tn_blocks, Out);
        RT_LAUNCH_CHECK();
    };
    auto solve = [&](const c128 *M, const c128 *B, c128 *X, int at) {   // X = M^-1 B on the kept directions
        hipLaunchKernelGGL(k_block_solve, dim3(1), dim3(kThreads), static_cast<uint32_t>(kBlockLds), ws.s, M, B, X,
                           keep_p, R, st, at, ws.history, ws.cap);
        RT_LAUNCH_CHECK();
    };
    auto combine = [&](c128 *Y, const c128 *W, const c128 *Cc, double sign, int mode, int accumulate) {
        hipLaunchKernelGGL(k_block_nn, dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                           ws.s, Y, W, Cc, sign, n, R, st, mode, accumulate);
        RT_LAUNCH_CHECK();
    };
    auto mm = [&](const c128 *L, const c128 *Rm, c128 *Out) {
        hipLaunchKernelGGL(k_block_mm, dim3(1), dim3(kThreads), 0, ws.s, L, Rm, Out, R);
        RT_LAUNCH_CHECK();
    };
    // Gram matrix of the block in `from`, its rank-revealing factors (C, T, keep), and `to` = from C.
    auto orthonormalize = [&](c128 *from, c128 *to, c128 *T_out, c128 *keep, int stop_empty) {
        transposed(from, from, G, true);
        hipLaunchKernelGGL(k_block_orth, dim3(1), dim3(kThreads), static_cast<uint32_t>(kBlockLds), ws.s, G, Cm,
                           T_out, keep, R, kBlockDrop, stop_empty, st, ws.history, ws.cap);
        RT_LAUNCH_CHECK();
        combine(to, from, Cm, 1.0, kBlockAll, 0);
    };
    auto precondition = [&](const c128 *in, c128 *out) {
        if (P) (*P)(in, out, nullptr);
        else ws.copy(in, out, nullptr);
    };
    ws.reset_state(skip_sys);
    ws.put_re(kTol, tol);
    ws.put_re(kBest, std::vector<double>(R, HUGE_VAL));
    const int max_iter = put_limits(ws, limits, skip_sys);
    // Systems left out (and those converged at the start) are zero columns of the first residual block; the vector
    // slots are shared with other loops, so every block starts cleared.
    for (c128 *v : {qb, zb, pd, ap}) RT_HIP(hipMemsetAsync(v, 0, sizeof(c128) * R * n, ws.s));
    if (has_x0) {
        A(x, ap, nullptr);
        ws.sub(b, ap, qb, st);
    } else {
        ws.zero(x);
        ws.copy(b, qb, st);
    }
    ws.nrm2(qb, kRnormCg, st);
    ws.nrm2(b, kBnorm, st);
    hipLaunchKernelGGL(k_initial_check, dim3((R + 63) / 64), dim3(64), 0, ws.s, ws.state, ws.scal, kRnormCg, kBnorm, R);
    RT_LAUNCH_CHECK();
    std::vector<LoopState> state;
    const auto first = ws.fetch_slots({kRnormCg, kBnorm}, &state);
    const std::vector<double> &r0 = first[0], &b0 = first[1];
    for (int j = 0; j < R; ++j)
        if (state[j].done) RT_HIP(hipMemsetAsync(qb + j * n, 0, sizeof(c128) * n, ws.s));
    // The best iterate so far (an inner solve of the refinement): the starting one, at its relative norm.
    c128 *xbest = stall > 0 ? vec(ctx, kSlotBlockBest, R * n) : nullptr;
    if (xbest) {
        ws.copy(x, xbest, nullptr);
        std::vector<double> best(R, HUGE_VAL);
        for (int j = 0; j < R; ++j)
            if (b0[j] > 0) best[j] = r0[j] / b0[j];
        ws.put_re(kBest, best);
    }
    orthonormalize(qb, zb, S, keep_r, 0);   // R0 = Q S
    std::swap(qb, zb);
    precondition(qb, zb);
    orthonormalize(zb, pd, Tp, keep_p, 1);
    int it = 0;
    state = ws.read_state();
    while (it < max_iter && !all_finished(state)) {
        const int window = std::min(kPoll, max_iter - it);
        for (int w = 0; w < window; ++w) {
            ++it;
            A(pd, ap, nullptr);
            transposed(pd, ap, D, false);
            transposed(pd, qb, F, false);
            solve(D, F, a, it);
            mm(a, S, ax);
            combine(x, pd, ax, 1.0, kBlockSearching, 1);
            combine(qb, ap, a, -1.0, kBlockAll, 1);   // W = Q - A P a
            transposed(qb, qb, G, true);
            hipLaunchKernelGGL(k_block_norms, dim3(1), dim3(kThreads), 0, ws.s, S, G, ws.partial, R);
            RT_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_cocg_finalize, dim3(1, R), dim3(kThreads), 0, ws.s, ws.partial, 1, ws.scal, it,
                               stall, ws.state, ws.history, ws.cap, 1);
            RT_LAUNCH_CHECK();
            if (xbest) {
                hipLaunchKernelGGL(k_block_keep_best, dim3(ws.blocks, R), dim3(kThreads), 0, ws.s, x, xbest, n, it, st);
                RT_LAUNCH_CHECK();
            }
            hipLaunchKernelGGL(k_block_orth, dim3(1), dim3(kThreads), static_cast<uint32_t>(kBlockLds), ws.s, G, Cm,
                               Tm, keep_r, R, kBlockDrop, 0, st, ws.history, ws.cap);
            RT_LAUNCH_CHECK();
            combine(zb, qb, Cm, 1.0, kBlockAll, 0);   // Q' = W C
            std::swap(qb, zb);
            mm(Tm, S, S2);   // S' = T S
            std::swap(S, S2);
            precondition(qb, zb);
            transposed(ap, zb, E, false);
            solve(D, E, beta, -1);
            combine(zb, pd, beta, -1.0, kBlockAll, 1);   // Z + P beta, beta = -D^-1 (A P)^T Z
            orthonormalize(zb, pd, Tp, keep_p, 1);
        }
        state = ws.read_state();
        ws.set_scale(state);
    }
    if (xbest) ws.copy(xbest, x, nullptr);
    ctx->timed_scale = 1.0;
    return collect(ws, state, skip_sys, r0, b0);
}

// ------------------------------------------------------------------------------------------------------------------
// GMRES(m) with m > kFastRestart, left preconditioned, CGS2 exactly as the reference orthogonalizes
// (iterate_.m:1-85, orthogonalize_.m:15-23), on an FP64 basis of m + 1 columns.
// ------------------------------------------------------------------------------------------------------------------
constexpr int kMaxRestart = 8192;   // the right-hand side of the cycle's triangular solve stays in one block's LDS

static InnerResult gmres_cgs(Workspace &ws, const LinOp &A, const LinOp *P, const c128 *b, c128 *x, bool has_x0,
                             const std::vector<double> &tol, const std::vector<char> &skip_sys, int max_iter, int m) {
    if (m > kMaxRestart)
        fail(RT_EINVAL, "rooftopMoM:solver:invalidRestart",
             "A GMRES cycle of %d steps exceeds the %d this GPU solver supports; set a smaller restart.", m,
             kMaxRestart);
    Context *ctx = ws.ctx;
    const int64_t n = ws.n;
    const int R = ws.R;
    const LoopState *st = ws.state;
    const int m1 = m + 1;
    c128 *V = vec(ctx, kSlotV, static_cast<int64_t>(R) * m1 * n);
    c128 *w = vec(ctx, kSlotW, R * n), *t = vec(ctx, kSlotT, R * n);
    c128 *H = vec(ctx, kSlotH, static_cast<int64_t>(R) * m1 * m);
    c128 *Hr = vec(ctx, kSlotHraw, static_cast<int64_t>(R) * m1 * m);
    c128 *g = vec(ctx, kSlotG, static_cast<int64_t>(R) * 4 * m1);
    c128 *hc = vec(ctx, kSlotHc, static_cast<int64_t>(R) * m1);
    c128 *Y = vec(ctx, kSlotY, static_cast<int64_t>(R) * m);
    c128 *partial = vec(ctx, kSlotPartialL, static_cast<int64_t>(R) * m1 * ws.blocks);
    const size_t solve_lds = static_cast<size_t>(m) * sizeof(c128);
    if (solve_lds > 48 * 1024) raise_lds(reinterpret_cast<const void *>(&k_cgs_solve), solve_lds);
    auto precondition = [&](const c128 *in, c128 *out) {
        if (P) (*P)(in, out, ws.skip());
        else ws.copy(in, out, st);
    };
    ws.reset_state(skip_sys);
    ws.put_re(kTol, tol);
    precondition(b, w);   // reference = norm(P b)
    ws.nrm2(w, kReference, st);
    if (has_x0) {
        A(x, t, ws.skip());
        ws.sub(b, t, t, st);
        precondition(t, w);
        ws.nrm2(w, kRnorm, st);
    } else {
        ws.zero(x);
        hipLaunchKernelGGL(k_copy_slot, dim3((R + 63) / 64), dim3(64), 0, ws.s, ws.scal, kReference, kRnorm, R);
        RT_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_initial_check, dim3((R + 63) / 64), dim3(64), 0, ws.s, ws.state, ws.scal, kRnorm, kReference,
                       R);
    RT_LAUNCH_CHECK();
    std::vector<LoopState> state;
    const auto first = ws.fetch_slots({kRnorm, kReference}, &state);
    const std::vector<double> &r0 = first[0], &ref = first[1];
    auto any_active = [&]() {
        for (const auto &q : state)
            if (!q.finished) return true;
        return false;
    };
    PollPlan plan;
    int it = 0;
    while (it < max_iter && any_active()) {
        hipLaunchKernelGGL(k_cycle_reset, dim3((R + 63) / 64), dim3(64), 0, ws.s, ws.state, R);
        RT_LAUNCH_CHECK();
        // v_0 = w / |w| into column 0 of the basis, g = (|w|, 0, ...).
        hipLaunchKernelGGL(k_gmres_start, dim3(vec_blocks(n), R), dim3(256), 0, ws.s, w, V, m1, n, ws.scal, g, st);
        RT_LAUNCH_CHECK();
        int step = 0;
        while (step < m && it < max_iter) {
            const int window = std::min({plan.next(state, tol), m - step, max_iter - it});
            for (int k = 0; k < window; ++k) {
                const int j = step;
                ++step;
                ++it;
                A(V + static_cast<int64_t>(j) * R * n, t, ws.skip());
                precondition(t, w);
                const int cnt = j + 1;
                for (int pass = 0; pass < 2; ++pass) {   // classical Gram-Schmidt, twice
                    hipLaunchKernelGGL(k_cgs_dots, dim3(ws.blocks, R, (cnt + 63) / 64), dim3(kThreads), 0, ws.s, V,
                                       cnt, w, n, ws.chunk, partial, m1, st);
                    RT_LAUNCH_CHECK();
                    hipLaunchKernelGGL(k_cgs_reduce, dim3((cnt + 3) / 4, R), dim3(kThreads), 0, ws.s, partial,
                                       ws.blocks, m1, cnt, hc, Hr, m, j, pass == 0 ? 1 : 0, st);
                    RT_LAUNCH_CHECK();
                    hipLaunchKernelGGL(k_cgs_gemv, dim3(vec_blocks(n), R), dim3(kThreads), 0, ws.s, V, cnt, hc, m1,
                                       -1.0, w, n, 1, st);
                    RT_LAUNCH_CHECK();
                }
                ws.nrm2(w, kSubSlot, st);
                hipLaunchKernelGGL(k_cgs_finalize, dim3(1, R), dim3(64), 0, ws.s, Hr, H, g, ws.scal, m, j, it,
                                   ws.state, ws.history, ws.cap);
                RT_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_cgs_next, dim3(vec_blocks(n), R), dim3(256), 0, ws.s, w, V, j + 1, n, ws.scal,
                                   st);
                RT_LAUNCH_CHECK();
            }
            state = ws.read_state();
            ws.set_scale(state);
            if (all_done(state)) break;
        }
        // x += V y with y = H \ g over the steps of this cycle (systems idle this cycle have 0 steps).
        hipLaunchKernelGGL(k_cgs_solve, dim3(1, R), dim3(kThreads), static_cast<uint32_t>(solve_lds), ws.s, H, g, m,
                           ws.state, Y);
        RT_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_cgs_gemv, dim3(vec_blocks(n), R), dim3(kThreads), 0, ws.s, V, -1, Y, m, 1.0, x, n, 0,
                           ws.state);
        RT_LAUNCH_CHECK();
        state = ws.read_state();
        if (it >= max_iter || !any_active()) break;
        // The next cycle starts from the true preconditioned residual (finished systems are skipped).
        hipLaunchKernelGGL(k_cycle_reset, dim3((R + 63) / 64), dim3(64), 0, ws.s, ws.state, R);
        RT_LAUNCH_CHECK();
        A(x, t, ws.skip());
        ws.sub(b, t, t, st);
        precondition(t, w);
        ws.nrm2(w, kRnorm, st);
    }
    RT_HIP(hipStreamSynchronize(ws.s));
    ws.ctx->timed_scale = 1.0;
    return collect(ws, state, skip_sys, r0, ref);
}

// ------------------------------------------------------------------------------------------------------------------
// GMRES(m), left preconditioned, CGS2 (iterate_.m:1-85), on the systems not marked `skip_sys`.
// ------------------------------------------------------------------------------------------------------------------
template <typename VT>
static InnerResult gmres(Workspace &ws, const LinOp &A, const LinOp *P, const c128 *b, c128 *x, bool has_x0,
                         const std::vector<double> &tol, const std::vector<char> &skip_sys,
                         const std::vector<int> &limits, int restart) {
    constexpr bool low = sizeof(VT) == sizeof(c64);
    Context *ctx = ws.ctx;
    const int64_t n = ws.n;
    const int R = ws.R;
    const LoopState *st = ws.state;
    const int max_iter = put_limits(ws, limits, skip_sys);
    int64_t m64 = restart > 0 ? restart : n;
    m64 = std::min<int64_t>({m64, n, static_cast<int64_t>(max_iter)});
    const int m = static_cast<int>(m64);
    if (m > kFastRestart) return gmres_cgs(ws, A, P, b, x, has_x0, tol, skip_sys, max_iter, m);
    const int m1 = m + 1;
    VT *V = static_cast<VT *>(ctx->scratch(kSlotV, sizeof(VT) * R * m * n, "the solver's vectors"));
    c128 *pv = vec(ctx, kSlotP, R * n);   // the pending vector p (FP64), the one the operator is applied to
    c128 *w = vec(ctx, kSlotW, R * n), *t = vec(ctx, kSlotT, R * n);
    c128 *H = vec(ctx, kSlotH, static_cast<int64_t>(R) * m1 * m);
    c128 *Hr = vec(ctx, kSlotHraw, static_cast<int64_t>(R) * m1 * m);
    c128 *g = vec(ctx, kSlotG, static_cast<int64_t>(R) * 4 * m1);
    auto precondition = [&](const c128 *in, c128 *out) {
        if (P) (*P)(in, out, ws.skip());
        else ws.copy(in, out, st);
    };
    ws.reset_state(skip_sys);
    ws.put_re(kTol, tol);
    precondition(b, w);   // reference = norm(P b)
    ws.nrm2(w, kReference, st);
    if (has_x0) {
        A(x, t, ws.skip());
        ws.sub(b, t, t, st);
        precondition(t, w);
        ws.nrm2(w, kRnorm, st);
    } else {
        ws.zero(x);
        hipLaunchKernelGGL(k_copy_slot, dim3((R + 63) / 64), dim3(64), 0, ws.s, ws.scal, kReference, kRnorm, R);
        RT_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_initial_check, dim3((R + 63) / 64), dim3(64), 0, ws.s, ws.state, ws.scal, kRnorm, kReference,
                       R);
    RT_LAUNCH_CHECK();
    std::vector<LoopState> state;
    const auto first = ws.fetch_slots({kRnorm, kReference}, &state);
    const std::vector<double> &r0 = first[0], &ref = first[1];
    auto any_active = [&]() {
        for (const auto &q : state)
            if (!q.finished) return true;
        return false;
    };
    PollPlan plan;
    size_t timed_mark = ctx->pending.size();
    int it = 0;
    while (it < max_iter && any_active()) {
        hipLaunchKernelGGL(k_cycle_reset, dim3((R + 63) / 64), dim3(64), 0, ws.s, ws.state, R);
        RT_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_gmres_start, dim3(vec_blocks(n), R), dim3(256), 0, ws.s, w, pv, m1, n, ws.scal, g, st);
        RT_LAUNCH_CHECK();
        int step = 0;
        while (step < m && it < max_iter) {
            const int window = std::min({plan.next(state, tol), m - step, max_iter - it});
            for (int k = 0; k < window; ++k) {
                const int j = step;
                ++step;
                ++it;
                ctx->timed_step = it;
                A(pv, t, ws.skip());
                precondition(t, w);
                // Algorithmic bytes of the two passes: the live basis columns once, p and w read, q_j and p written.
                const double col = static_cast<double>(sizeof(VT)) * n * R, wb = 16.0 * n * R;
                KernelTimer t1(ctx, RT_TIMED_GS_DOTS, j * col + 2 * wb);
                const size_t tile_lds = std::max(sizeof(VT) * j * kTileStride, kDotsRedBytes);
                if (tile_lds > 32 * 1024) raise_lds(reinterpret_cast<const void *>(&k_dcgs_dots<VT>), tile_lds);
                launch_timed(&t1, k_dcgs_dots<VT>, dim3(ws.dot_blocks, R), dim3(kThreads),
                             static_cast<uint32_t>(tile_lds), ws.s, V, j, pv, w, n, ws.dot_per, ws.dot_strided,
                             ws.partial, st);
                t1.done();
                // Algorithmic bytes: the dot pass's partial sums once (the Hessenberg entries and scalars are few).
                KernelTimer tf(ctx, RT_TIMED_GS_SUMS, 16.0 * (2 * j + 3) * ws.dot_blocks * R);
                launch_timed(&tf, k_dgmres_sums, dim3((2 * j + 3 + 3) / 4, R), dim3(kThreads), 0, ws.s,
                             static_cast<const c128 *>(ws.partial), ws.dot_blocks, ws.scal, j, it, ws.state);
                tf.done();
                KernelTimer t2(ctx, RT_TIMED_GS_UPDATE, j * col + 3 * wb + col);
                // A few systems leave the chip short of rows: four lanes per row then. The first block per system
                // finalizes the step.
                if (ws.gs_tile_update)
                    launch_timed(&t2, k_dcgs_update<VT, 4>, dim3(ws.gs_blocks + kLeadBlocks, R), dim3(kThreads), 0,
                                 ws.s, V, j, pv, w, n, ws.gs_chunk, ws.scal, it, ws.state, Hr, H, g, m, ws.history,
                                 ws.cap);
                else
                    launch_timed(&t2, k_dcgs_update<VT, 1>, dim3(ws.blocks + kLeadBlocks, R), dim3(kThreads), 0, ws.s,
                                 V, j, pv, w, n, ws.chunk, ws.scal, it, ws.state, Hr, H, g, m, ws.history, ws.cap);
                t2.done();
            }
            ctx->timed_step = -1;
            state = ws.read_state();
            timed_mark = settle_timed(ctx, state, timed_mark);
            ws.set_scale(state);
            if (all_done(state)) break;
        }
        // x += V y with y = H \ g over the steps of this cycle (systems idle this cycle have 0 steps).
        hipLaunchKernelGGL(k_gmres_solve, dim3(1, R), dim3(64), 0, ws.s, H, g, m, ws.state, ws.scal);
        RT_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_gemv_add<VT>, dim3(vec_blocks(n), R), dim3(256), 0, ws.s, V, ws.state, ws.scal, x, n);
        RT_LAUNCH_CHECK();
        state = ws.read_state();
        if (it >= max_iter || !any_active()) break;
        // The next cycle starts from the true preconditioned residual (finished systems are skipped).
        hipLaunchKernelGGL(k_cycle_reset, dim3((R + 63) / 64), dim3(64), 0, ws.s, ws.state, R);
        RT_LAUNCH_CHECK();
        A(x, t, ws.skip());
        ws.sub(b, t, t, st);
        precondition(t, w);
        ws.nrm2(w, kRnorm, st);
    }
    RT_HIP(hipStreamSynchronize(ws.s));
    ctx->timed_scale = 1.0;
    return collect(ws, state, skip_sys, r0, ref);
}

// ------------------------------------------------------------------------------------------------------------------
// IterativeMethod.solve (IterativeMethod.m:27-70) on R systems, with the optional FP64 refinement loop on the true
// residual.
// ------------------------------------------------------------------------------------------------------------------
std::vector<SolveResult> solve_batch(Context *ctx, Product &op, Product *pc, const SolveOptions &o, const c128 *b,
                                     const c128 *x0, c128 *x, int R) {
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t n = op.dims.n;
    Workspace ws(ctx, n, R, o.max_iterations);
    hipStream_t s = ctx->stream;
    std::vector<SolveResult> out(R);
    ws.nrm2(b, kOuterB, nullptr);
    const std::vector<double> bnorm = ws.fetch_re(kOuterB);
    std::vector<char> zero_rhs(R, 0);
    bool any = false;
    for (int r = 0; r < R; ++r) {
        zero_rhs[r] = !(bnorm[r] > 0);
        any = any || !zero_rhs[r];
    }
    c128 *t = vec(ctx, kSlotOuterT, R * n);
    const int outer_prec = op.has(RT_C128) ? RT_C128 : RT_C64;   // the outer residual in FP64 when held
    // norm(b - A x) / norm(b) of every system; t keeps b - A x.
    auto true_residual = [&](const c128 *xx) {
        op.apply(xx, t, R, outer_prec, s);
        ws.sub(b, t, t, nullptr);
        ws.nrm2(t, kOuterRes, nullptr);
        std::vector<double> rel = ws.fetch_re(kOuterRes);
        for (int r = 0; r < R; ++r) rel[r] = zero_rhs[r] ? 0.0 : rel[r] / bnorm[r];
        return rel;
    };
    int inner_prec = o.precision;
    LinOp A = [&](const c128 *in, c128 *res, const int *skip) {
        op.apply(in, res, R, inner_prec, s, skip, kStateInts);
    };
    LinOp Pop = [&](const c128 *in, c128 *res, const int *skip) {
        pc->apply(in, res, R, inner_prec, s, skip, kStateInts);
    };
    const LinOp *P = pc ? &Pop : nullptr;
    auto run = [&](const c128 *rhs, c128 *sol, bool has_x0, const std::vector<double> &tol,
                   const std::vector<char> &skip_sys, const std::vector<int> &limits, int stall) {
        if (o.method == RT_METHOD_CG) return cocg(ws, A, P, rhs, sol, has_x0, tol, skip_sys, limits, stall);
        if (o.method == RT_METHOD_BLOCK_CG)
            return block_cocg(ws, A, P, rhs, sol, has_x0, tol, skip_sys, limits, stall);
        // A complex64 inner GMRES keeps its basis in complex64: half the basis traffic of the Gram-Schmidt passes.
        if (inner_prec == RT_C64 && !o.fp64_basis)
            return gmres<c64>(ws, A, P, rhs, sol, has_x0, tol, skip_sys, limits, o.restart);
        return gmres<c128>(ws, A, P, rhs, sol, has_x0, tol, skip_sys, limits, o.restart);
    };
    if (!any) {
        ws.zero(x);
        for (auto &r : out) {
            r.converged = true;
            r.norms.push_back(0.0);
        }
    } else if (!o.refine) {
        if (x0 && x0 != x) ws.copy(x0, x, nullptr);
        std::vector<double> tol(R, o.tolerance);
        InnerResult ir = run(b, x, x0 != nullptr, tol, zero_rhs, std::vector<int>(R, o.max_iterations), 0);
        const std::vector<double> rel = true_residual(x);
        for (int r = 0; r < R; ++r) {
            if (zero_rhs[r]) {
                out[r].converged = true;
                out[r].norms.push_back(0.0);
                continue;
            }
            out[r].converged = ir.converged[r];
            out[r].iterations = ir.iterations[r];
            out[r].norms = ir.norms[r];
            out[r].residual = rel[r];
        }
    } else {
        // x_{k+1} = x_k + d_k, A d_k = b - A x_k solved by the inner method in its precision.
        c128 *rv = vec(ctx, kSlotOuterR, R * n), *d = vec(ctx, kSlotOuterD, R * n);
        std::vector<double> rel(R, 1.0);
        if (x0) {
            ws.copy(x0, x, nullptr);
            rel = true_residual(x);
        } else {   // x = 0: the residual is b itself, at relative norm 1
            ws.zero(x);
            ws.copy(b, t, nullptr);
            for (int r = 0; r < R; ++r)
                if (zero_rhs[r]) rel[r] = 0.0;
        }
        // Inner tolerance: complex64 GMRES reaches about 1e-6 before the rounding of its products shows; COCG's
        // short recurrence loses the complex symmetry of complex64 products near 1e-4 (sooner on systems whose
        // residual norm swings high on the way), so a mixed COCG hands what is left after its first inner solve to
        // FP64 inner solves when the products hold FP64 (and, without them, refines with complex64 inner solves to
        // 1e-4 at most).
        double inner_tol = o.tolerance;
        if (o.inner_tolerance > 0) inner_tol = o.inner_tolerance;
        const bool cg_like = o.method == RT_METHOD_CG || o.method == RT_METHOD_BLOCK_CG;
        const bool fallback = kCocgFallback && cg_like && o.precision == RT_C64 &&
                              op.has(RT_C128) && (!pc || pc->has(RT_C128));
        if (o.inner_tolerance <= 0 && o.precision == RT_C64)
            inner_tol = std::max(o.tolerance, cg_like ? (fallback ? kCocgInner : 1e-4) : 1e-6);
        // A block COCG cycle reduces its residual by kBlockInner at most, in either precision: a long block recurrence
        // loses its conjugacy and then its recursive residual parts from the true one.
        const bool block = o.method == RT_METHOD_BLOCK_CG && o.inner_tolerance <= 0;
        if (block) inner_tol = std::max(o.tolerance, kBlockInner);
        std::vector<char> done(R);
        std::vector<int> total(R, 0);
        for (int r = 0; r < R; ++r) {
            done[r] = zero_rhs[r] || rel[r] < o.tolerance;
            out[r].norms.push_back(rel[r]);
        }
        // An inner complex64 COCG that stagnates (no new best norm for kStall iterations) hands back to this loop; so
        // does every inner block COCG (whose recurrence can lose its conjugacy in any precision) and every FP64 inner
        // COCG: each restart begins a fresh recurrence from the true residual.
        int stall = cg_like ? kStall : 0;
        // Block COCG: a system whose residual a block cycle does not halve goes on alone (COCG, the block recurrence
        // of one column) from its iterate in the later cycles. On an indefinite system the recurrence of many
        // right-hand sides can lose its conjugacy, and a system caught in that would not converge in the block.
        std::vector<char> solo(R, 0);
        c128 *d_solo = nullptr;   // the lone systems' corrections (every inner solve zeroes its whole output)
        int outer = 0;
        while (outer < o.max_outer) {
            bool active = false;
            std::vector<int> limits(R, 0);
            for (int r = 0; r < R; ++r) {
                if (!done[r] && total[r] >= o.max_iterations) done[r] = 1;   // this system's budget is spent
                limits[r] = o.max_iterations - total[r];
                active = active || !done[r];
            }
            if (!active) break;
            ++outer;
            if (fallback && outer == 2) {   // the rest of a mixed COCG in FP64
                inner_prec = RT_C128;
                if (!block) inner_tol = o.tolerance;
            }
            ws.copy(t, rv, nullptr);   // t holds b - A x from true_residual
            // Each system asks the inner solve for what is left to reach the tolerance, no less than the floor.
            std::vector<double> want(R, 0.5);
            for (int r = 0; r < R; ++r)
                if (!done[r]) want[r] = std::min(0.5, std::max(inner_tol, 0.5 * o.tolerance / rel[r]));
            bool any_solo = false, any_block = false;
            std::vector<char> skip_block(R), skip_solo(R);
            for (int r = 0; r < R; ++r) {
                skip_block[r] = done[r] || solo[r];
                skip_solo[r] = done[r] || !solo[r];
                any_block = any_block || !skip_block[r];
                any_solo = any_solo || !skip_solo[r];
            }
            InnerResult ir;
            if (!any_solo) {
                ir = run(rv, d, false, want, done, limits, stall);
            } else {
                if (!d_solo) d_solo = vec(ctx, kSlotOuterSolo, R * n);
                const InnerResult lone = cocg(ws, A, P, rv, d_solo, false, want, skip_solo, limits, stall);
                if (any_block) {
                    ir = block_cocg(ws, A, P, rv, d, false, want, skip_block, limits, stall);
                } else {
                    ws.zero(d);
                    ir = lone;
                }
                for (int r = 0; r < R; ++r)
                    if (solo[r]) {
                        ir.converged[r] = lone.converged[r];
                        ir.iterations[r] = lone.iterations[r];
                        ir.norms[r] = lone.norms[r];
                    }
                ws.add(d_solo, x, nullptr);   // zero for the systems in the block
            }
            for (int r = 0; r < R; ++r) {
                if (done[r]) continue;
                total[r] += ir.iterations[r];
                out[r].outer = outer;
                for (size_t k = 1; k < ir.norms[r].size(); ++k) out[r].norms.push_back(ir.norms[r][k] * rel[r]);
            }
            ws.add(d, x, nullptr);   // d is zero for the systems the inner solve left out
            const std::vector<double> before = rel;
            rel = true_residual(x);
            for (int r = 0; r < R; ++r) {
                if (!done[r] && (rel[r] < o.tolerance || !std::isfinite(rel[r]))) done[r] = 1;
                if (block && !done[r] && !solo[r] && !(rel[r] < 0.5 * before[r])) solo[r] = 1;
            }
        }
        for (int r = 0; r < R; ++r) {
            out[r].iterations = total[r];
            out[r].residual = rel[r];
            out[r].converged = zero_rhs[r] || rel[r] < o.tolerance;
        }
    }
    // A zero right-hand side has the zero solution (IterativeMethod.m:50-57), whatever the initial guess.
    for (int r = 0; r < R; ++r)
        if (zero_rhs[r]) RT_HIP(hipMemsetAsync(x + static_cast<int64_t>(r) * n, 0, sizeof(c128) * n, s));
    RT_HIP(hipStreamSynchronize(s));
    const double time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (auto &r : out) r.time = time;
    return out;
}

}  // namespace rt

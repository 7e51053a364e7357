// Pass B of the matvec (passes.hpp): for `tile` consecutive x positions, the forward y transforms of both components
// (contiguous work columns, zero-padded to My), the three-plane spectral multiply, the inverse y transforms and the
// cut back to the columns. A team of T threads owns one x position and both of its columns; LDS holds
// [position][component][y (padded)].
#include "passes.hpp"

namespace rt {

// The last forward round, the multiply and the first inverse round on the registers of one butterfly at a time.
// Spectra: row = bspec + fold(kx) My 3, position py holds (Mxx, Myy, E) at row[3 py ..]; dy in position order.
// The multiply runs in chunks of CH positions behind scheduling barriers, which bounds the spectrum loads in flight
// (and the registers they hold). FOUR: a general 2x2 block per frequency instead, row = bspec + px My 4 holding
// (XX, XY, YX, YY) at row[4 py ..] (the blockwise circulant preconditioner, whose coupling is not a multiple of one
// plane).
template <typename TC, class P, bool DIRECT, bool EMB, bool FOUR = false>
__device__ __forceinline__ void multiply_round(int t, cplx<TC> *col0, cplx<TC> *col1, cplx<TC> *__restrict__ w0,
                                               cplx<TC> *__restrict__ w1, int n0, int n1,
                                               const cplx<TC> *__restrict__ row, cplx<TC> dxv,
                                               const cplx<TC> *__restrict__ dy) {
    constexpr int S = P::K - 1;
    constexpr int R = 1 << P::lr(S), U = P::E / R;
    constexpr int CH = sizeof(TC) == sizeof(float) ? 8 : 4;
#pragma unroll 1
    for (int u = 0; u < U; ++u) {
        const int base = (t + u * P::T) * R;   // the last round has stride 1: positions base + q
        const int pb = P::pad(base);
        cplx<TC> X[R], Y[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            if constexpr (DIRECT) {
                const bool in_half = !EMB || R == 1 || q < R / 2;   // an embedding's upper half is zero
                // Loads from a valid column (clamped), then a select: a load under a branch is waited for inside it.
                const cplx<TC> x = w0[min(base + q, max(n0 - 1, 0))], y = w1[min(base + q, max(n1 - 1, 0))];
                X[q] = in_half && base + q < n0 ? x : mk<TC>(0, 0);
                Y[q] = in_half && base + q < n1 ? y : mk<TC>(0, 0);
            } else {
                X[q] = col0[pb + P::pad(q)];
                Y[q] = col1[pb + P::pad(q)];
            }
        }
        dft<TC, R, false, false>(X);
        dft<TC, R, false, false>(Y);
        constexpr int NP = FOUR ? 4 : 3;
        const cplx<TC> *sp = row + NP * base;
        const cplx<TC> *dyp = dy + base;
#pragma unroll
        for (int q0 = 0; q0 < R; q0 += CH) {
#pragma unroll
            for (int q = q0; q < q0 + CH && q < R; ++q) {
                if constexpr (FOUR) {
                    const cplx<TC> xx = sp[4 * q], xy = sp[4 * q + 1], yx = sp[4 * q + 2], yy = sp[4 * q + 3];
                    const cplx<TC> x = X[q];
                    X[q] = xx * x + xy * Y[q];
                    Y[q] = yx * x + yy * Y[q];
                } else {
                    const cplx<TC> a = sp[3 * q], b = sp[3 * q + 1], el = sp[3 * q + 2];
                    const cplx<TC> dyv = dyp[q];
                    const cplx<TC> phi = el * (cmulc(dxv, X[q]) + cmulc(dyv, Y[q]));
                    X[q] = a * X[q] - dxv * phi;
                    Y[q] = b * Y[q] - dyv * phi;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        dft<TC, R, true, false>(X);
        dft<TC, R, true, false>(Y);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            if constexpr (DIRECT) {
                if (base + q < n0) w0[base + q] = X[q];
                if (base + q < n1) w1[base + q] = Y[q];
            } else {
                col0[pb + P::pad(q)] = X[q];
                col1[pb + P::pad(q)] = Y[q];
            }
        }
    }
}

// EMB: the columns are the lower halves of an embedding (the upper half of every input zero, only the lower half of
// every output kept), as in the Toeplitz product; false for full-period columns (a circulant on the face grid).
// FOUR: the four-plane multiply (unfolded blocks per x position).
template <typename TC, int LOG2M, int BITS, int MAXT, bool EMB = true, bool FOUR = false>
__global__ void __launch_bounds__(MAXT) pass_b_k(cplx<TC> *__restrict__ work, const cplx<TC> *__restrict__ bspec,
                                                 const int *__restrict__ fold_x, const int *__restrict__ border,
                                                 const cplx<TC> *__restrict__ dx,
                                                 const cplx<TC> *__restrict__ dy, ProdDims d,
                                                 const cplx<TC> *__restrict__ tw, int tile, int R, int nblocks,
                                                 const int *__restrict__ skip, int skip_stride) {
    using P = BPlan<LOG2M, BITS, sizeof(cplx<TC>)>;
    const int lb = xcd_block(blockIdx.x, nblocks);
    if (lb < 0) return;
    const int r = lb % R, pt = lb / R;
    if (skip && skip[r * skip_stride]) return;
    extern __shared__ unsigned char smem_raw[];
    constexpr int Mp = P::pad(P::M - 1) + 1;
    const int tt = threadIdx.x / P::T, t = threadIdx.x - tt * P::T;
    const int px = border[pt * tile + tt];   // partners k and M - k run together
    cplx<TC> *col0 = reinterpret_cast<cplx<TC> *>(smem_raw) + 2 * tt * Mp;
    cplx<TC> *col1 = col0 + Mp;
    cplx<TC> *w0 = work + work_index(d, r, 0, px, 0);
    cplx<TC> *w1 = work + work_index(d, r, 1, px, 0);
    const int n0 = d.cols_x, n1 = d.cols_y;
    const cplx<TC> *row = FOUR ? bspec + static_cast<int64_t>(px) * d.My * 4
                               : bspec + r * d.spec_stride + static_cast<int64_t>(fold_x[px]) * d.My * 3;
    dx += r * d.dx_stride;
    dy += r * d.dy_stride;
    const cplx<TC> dxv = dx[px];
    if constexpr (P::K == 1) {   // one round: straight from the work array and back
        multiply_round<TC, P, true, EMB, FOUR>(t, col0, col1, w0, w1, n0, n1, row, dxv, dy);
    } else {
        // Forward round 0 from the work array (the upper half of the embedding is zero).
        cround<TC, P, 0, false, EMB, true>(
            t, tw, [&](int pos, int) { return clamped_load(w0, pos, n0); },
            [&](int, int pp, cplx<TC> v) { col0[pp] = v; });
        cround<TC, P, 0, false, EMB, true>(
            t, tw, [&](int pos, int) { return clamped_load(w1, pos, n1); },
            [&](int, int pp, cplx<TC> v) { col1[pp] = v; });
        __syncthreads();
        static_for<1, P::K - 1>([&](auto S) {
            constexpr int s = decltype(S)::value;
            cround<TC, P, s, false, false, true>(
                t, tw, [&](int, int pp) { return col0[pp]; }, [&](int, int pp, cplx<TC> v) { col0[pp] = v; });
            cround<TC, P, s, false, false, true>(
                t, tw, [&](int, int pp) { return col1[pp]; }, [&](int, int pp, cplx<TC> v) { col1[pp] = v; });
            __syncthreads();
        });
        multiply_round<TC, P, false, EMB, FOUR>(t, col0, col1, w0, w1, n0, n1, row, dxv, dy);
        __syncthreads();
        tw = opaque(tw);
        static_for_down<1, P::K - 1>([&](auto S) {
            constexpr int s = decltype(S)::value;
            cround<TC, P, s, true, false, true>(
                t, tw, [&](int, int pp) { return col0[pp]; }, [&](int, int pp, cplx<TC> v) { col0[pp] = v; });
            cround<TC, P, s, true, false, true>(
                t, tw, [&](int, int pp) { return col1[pp]; }, [&](int, int pp, cplx<TC> v) { col1[pp] = v; });
            __syncthreads();
        });
        // Inverse round 0 into the work array, cut to the columns (the upper half of the period is never kept).
        cround<TC, P, 0, true, EMB, true>(
            t, tw, [&](int, int pp) { return col0[pp]; },
            [&](int pos, int, cplx<TC> v) {
                if (pos < n0) w0[pos] = v;
            });
        cround<TC, P, 0, true, EMB, true>(
            t, tw, [&](int, int pp) { return col1[pp]; },
            [&](int pos, int, cplx<TC> v) {
                if (pos < n1) w1[pos] = v;
            });
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Pass B with a team per component: the 2 T threads of an x position form a team for each of its two columns, which
// doubles the waves that share the position's LDS (the LDS, not the registers, bounds the blocks per CU for long
// columns). Each team runs its column's rounds; the last forward round, the multiply and the first inverse round
// stay in the registers of the team's own column, and the multiply reads the other column's transformed values
// from the LDS. Needs T >= 64, so that each wave belongs to one team and the role is wave-uniform.
// ------------------------------------------------------------------------------------------------------------------
template <typename TC, class P, int C>
__device__ __forceinline__ void multiply_team(int t, cplx<TC> *mine, const cplx<TC> *other,
                                              const cplx<TC> *__restrict__ row, cplx<TC> dxv,
                                              const cplx<TC> *__restrict__ dy) {
    constexpr int S = P::K - 1;
    constexpr int R = 1 << P::lr(S), U = P::E / R;
    constexpr int CH = sizeof(TC) == sizeof(float) ? 8 : 4;
    cplx<TC> v[U][R];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int base = (t + u * P::T) * R;   // the last round has stride 1: positions base + q
        const int pb = P::pad(base);
#pragma unroll
        for (int q = 0; q < R; ++q) v[u][q] = mine[pb + P::pad(q)];
        dft<TC, R, false, false>(v[u]);
#pragma unroll
        for (int q = 0; q < R; ++q) mine[pb + P::pad(q)] = v[u][q];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int base = (t + u * P::T) * R;
        const int pb = P::pad(base);
        const cplx<TC> *sp = row + 3 * base;
        const cplx<TC> *dyp = dy + base;
#pragma unroll
        for (int q0 = 0; q0 < R; q0 += CH) {
#pragma unroll
            for (int q = q0; q < q0 + CH && q < R; ++q) {
                const cplx<TC> o = other[pb + P::pad(q)];
                const cplx<TC> self = sp[3 * q + C], el = sp[3 * q + 2];
                const cplx<TC> dyv = dyp[q];
                cplx<TC> &w = v[u][q];
                if constexpr (C == 0) {
                    const cplx<TC> phi = el * (cmulc(dxv, w) + cmulc(dyv, o));
                    w = self * w - dxv * phi;
                } else {
                    const cplx<TC> phi = el * (cmulc(dxv, o) + cmulc(dyv, w));
                    w = self * w - dyv * phi;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        dft<TC, R, true, false>(v[u]);
    }
    __syncthreads();   // the other team has read this column
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int base = (t + u * P::T) * R;
        const int pb = P::pad(base);
#pragma unroll
        for (int q = 0; q < R; ++q) mine[pb + P::pad(q)] = v[u][q];
    }
}

template <typename TC, class P, int C>
__device__ __forceinline__ void pass_b_team(int t, cplx<TC> *mine, const cplx<TC> *other, cplx<TC> *w, int nc,
                                            const cplx<TC> *tw, const cplx<TC> *row, cplx<TC> dxv,
                                            const cplx<TC> *dy) {
    cround<TC, P, 0, false, true, true>(
        t, tw, [&](int pos, int) { return clamped_load(w, pos, nc); },
        [&](int, int pp, cplx<TC> v) { mine[pp] = v; });
    __syncthreads();
    static_for<1, P::K - 1>([&](auto S) {
        constexpr int s = decltype(S)::value;
        cround<TC, P, s, false, false, true>(
            t, tw, [&](int, int pp) { return mine[pp]; }, [&](int, int pp, cplx<TC> v) { mine[pp] = v; });
        __syncthreads();
    });
    multiply_team<TC, P, C>(t, mine, other, row, dxv, dy);
    __syncthreads();
    tw = opaque(tw);
    static_for_down<1, P::K - 1>([&](auto S) {
        constexpr int s = decltype(S)::value;
        cround<TC, P, s, true, false, true>(
            t, tw, [&](int, int pp) { return mine[pp]; }, [&](int, int pp, cplx<TC> v) { mine[pp] = v; });
        __syncthreads();
    });
    cround<TC, P, 0, true, true, true>(
        t, tw, [&](int, int pp) { return mine[pp]; },
        [&](int pos, int, cplx<TC> v) {
            if (pos < nc) w[pos] = v;
        });
}

template <typename TC, int LOG2M, int BITS, int MAXT>
__global__ void __launch_bounds__(MAXT) pass_b2_k(cplx<TC> *__restrict__ work, const cplx<TC> *__restrict__ bspec,
                                                  const int *__restrict__ fold_x, const int *__restrict__ border,
                                                  const cplx<TC> *__restrict__ dx,
                                                  const cplx<TC> *__restrict__ dy, ProdDims d,
                                                  const cplx<TC> *__restrict__ tw, int tile, int R, int nblocks,
                                                  const int *__restrict__ skip, int skip_stride) {
    using P = BPlan<LOG2M, BITS, sizeof(cplx<TC>)>;
    static_assert(P::K >= 2 && P::T % 64 == 0, "a team per component needs two rounds and whole waves");
    const int lb = xcd_block(blockIdx.x, nblocks);
    if (lb < 0) return;
    const int r = lb % R, pt = lb / R;
    if (skip && skip[r * skip_stride]) return;
    extern __shared__ unsigned char smem_raw[];
    constexpr int Mp = P::pad(P::M - 1) + 1;
    const int tt = threadIdx.x / (2 * P::T);
    const int rem = threadIdx.x - tt * 2 * P::T;
    const int c = __builtin_amdgcn_readfirstlane(rem / P::T);   // the team's component, uniform in the wave
    const int t = rem - c * P::T;
    const int px = border[pt * tile + tt];   // partners k and M - k run together
    cplx<TC> *col0 = reinterpret_cast<cplx<TC> *>(smem_raw) + 2 * tt * Mp;
    cplx<TC> *col1 = col0 + Mp;
    const cplx<TC> *row = bspec + r * d.spec_stride + static_cast<int64_t>(fold_x[px]) * d.My * 3;
    dx += r * d.dx_stride;
    dy += r * d.dy_stride;
    const cplx<TC> dxv = dx[px];
    if (c == 0)
        pass_b_team<TC, P, 0>(t, col0, col1, work + work_index(d, r, 0, px, 0), d.cols_x, tw, row, dxv, dy);
    else
        pass_b_team<TC, P, 1>(t, col1, col0, work + work_index(d, r, 1, px, 0), d.cols_y, tw, row, dxv, dy);
}

// ------------------------------------------------------------------------------------------------------------------
// Pass B through one column of LDS: the team transforms the X column of its x position, keeps the outputs of the
// last forward round in registers (E per thread), transforms the Y column through the same LDS, multiplies in
// registers and runs the inverse transforms, Y first, back through the LDS. Half the LDS per position of pass_b_k,
// for twice the barriers and a column held in registers: long columns fit twice the blocks per CU, and 8192-point
// complex128 columns fit one launch.
// ------------------------------------------------------------------------------------------------------------------
template <typename TC, class P>
struct HeldColumn {
    static constexpr int R = 1 << P::lr(P::K - 1), U = P::E / R;   // the last round: radix R, U butterflies
    cplx<TC> v[U][R];                                               // position (t + u T) R + q at v[u][q]
};

// Forward y transform of the work column w (n values, zero beyond) through buf; the last round stays in h.
template <typename TC, class P>
__device__ __forceinline__ void b1_forward(int t, cplx<TC> *buf, const cplx<TC> *__restrict__ w, int n,
                                           const cplx<TC> *__restrict__ tw, HeldColumn<TC, P> &h) {
    constexpr int R = HeldColumn<TC, P>::R, U = HeldColumn<TC, P>::U;
    cround<TC, P, 0, false, true, true>(
        t, tw, [&](int pos, int) { return clamped_load(w, pos, n); }, [&](int, int pp, cplx<TC> v) { buf[pp] = v; });
    __syncthreads();
    static_for<1, P::K - 1>([&](auto S) {
        cround<TC, P, decltype(S)::value, false, false, true>(
            t, tw, [&](int, int pp) { return buf[pp]; }, [&](int, int pp, cplx<TC> v) { buf[pp] = v; });
        __syncthreads();
    });
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int pb = P::pad((t + u * P::T) * R);
#pragma unroll
        for (int q = 0; q < R; ++q) h.v[u][q] = buf[pb + P::pad(q)];
        dft<TC, R, false, false>(h.v[u]);
    }
    __syncthreads();   // buf is free for the next column
}

// Inverse y transform of the held column h (its first inverse round already applied) through buf into the work
// column w, cut to n values.
template <typename TC, class P>
__device__ __forceinline__ void b1_inverse(int t, cplx<TC> *buf, cplx<TC> *__restrict__ w, int n,
                                           const cplx<TC> *__restrict__ tw, const HeldColumn<TC, P> &h) {
    constexpr int R = HeldColumn<TC, P>::R, U = HeldColumn<TC, P>::U;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int pb = P::pad((t + u * P::T) * R);
#pragma unroll
        for (int q = 0; q < R; ++q) buf[pb + P::pad(q)] = h.v[u][q];
    }
    __syncthreads();
    static_for_down<1, P::K - 1>([&](auto S) {
        cround<TC, P, decltype(S)::value, true, false, true>(
            t, tw, [&](int, int pp) { return buf[pp]; }, [&](int, int pp, cplx<TC> v) { buf[pp] = v; });
        __syncthreads();
    });
    cround<TC, P, 0, true, true, true>(
        t, tw, [&](int, int pp) { return buf[pp]; },
        [&](int pos, int, cplx<TC> v) {
            if (pos < n) w[pos] = v;
        });
    __syncthreads();   // buf is free for the next column
}

template <typename TC, int LOG2M, int BITS, int MAXT>
__global__ void __launch_bounds__(MAXT) pass_b1_k(cplx<TC> *__restrict__ work, const cplx<TC> *__restrict__ bspec,
                                                  const int *__restrict__ fold_x, const int *__restrict__ border,
                                                  const cplx<TC> *__restrict__ dx,
                                                  const cplx<TC> *__restrict__ dy, ProdDims d,
                                                  const cplx<TC> *__restrict__ tw, int tile, int R, int nblocks,
                                                  const int *__restrict__ skip, int skip_stride) {
    using P = BPlan<LOG2M, BITS, sizeof(cplx<TC>)>;
    static_assert(P::K >= 2, "one column of LDS needs at least two rounds");
    constexpr int RL = HeldColumn<TC, P>::R, U = HeldColumn<TC, P>::U;
    constexpr int CH = sizeof(TC) == sizeof(float) ? 4 : 2;   // both columns are live: fewer spectra in flight
    const int lb = xcd_block(blockIdx.x, nblocks);
    if (lb < 0) return;
    const int r = lb % R, pt = lb / R;
    if (skip && skip[r * skip_stride]) return;
    extern __shared__ unsigned char smem_raw[];
    constexpr int Mp = P::pad(P::M - 1) + 1;
    const int tt = threadIdx.x / P::T, t = threadIdx.x - tt * P::T;
    const int px = border[pt * tile + tt];   // partners k and M - k run together
    cplx<TC> *buf = reinterpret_cast<cplx<TC> *>(smem_raw) + tt * Mp;
    cplx<TC> *w0 = work + work_index(d, r, 0, px, 0);
    cplx<TC> *w1 = work + work_index(d, r, 1, px, 0);
    const cplx<TC> *row = bspec + r * d.spec_stride + static_cast<int64_t>(fold_x[px]) * d.My * 3;
    dx += r * d.dx_stride;
    dy += r * d.dy_stride;
    const cplx<TC> dxv = dx[px];
    // Each transform reads the twiddles through its own opaque pointer: the loads of one column's rounds are not
    // merged with another's, which would hold every round's twiddles in registers across a whole transform.
    HeldColumn<TC, P> X, Y;
    b1_forward<TC, P>(t, buf, w0, d.cols_x, opaque(tw), X);
    b1_forward<TC, P>(t, buf, w1, d.cols_y, opaque(tw), Y);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int base = (t + u * P::T) * RL;
        const cplx<TC> *sp = row + 3 * base;
        const cplx<TC> *dyp = dy + base;
#pragma unroll
        for (int q0 = 0; q0 < RL; q0 += CH) {
#pragma unroll
            for (int q = q0; q < q0 + CH && q < RL; ++q) {
                const cplx<TC> a = sp[3 * q], b = sp[3 * q + 1], el = sp[3 * q + 2];
                const cplx<TC> dyv = dyp[q];
                const cplx<TC> phi = el * (cmulc(dxv, X.v[u][q]) + cmulc(dyv, Y.v[u][q]));
                X.v[u][q] = a * X.v[u][q] - dxv * phi;
                Y.v[u][q] = b * Y.v[u][q] - dyv * phi;
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        dft<TC, RL, true, false>(X.v[u]);
        dft<TC, RL, true, false>(Y.v[u]);
    }
    b1_inverse<TC, P>(t, buf, w1, d.cols_y, opaque(tw), Y);
    b1_inverse<TC, P>(t, buf, w0, d.cols_x, opaque(tw), X);
}

template <typename TC>
void launch_pass_b(const Product &op, const PlanTables &tx, const PlanTables &ty, cplx<TC> *work,
                   const cplx<TC> *bspec, int nrhs, int tile, int teams, hipStream_t s, const int *skip,
                   int skip_stride, const KernelTimer *timer) {
    const FftPlan &p = ty.plan;
    constexpr int es = sizeof(cplx<TC>);
    const int64_t nb = static_cast<int64_t>(op.dims.Mx / tile) * nrhs;
    const size_t lds = lds_b_bytes(p, tile, es, teams);
    const int threads = team_threads(p) * tile;
    const bool dbl = sizeof(TC) == sizeof(double);
    const cplx<TC> *tw = dbl ? reinterpret_cast<const cplx<TC> *>(ty.tw.ptr)
                             : reinterpret_cast<const cplx<TC> *>(ty.tw32.ptr);
    const cplx<TC> *dx = dbl ? reinterpret_cast<const cplx<TC> *>(tx.d.ptr)
                             : reinterpret_cast<const cplx<TC> *>(tx.d32.ptr);
    const cplx<TC> *dy = dbl ? reinterpret_cast<const cplx<TC> *>(ty.d.ptr)
                             : reinterpret_cast<const cplx<TC> *>(ty.d32.ptr);
    if (!b_threads_fit(p, tile, teams)) fail(RT_EINVAL, "rooftop:product", "pass B shape does not fit this plan");
    if ((!op.embedding || op.four_plane) && (p.bits != 3 || teams != 1))
        fail(RT_EINVAL, "rooftop:product",
             "a face-grid or four-plane product runs the radix-8 plan with one team per position");
#define RT_LAUNCH_B(L, B)                                                                                          \
    {                                                                                                              \
        using PL = CPlan<L, B>;                                                                                    \
        if (op.four_plane) {                                                                                       \
            using PF = CPlan<L, 3>;                                                                                \
            constexpr int MT = PF::T > 256 ? PF::T : 256;                                                          \
            if (op.embedding) {                                                                                    \
                set_lds(reinterpret_cast<const void *>(&pass_b_k<TC, L, 3, MT, true, true>), lds);                 \
                launch_timed(timer, pass_b_k<TC, L, 3, MT, true, true>, dim3(xcd_grid(nb)), dim3(threads),         \
                             static_cast<uint32_t>(lds), s, work, bspec, tx.fold.ptr, tx.border.ptr, dx, dy, op.dims, tw, tile, \
                             nrhs, static_cast<int>(nb), skip, skip_stride);                                       \
            } else {                                                                                               \
                set_lds(reinterpret_cast<const void *>(&pass_b_k<TC, L, 3, MT, false, true>), lds);                \
                launch_timed(timer, pass_b_k<TC, L, 3, MT, false, true>, dim3(xcd_grid(nb)), dim3(threads),        \
                             static_cast<uint32_t>(lds), s, work, bspec, tx.fold.ptr, tx.border.ptr, dx, dy, op.dims, tw, tile, \
                             nrhs, static_cast<int>(nb), skip, skip_stride);                                       \
            }                                                                                                      \
        } else if (!op.embedding) {                                                                                \
            using PF = CPlan<L, 3>;                                                                                \
            constexpr int MT = PF::T > 256 ? PF::T : 256;                                                          \
            set_lds(reinterpret_cast<const void *>(&pass_b_k<TC, L, 3, MT, false>), lds);                          \
            launch_timed(timer, pass_b_k<TC, L, 3, MT, false>, dim3(xcd_grid(nb)), dim3(threads),                  \
                         static_cast<uint32_t>(lds), s, work, bspec, tx.fold.ptr, tx.border.ptr, dx, dy, op.dims, tw, tile, nrhs, \
                         static_cast<int>(nb), skip, skip_stride);                                                 \
        } else if (teams == 3) {                                                                                   \
            if constexpr (PL::K >= 2) {                                                                            \
                constexpr int MT = PL::T > 256 ? PL::T : 256;                                                      \
                set_lds(reinterpret_cast<const void *>(&pass_b1_k<TC, L, B, MT>), lds);                            \
                launch_timed(timer, pass_b1_k<TC, L, B, MT>, dim3(xcd_grid(nb)), dim3(threads),                    \
                             static_cast<uint32_t>(lds), s, work, bspec, tx.fold.ptr, tx.border.ptr, dx, dy, op.dims, tw, \
                             tile, nrhs, static_cast<int>(nb), skip, skip_stride);                                 \
            }                                                                                                      \
        } else if (teams == 1) {                                                                                   \
            constexpr int MT = PL::T > 256 ? PL::T : 256;                                                          \
            set_lds(reinterpret_cast<const void *>(&pass_b_k<TC, L, B, MT>), lds);                                 \
            launch_timed(timer, pass_b_k<TC, L, B, MT>, dim3(xcd_grid(nb)), dim3(threads),                         \
                         static_cast<uint32_t>(lds), s, work, bspec, tx.fold.ptr, tx.border.ptr, dx, dy, op.dims, tw, tile, nrhs, \
                         static_cast<int>(nb), skip, skip_stride);                                                 \
        } else if constexpr (PL::K >= 2 && PL::T % 64 == 0 && 2 * PL::T <= 1024) {                                 \
            constexpr int MT = 2 * PL::T > 256 ? 2 * PL::T : 256;                                                  \
            set_lds(reinterpret_cast<const void *>(&pass_b2_k<TC, L, B, MT>), lds);                                \
            launch_timed(timer, pass_b2_k<TC, L, B, MT>, dim3(xcd_grid(nb)), dim3(2 * threads),                    \
                         static_cast<uint32_t>(lds), s, work, bspec, tx.fold.ptr, tx.border.ptr, dx, dy, op.dims, tw, tile, nrhs, \
                         static_cast<int>(nb), skip, skip_stride);                                                 \
        }                                                                                                          \
    }
    RT_BY_PLAN(p.log2M, p.bits, RT_LAUNCH_B);
#undef RT_LAUNCH_B
}

}  // namespace rt

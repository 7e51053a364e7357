// Pass C of the matvec (passes.hpp): inverse x transforms of `group` consecutive columns read back from the work
// array; the first rows of each column are the product (the cut of FFTApplicator.m:80-84).
#include "passes.hpp"

namespace rt {

// EMB: only the lower half of every column is kept (the cut of an embedding); false keeps the full period.
template <typename TC, int LOG2M, int BITS, int MAXT, bool EMB = true>
__global__ void __launch_bounds__(MAXT) pass_c_k(const cplx<TC> *__restrict__ work, c128 *__restrict__ out,
                                                 const uint8_t *__restrict__ exists, ProdDims d,
                                                 const cplx<TC> *__restrict__ tw, int G, int CS, int nblocks,
                                                 const int *__restrict__ skip, int skip_stride) {
    using P = CPlan<LOG2M, BITS>;
    const int lb = xcd_block(blockIdx.x, nblocks);
    if (lb < 0) return;
    int r, comp, j0;
    group_of(d, G, lb, r, comp, j0);
    if (skip && skip[r * skip_stride]) return;
    extern __shared__ unsigned char smem_raw[];
    cplx<TC> *buf = reinterpret_cast<cplx<TC> *>(smem_raw);
    const int cols = comp ? d.cols_y : d.cols_x;
    const int rows = comp ? d.rows_y : d.rows_x;
    const int ncols = min(G, cols - j0);
    {
        const int c = threadIdx.x % G, t = threadIdx.x / G;
        const bool live = c < ncols;
        const cplx<TC> *src = work + work_index(d, r, comp, 0, j0 + (live ? c : 0));
        const int64_t ldj = d.ldj;
        cplx<TC> *lds = buf + c * CS;
        if constexpr (P::K > 1) {
            // Round K-1 (no twiddles), threads along the columns, read from the work array.
            cround<TC, P, P::K - 1, true, false, false>(
                t, tw,
                [&](int pos, int) {   // src is a valid column: load, then select (no load under a branch)
                    const cplx<TC> v = src[pos * ldj];
                    return live ? v : mk<TC>(0, 0);
                },
                [&](int, int pp, cplx<TC> v) { lds[pp] = v; });
            __syncthreads();
            static_for_down<1, P::K - 1>([&](auto S) {
                cround<TC, P, decltype(S)::value, true, false, true>(
                    t, tw, [&](int, int pp) { return lds[pp]; }, [&](int, int pp, cplx<TC> v) { lds[pp] = v; });
                __syncthreads();
            });
        } else {
            for (int pos = t; pos < P::M; pos += P::T) {
                const cplx<TC> v = src[pos * ldj];
                lds[P::pad(pos)] = live ? v : mk<TC>(0, 0);
            }
            __syncthreads();
        }
    }
    // Round 0, threads along x: writes the rooftop rows of column c.
    const int c = threadIdx.x / P::T, t = threadIdx.x - c * P::T;
    if (c >= ncols) return;
    const int64_t off = static_cast<int64_t>(r) * d.n + (comp ? d.nx_count : 0);
    const int64_t col = static_cast<int64_t>(j0 + c) * rows;
    c128 *dst = out + off + col;
    const cplx<TC> *lds = buf + c * CS;
    auto load = [&](int, int pp) { return lds[pp]; };
    if (exists) {   // a masked plate: the mask bytes are loaded before the branch around each store
        const uint8_t *exc = exists + r * d.mask_stride + (comp ? d.nx_count : 0) + col;
        cround<TC, P, 0, true, EMB && (P::M >= 2), (P::K > 1)>(t, tw, load, [&](int pos, int, cplx<TC> v) {
            const bool keep = exc[min(pos, rows - 1)] != 0;
            if (pos < rows) dst[pos] = keep ? cvt<double>(v) : mk(0.0, 0.0);
        });
    } else {
        cround<TC, P, 0, true, EMB && (P::M >= 2), (P::K > 1)>(t, tw, load, [&](int pos, int, cplx<TC> v) {
            if (pos < rows) dst[pos] = cvt<double>(v);
        });
    }
}

// Pass C with the columns in turn, the mirror of pass_a1_k: the team reads the rows of its G columns together (the
// first inverse round's inputs, held in registers), then runs each column's inverse transform through one column of
// LDS and writes its rooftop rows.
template <typename TC, int LOG2M, int BITS, int MAXT, int G>
__global__ void __launch_bounds__(MAXT) pass_c1_k(const cplx<TC> *__restrict__ work, c128 *__restrict__ out,
                                                  const uint8_t *__restrict__ exists, ProdDims d,
                                                  const cplx<TC> *__restrict__ tw, int nblocks,
                                                  const int *__restrict__ skip, int skip_stride) {
    using P = CPlan<LOG2M, BITS>;
    static_assert(P::K >= 2, "columns in turn need at least two rounds");
    constexpr int R = 1 << P::lr(P::K - 1), U = P::E / R;   // the first inverse round: radix R, U butterflies
    const int lb = xcd_block(blockIdx.x, nblocks);
    if (lb < 0) return;
    int r, comp, j0;
    group_of(d, G, lb, r, comp, j0);
    if (skip && skip[r * skip_stride]) return;
    extern __shared__ unsigned char smem_raw[];
    cplx<TC> *buf = reinterpret_cast<cplx<TC> *>(smem_raw);
    const int cols = comp ? d.cols_y : d.cols_x;
    const int rows = comp ? d.rows_y : d.rows_x;
    const int ncols = min(G, cols - j0);
    const int t = threadIdx.x;
    // Every input first: position (t + u T) R + q of the G columns (the columns past the last read the last again).
    const cplx<TC> *src = work + work_index(d, r, comp, 0, j0);
    const int64_t ldj = d.ldj;
    cplx<TC> held[G][U][R];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const cplx<TC> *row = src + static_cast<int64_t>((t + u * P::T) * R + q) * ldj;
#pragma unroll
            for (int c = 0; c < G; ++c) held[c][u][q] = row[min(c, ncols - 1)];
        }
    const int64_t off = static_cast<int64_t>(r) * d.n + (comp ? d.nx_count : 0);
#pragma unroll
    for (int c = 0; c < G; ++c) {
        if (c >= ncols) break;   // uniform over the block
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int pb = P::pad((t + u * P::T) * R);
            dft<TC, R, true, false>(held[c][u]);
#pragma unroll
            for (int q = 0; q < R; ++q) buf[pb + P::pad(q)] = held[c][u][q];
        }
        __syncthreads();
        static_for_down<1, P::K - 1>([&](auto S) {
            cround<TC, P, decltype(S)::value, true, false, true>(
                t, tw, [&](int, int pp) { return buf[pp]; }, [&](int, int pp, cplx<TC> v) { buf[pp] = v; });
            __syncthreads();
        });
        const int64_t col = static_cast<int64_t>(j0 + c) * rows;
        c128 *dst = out + off + col;
        auto load = [&](int, int pp) { return buf[pp]; };
        if (exists) {   // a masked plate: the mask bytes are loaded before the branch around each store
            const uint8_t *exc = exists + r * d.mask_stride + (comp ? d.nx_count : 0) + col;
            cround<TC, P, 0, true, true, true>(t, tw, load, [&](int pos, int, cplx<TC> v) {
                const bool keep = exc[min(pos, rows - 1)] != 0;
                if (pos < rows) dst[pos] = keep ? cvt<double>(v) : mk(0.0, 0.0);
            });
        } else {
            cround<TC, P, 0, true, true, true>(t, tw, load, [&](int pos, int, cplx<TC> v) {
                if (pos < rows) dst[pos] = cvt<double>(v);
            });
        }
        __syncthreads();   // buf is free for the next column
    }
}

template <typename TC>
void launch_pass_c(const Product &op, const PlanTables &t, const cplx<TC> *work, c128 *out, int nrhs, int g,
                   hipStream_t s, const int *skip, int skip_stride, const KernelTimer *timer) {
    const FftPlan &p = t.plan;
    constexpr int es = sizeof(cplx<TC>);
    const cplx<TC> *tw = sizeof(TC) == sizeof(double) ? reinterpret_cast<const cplx<TC> *>(t.tw.ptr)
                                                      : reinterpret_cast<const cplx<TC> *>(t.tw32.ptr);
    if (!op.embedding && p.bits != 3) fail(RT_EINVAL, "rooftop:product", "a face-grid product runs the radix-8 plan");
    if (g < 0) {   // -G: G columns in turn through one column of LDS
        if (!op.embedding || (g != -2 && g != -4) || p.K < 2)
            fail(RT_EINVAL, "rooftop:product", "columns in turn are 2 or 4 columns of an embedding's long transforms");
        const int64_t nb = blocks_ac(op.dims, nrhs, -g);
        const size_t lds = lds_ac_bytes(p, g, es);
#define RT_LAUNCH_C1(L, B)                                                                                  \
    {                                                                                                       \
        constexpr int MT = CPlan<L, B>::T > 64 ? CPlan<L, B>::T : 64;                                       \
        if constexpr (CPlan<L, B>::K >= 2) {                                                                \
            if (g == -2) {                                                                                  \
                set_lds(reinterpret_cast<const void *>(&pass_c1_k<TC, L, B, MT, 2>), lds);                  \
                launch_timed(timer, pass_c1_k<TC, L, B, MT, 2>, dim3(xcd_grid(nb)), dim3(team_threads(p)),  \
                             static_cast<uint32_t>(lds), s, work, out, op.problem->exists.ptr, op.dims, tw, \
                             static_cast<int>(nb), skip, skip_stride);                                      \
            } else {                                                                                        \
                set_lds(reinterpret_cast<const void *>(&pass_c1_k<TC, L, B, MT, 4>), lds);                  \
                launch_timed(timer, pass_c1_k<TC, L, B, MT, 4>, dim3(xcd_grid(nb)), dim3(team_threads(p)),  \
                             static_cast<uint32_t>(lds), s, work, out, op.problem->exists.ptr, op.dims, tw, \
                             static_cast<int>(nb), skip, skip_stride);                                      \
            }                                                                                               \
        }                                                                                                   \
    }
        RT_BY_PLAN(p.log2M, p.bits, RT_LAUNCH_C1);
#undef RT_LAUNCH_C1
        return;
    }
    const int64_t nb = blocks_ac(op.dims, nrhs, g);
    const int cs = column_stride(p, g, es);
    const size_t lds = lds_ac_bytes(p, g, es);
    const int threads = team_threads(p) * g;
#define RT_LAUNCH_C(L, B)                                                                                   \
    {                                                                                                       \
        constexpr int MT = (es == 16 && CPlan<L, B>::log2E >= 4) ? 512 : 1024;                              \
        constexpr int MF = (es == 16 && CPlan<L, 3>::log2E >= 4) ? 512 : 1024;                              \
        if (op.embedding) {                                                                                 \
            set_lds(reinterpret_cast<const void *>(&pass_c_k<TC, L, B, MT>), lds);                          \
            launch_timed(timer, pass_c_k<TC, L, B, MT>, dim3(xcd_grid(nb)), dim3(threads),                  \
                         static_cast<uint32_t>(lds), s, work, out, op.problem->exists.ptr, op.dims, tw, g, cs, \
                         static_cast<int>(nb), skip, skip_stride);                                          \
        } else {                                                                                            \
            set_lds(reinterpret_cast<const void *>(&pass_c_k<TC, L, 3, MF, false>), lds);                   \
            launch_timed(timer, pass_c_k<TC, L, 3, MF, false>, dim3(xcd_grid(nb)), dim3(threads),           \
                         static_cast<uint32_t>(lds), s, work, out, op.problem->exists.ptr, op.dims, tw, g, cs, \
                         static_cast<int>(nb), skip, skip_stride);                                          \
        }                                                                                                   \
    }
    RT_BY_PLAN(p.log2M, p.bits, RT_LAUNCH_C);
#undef RT_LAUNCH_C
}

}  // namespace rt

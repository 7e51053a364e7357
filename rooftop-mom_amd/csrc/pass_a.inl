// Pass A of the matvec (passes.hpp): forward x transforms of `group` consecutive rooftop columns of one component,
// zero-padded from their rows to Mx, written into the work array. LDS holds [column][x position (padded)].
#include "passes.hpp"

namespace rt {

// EMB: the rows are the lower half of an embedding (the upper half of the period zero), as in the Toeplitz product;
// false for rows that fill the period (a circulant on the face grid).
template <typename TC, int LOG2M, int BITS, int MAXT, bool EMB = true>
__global__ void __launch_bounds__(MAXT) pass_a_k(const c128 *__restrict__ in, cplx<TC> *__restrict__ work,
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
    const int64_t off = static_cast<int64_t>(r) * d.n + (comp ? d.nx_count : 0);
    {   // Round 0, threads along x: reads the rooftop rows of column c. Every thread loads from a valid row of a valid
        // column (the last one again past the end) and zeroes what lies outside: a load under a branch would be
        // waited for inside it, one at a time.
        const int c = threadIdx.x / P::T, t = threadIdx.x - c * P::T;
        const bool live = c < ncols;
        const int64_t col = static_cast<int64_t>(j0 + (live ? c : 0)) * rows;
        const c128 *src = in + off + col;
        cplx<TC> *lds = buf + c * CS;
        auto store = [&](int, int pp, cplx<TC> v) { lds[pp] = v; };
        if (exists) {   // a masked plate: the mask bytes are loaded beside the values
            const uint8_t *exc = exists + r * d.mask_stride + (comp ? d.nx_count : 0) + col;
            cround<TC, P, 0, false, EMB, (P::K > 1)>(
                t, tw,
                [&](int pos, int) {
                    const int pc = min(pos, rows - 1);
                    const cplx<TC> v = cvt<TC>(src[pc]);
                    return live && pos < rows && exc[pc] ? v : mk<TC>(0, 0);
                },
                store);
        } else {
            cround<TC, P, 0, false, EMB, (P::K > 1)>(
                t, tw,
                [&](int pos, int) {
                    const cplx<TC> v = cvt<TC>(src[min(pos, rows - 1)]);
                    return live && pos < rows ? v : mk<TC>(0, 0);
                },
                store);
        }
    }
    __syncthreads();
    // Rounds 1 .. K-1, threads along the columns; the last one writes the work array.
    const int c = threadIdx.x % G, t = threadIdx.x / G;
    cplx<TC> *lds = buf + c * CS;
    static_for<1, P::K - 1>([&](auto S) {
        cround<TC, P, decltype(S)::value, false, false, true>(
            t, tw, [&](int, int pp) { return lds[pp]; }, [&](int, int pp, cplx<TC> v) { lds[pp] = v; });
        __syncthreads();
    });
    if (c >= ncols) return;
    cplx<TC> *dst = work + work_index(d, r, comp, 0, j0 + c);
    const int64_t ldj = d.ldj;
    if constexpr (P::K > 1) {
        cround<TC, P, P::K - 1, false, false, false>(
            t, tw, [&](int, int pp) { return lds[pp]; }, [&](int pos, int, cplx<TC> v) { dst[pos * ldj] = v; });
    } else {
        for (int pos = t; pos < P::M; pos += P::T) dst[pos * ldj] = lds[P::pad(pos)];
    }
}

// Pass A with the columns in turn (long columns): one team of T threads transforms the G columns of its group one
// after another through one column of LDS, holds each column's last-round outputs in registers (E per thread) and
// writes the G columns' rows together: rows of G elements, as the side-by-side form writes them, from one column of
// LDS instead of G (8192-point complex64 columns: four columns' 32-byte rows from 66 KB).
template <typename TC, int LOG2M, int BITS, int MAXT, int G>
__global__ void __launch_bounds__(MAXT) pass_a1_k(const c128 *__restrict__ in, cplx<TC> *__restrict__ work,
                                                  const uint8_t *__restrict__ exists, ProdDims d,
                                                  const cplx<TC> *__restrict__ tw, int nblocks,
                                                  const int *__restrict__ skip, int skip_stride) {
    using P = CPlan<LOG2M, BITS>;
    static_assert(P::K >= 2, "columns in turn need at least two rounds");
    constexpr int R = 1 << P::lr(P::K - 1), U = P::E / R;   // the last round: radix R, U butterflies per thread
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
    const int64_t off = static_cast<int64_t>(r) * d.n + (comp ? d.nx_count : 0);
    const int t = threadIdx.x;
    cplx<TC> held[G][U][R];
#pragma unroll
    for (int c = 0; c < G; ++c) {
        if (c >= ncols) break;   // uniform over the block
        const int64_t col = static_cast<int64_t>(j0 + c) * rows;
        const c128 *src = in + off + col;
        auto store = [&](int, int pp, cplx<TC> v) { buf[pp] = v; };
        if (exists) {   // a masked plate: the mask bytes are loaded beside the values
            const uint8_t *exc = exists + r * d.mask_stride + (comp ? d.nx_count : 0) + col;
            cround<TC, P, 0, false, true, true>(
                t, tw,
                [&](int pos, int) {
                    const int pc = min(pos, rows - 1);
                    const cplx<TC> v = cvt<TC>(src[pc]);
                    return pos < rows && exc[pc] ? v : mk<TC>(0, 0);
                },
                store);
        } else {
            cround<TC, P, 0, false, true, true>(
                t, tw,
                [&](int pos, int) {
                    const cplx<TC> v = cvt<TC>(src[min(pos, rows - 1)]);
                    return pos < rows ? v : mk<TC>(0, 0);
                },
                store);
        }
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
            for (int q = 0; q < R; ++q) held[c][u][q] = buf[pb + P::pad(q)];
            dft<TC, R, false, false>(held[c][u]);
        }
        __syncthreads();   // buf is free for the next column
    }
    // Rows of the work array: position (t + u T) R + q of the G columns, adjacent.
    cplx<TC> *dst = work + work_index(d, r, comp, 0, j0);
    const int64_t ldj = d.ldj;
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int q = 0; q < R; ++q) {
            cplx<TC> *row = dst + static_cast<int64_t>((t + u * P::T) * R + q) * ldj;
#pragma unroll
            for (int c = 0; c < G; ++c)
                if (c < ncols) row[c] = held[c][u][q];
        }
}

template <typename TC>
void launch_pass_a(const Product &op, const PlanTables &t, const c128 *in, cplx<TC> *work, int nrhs, int g,
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
#define RT_LAUNCH_A1(L, B)                                                                                  \
    {                                                                                                       \
        constexpr int MT = CPlan<L, B>::T > 64 ? CPlan<L, B>::T : 64;                                       \
        if constexpr (CPlan<L, B>::K >= 2) {                                                                \
            if (g == -2) {                                                                                  \
                set_lds(reinterpret_cast<const void *>(&pass_a1_k<TC, L, B, MT, 2>), lds);                  \
                launch_timed(timer, pass_a1_k<TC, L, B, MT, 2>, dim3(xcd_grid(nb)), dim3(team_threads(p)),  \
                             static_cast<uint32_t>(lds), s, in, work, op.problem->exists.ptr, op.dims, tw,  \
                             static_cast<int>(nb), skip, skip_stride);                                      \
            } else {                                                                                        \
                set_lds(reinterpret_cast<const void *>(&pass_a1_k<TC, L, B, MT, 4>), lds);                  \
                launch_timed(timer, pass_a1_k<TC, L, B, MT, 4>, dim3(xcd_grid(nb)), dim3(team_threads(p)),  \
                             static_cast<uint32_t>(lds), s, in, work, op.problem->exists.ptr, op.dims, tw,  \
                             static_cast<int>(nb), skip, skip_stride);                                      \
            }                                                                                               \
        }                                                                                                   \
    }
        RT_BY_PLAN(p.log2M, p.bits, RT_LAUNCH_A1);
#undef RT_LAUNCH_A1
        return;
    }
    const int64_t nb = blocks_ac(op.dims, nrhs, g);
    const int cs = column_stride(p, g, es);
    const size_t lds = lds_ac_bytes(p, g, es);
    const int threads = team_threads(p) * g;
#define RT_LAUNCH_A(L, B)                                                                                   \
    {                                                                                                       \
        constexpr int MT = (es == 16 && CPlan<L, B>::log2E >= 4) ? 512 : 1024;                              \
        constexpr int MF = (es == 16 && CPlan<L, 3>::log2E >= 4) ? 512 : 1024;                              \
        if (op.embedding) {                                                                                 \
            set_lds(reinterpret_cast<const void *>(&pass_a_k<TC, L, B, MT>), lds);                          \
            launch_timed(timer, pass_a_k<TC, L, B, MT>, dim3(xcd_grid(nb)), dim3(threads),                  \
                         static_cast<uint32_t>(lds), s, in, work, op.problem->exists.ptr, op.dims, tw, g, cs, \
                         static_cast<int>(nb), skip, skip_stride);                                          \
        } else {                                                                                            \
            set_lds(reinterpret_cast<const void *>(&pass_a_k<TC, L, 3, MF, false>), lds);                   \
            launch_timed(timer, pass_a_k<TC, L, 3, MF, false>, dim3(xcd_grid(nb)), dim3(threads),           \
                         static_cast<uint32_t>(lds), s, in, work, op.problem->exists.ptr, op.dims, tw, g, cs, \
                         static_cast<int>(nb), skip, skip_stride);                                          \
        }                                                                                                   \
    }
    RT_BY_PLAN(p.log2M, p.bits, RT_LAUNCH_A);
#undef RT_LAUNCH_A
}

}  // namespace rt

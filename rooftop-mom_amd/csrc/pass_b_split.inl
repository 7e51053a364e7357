// Pass B in three launches, for y transforms whose two columns do not fit the LDS of one block together (complex128
// at 8192-point columns): B1 forward y transform of each work column into a full-period spectrum buffer, B2 the
// three-plane multiply of the two components position by position, B3 inverse y transform back into the work column
// (cut). One column per block; the spectrum buffer is [r][component][x position][My].
#include "passes.hpp"

namespace rt {

template <typename TC, int LOG2M, int BITS, int MAXT>
__global__ void __launch_bounds__(MAXT) pass_by_fwd_k(const cplx<TC> *__restrict__ work, cplx<TC> *__restrict__ spec,
                                                      ProdDims d, const cplx<TC> *__restrict__ tw, int nblocks,
                                                      const int *__restrict__ skip, int skip_stride) {
    using P = CPlan<LOG2M, BITS>;
    const int lb = xcd_block(blockIdx.x, nblocks);
    if (lb < 0) return;
    const int px = lb % d.Mx, comp = (lb / d.Mx) & 1, r = lb / (2 * d.Mx);
    if (skip && skip[r * skip_stride]) return;
    extern __shared__ unsigned char smem_raw[];
    cplx<TC> *col = reinterpret_cast<cplx<TC> *>(smem_raw);
    const int t = threadIdx.x;
    const int cols = comp ? d.cols_y : d.cols_x;
    const cplx<TC> *w = work + work_index(d, r, comp, px, 0);
    cplx<TC> *out = spec + ((static_cast<int64_t>(r) * 2 + comp) * d.Mx + px) * d.My;
    if constexpr (P::K == 1) {
        cround<TC, P, 0, false, true, false>(
            t, tw, [&](int pos, int) { return clamped_load(w, pos, cols); },
            [&](int pos, int, cplx<TC> v) { out[pos] = v; });
    } else {
        cround<TC, P, 0, false, true, true>(
            t, tw, [&](int pos, int) { return clamped_load(w, pos, cols); },
            [&](int, int pp, cplx<TC> v) { col[pp] = v; });
        __syncthreads();
        static_for<1, P::K - 1>([&](auto S) {
            cround<TC, P, decltype(S)::value, false, false, true>(
                t, tw, [&](int, int pp) { return col[pp]; }, [&](int, int pp, cplx<TC> v) { col[pp] = v; });
            __syncthreads();
        });
        cround<TC, P, P::K - 1, false, false, false>(
            t, tw, [&](int, int pp) { return col[pp]; }, [&](int pos, int, cplx<TC> v) { out[pos] = v; });
    }
}

template <typename TC, int LOG2M, int BITS, int MAXT>
__global__ void __launch_bounds__(MAXT) pass_by_inv_k(cplx<TC> *__restrict__ work, const cplx<TC> *__restrict__ spec,
                                                      ProdDims d, const cplx<TC> *__restrict__ tw, int nblocks,
                                                      const int *__restrict__ skip, int skip_stride) {
    using P = CPlan<LOG2M, BITS>;
    const int lb = xcd_block(blockIdx.x, nblocks);
    if (lb < 0) return;
    const int px = lb % d.Mx, comp = (lb / d.Mx) & 1, r = lb / (2 * d.Mx);
    if (skip && skip[r * skip_stride]) return;
    extern __shared__ unsigned char smem_raw[];
    cplx<TC> *col = reinterpret_cast<cplx<TC> *>(smem_raw);
    const int t = threadIdx.x;
    const int cols = comp ? d.cols_y : d.cols_x;
    cplx<TC> *w = work + work_index(d, r, comp, px, 0);
    const cplx<TC> *in = spec + ((static_cast<int64_t>(r) * 2 + comp) * d.Mx + px) * d.My;
    if constexpr (P::K == 1) {
        cround<TC, P, 0, true, true, false>(
            t, tw, [&](int pos, int) { return in[pos]; },
            [&](int pos, int, cplx<TC> v) {
                if (pos < cols) w[pos] = v;
            });
    } else {
        cround<TC, P, P::K - 1, true, false, false>(
            t, tw, [&](int pos, int) { return in[pos]; }, [&](int, int pp, cplx<TC> v) { col[pp] = v; });
        __syncthreads();
        static_for_down<1, P::K - 1>([&](auto S) {
            cround<TC, P, decltype(S)::value, true, false, true>(
                t, tw, [&](int, int pp) { return col[pp]; }, [&](int, int pp, cplx<TC> v) { col[pp] = v; });
            __syncthreads();
        });
        cround<TC, P, 0, true, true, true>(
            t, tw, [&](int, int pp) { return col[pp]; },
            [&](int pos, int, cplx<TC> v) {
                if (pos < cols) w[pos] = v;
            });
    }
}

// The three-plane multiply at every (x position, y position) of every right-hand side (the spectra of pass B).
template <typename TC>
__global__ void __launch_bounds__(256) pass_by_mul_k(cplx<TC> *__restrict__ spec, const cplx<TC> *__restrict__ bspec,
                                                     const int *__restrict__ fold_x, const int *__restrict__ border,
                                                     const cplx<TC> *__restrict__ dx,
                                                     const cplx<TC> *__restrict__ dy, int Mx, int My, int R,
                                                     const int *__restrict__ skip, int skip_stride) {
    const int64_t plane = static_cast<int64_t>(Mx) * My;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < plane * R;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t r = e / plane;
        if (skip && skip[r * skip_stride]) continue;
        const int64_t p = e - r * plane;
        const int q = static_cast<int>(p / My), py = static_cast<int>(p - static_cast<int64_t>(q) * My);
        const int px = border[q];   // partners k and M - k run together and share their spectrum row
        cplx<TC> *X = spec + (r * 2) * plane + static_cast<int64_t>(px) * My + py;
        cplx<TC> *Y = X + plane;
        const cplx<TC> *sp = bspec + (static_cast<int64_t>(fold_x[px]) * My + py) * 3;
        const cplx<TC> dxv = dx[px], dyv = dy[py];
        const cplx<TC> x = *X, y = *Y;
        const cplx<TC> phi = sp[2] * (cmulc(dxv, x) + cmulc(dyv, y));
        *X = sp[0] * x - dxv * phi;
        *Y = sp[1] * y - dyv * phi;
    }
}

template <typename TC>
void launch_pass_b_split(const Product &op, const PlanTables &tx, const PlanTables &ty, cplx<TC> *work,
                         cplx<TC> *spec, const cplx<TC> *bspec, int nrhs, hipStream_t s, const int *skip,
                         int skip_stride) {
    const FftPlan &p = ty.plan;
    constexpr int es = sizeof(cplx<TC>);
    const int64_t nb = static_cast<int64_t>(2) * op.dims.Mx * nrhs;
    const size_t lds = static_cast<size_t>(padded_length(p)) * es;
    const int threads = team_threads(p);
    const bool dbl = sizeof(TC) == sizeof(double);
    const cplx<TC> *tw = dbl ? reinterpret_cast<const cplx<TC> *>(ty.tw.ptr)
                             : reinterpret_cast<const cplx<TC> *>(ty.tw32.ptr);
    const cplx<TC> *dx = dbl ? reinterpret_cast<const cplx<TC> *>(tx.d.ptr)
                             : reinterpret_cast<const cplx<TC> *>(tx.d32.ptr);
    const cplx<TC> *dy = dbl ? reinterpret_cast<const cplx<TC> *>(ty.d.ptr)
                             : reinterpret_cast<const cplx<TC> *>(ty.d32.ptr);
#define RT_LAUNCH_BY(L, B)                                                                                     \
    {                                                                                                          \
        constexpr int MT = CPlan<L, B>::T > 256 ? CPlan<L, B>::T : 256;                                        \
        set_lds(reinterpret_cast<const void *>(&pass_by_fwd_k<TC, L, B, MT>), lds);                            \
        set_lds(reinterpret_cast<const void *>(&pass_by_inv_k<TC, L, B, MT>), lds);                            \
        launch_timed(nullptr, pass_by_fwd_k<TC, L, B, MT>, dim3(xcd_grid(nb)), dim3(threads),                  \
                     static_cast<uint32_t>(lds), s, static_cast<const cplx<TC> *>(work), spec, op.dims, tw,    \
                     static_cast<int>(nb), skip, skip_stride);                                                 \
        launch_timed(nullptr, pass_by_mul_k<TC>, dim3(blocks_for(static_cast<size_t>(op.dims.Mx) * op.dims.My * \
                                                                 nrhs, 256, 65536)),                          \
                     dim3(256), 0u, s, spec, bspec, tx.fold.ptr, tx.border.ptr, dx, dy, op.dims.Mx, op.dims.My,  \
                     nrhs, skip,                                                                               \
                     skip_stride);                                                                             \
        launch_timed(nullptr, pass_by_inv_k<TC, L, B, MT>, dim3(xcd_grid(nb)), dim3(threads),                  \
                     static_cast<uint32_t>(lds), s, work, static_cast<const cplx<TC> *>(spec), op.dims, tw,    \
                     static_cast<int>(nb), skip, skip_stride);                                                 \
    }
    RT_BY_PLAN(p.log2M, p.bits, RT_LAUNCH_BY);
#undef RT_LAUNCH_BY
}

}  // namespace rt

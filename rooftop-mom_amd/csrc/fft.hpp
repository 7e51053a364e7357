// In-LDS radix-4/2 FFT for gfx950: decimation in frequency forward (natural order in, digit-reversed order out) and
// decimation in time inverse (digit-reversed in, natural out). A convolution never needs the natural frequency
// order, so no permutation pass is run: the spectra are stored in the same digit-reversed order, and per-position
// tables (frequency fold index, divergence factor) carry the true frequency where the multiply needs it.
#pragma once

#include <vector>

#include "common.hpp"

namespace rt {

constexpr int kMaxStages = 16;

// Plan of one transform length M (a power of two, or 1). Stage s has radix radix[s] and acts on blocks of
// len[s] = M / (radix[0] ... radix[s-1]) elements with the butterfly stride len[s] / radix[s].
struct FftPlan {
    int M = 1;
    int log2M = 0;
    int nstages = 0;
    int radix[kMaxStages] = {};
    int log2len[kMaxStages] = {};    // log2 of len[s]
    int log2radix[kMaxStages] = {};
};

FftPlan make_plan(int M);

// Frequency of each output position of the forward transform (host).
std::vector<int> output_frequencies(const FftPlan &p);

inline int next_pow2(long n) {
    int m = 1;
    while (m < n) m <<= 1;
    return m;
}

// Radix-4 / radix-2 butterflies. Forward: y_j = w^(j n) sum_m x_m W4^(jm), W4 = exp(-2 pi i/4). Inverse undoes one
// forward stage up to the factor of its radix: x_m = sum_j W4^(-jm) conj(w^(j n)) y_j.
template <typename T>
__device__ __forceinline__ void bfly4_fwd(cplx<T> &x0, cplx<T> &x1, cplx<T> &x2, cplx<T> &x3) {
    cplx<T> a0 = x0 + x2, a1 = x0 - x2, a2 = x1 + x3, a3 = x1 - x3;
    cplx<T> ma3 = mk<T>(a3.y, -a3.x);  // -i a3
    x0 = a0 + a2;
    x2 = a0 - a2;
    x1 = a1 + ma3;
    x3 = a1 - ma3;
}

template <typename T>
__device__ __forceinline__ void bfly4_inv(cplx<T> &y0, cplx<T> &y1, cplx<T> &y2, cplx<T> &y3) {
    cplx<T> b0 = y0 + y2, b1 = y0 - y2, b2 = y1 + y3, b3 = y1 - y3;
    cplx<T> ib3 = mk<T>(-b3.y, b3.x);  // +i b3
    y0 = b0 + b2;
    y2 = b0 - b2;
    y1 = b1 + ib3;
    y3 = b1 - ib3;
}

// One transform of `ncols` columns held in LDS: column c starts at buf + c * col_stride, elements contiguous.
// tw is the table exp(-2 pi i t / M), t < M (global memory, cached). All threads of the block must call it.
template <typename T, bool INVERSE>
__device__ void lds_fft(cplx<T> *buf, int ncols, int col_stride, const FftPlan &p, const cplx<T> *__restrict__ tw,
                        int tid, int nthr) {
    for (int ss = 0; ss < p.nstages; ++ss) {
        const int s = INVERSE ? p.nstages - 1 - ss : ss;
        const int lr = p.log2radix[s];
        const int ll = p.log2len[s];
        const int ls = ll - lr;                 // log2 of the butterfly stride S
        const int S = 1 << ls;
        const int tw_shift = p.log2M - ll;      // twiddle index = j * n * (M / len)
        const int lnb = p.log2M - lr;           // log2 of the butterflies per column
        const int total = ncols << lnb;
        for (int b = tid; b < total; b += nthr) {
            const int col = b >> lnb;
            const int bb = b & ((1 << lnb) - 1);
            const int blk = bb >> ls;
            const int n = bb & (S - 1);
            cplx<T> *base = buf + col * col_stride + (blk << ll) + n;
            if (lr == 2) {
                cplx<T> x0 = base[0], x1 = base[S], x2 = base[2 * S], x3 = base[3 * S];
                if (!INVERSE) {
                    bfly4_fwd(x0, x1, x2, x3);
                    if (n) {
                        const int t = n << tw_shift;
                        x1 = x1 * tw[t];
                        x2 = x2 * tw[2 * t];
                        x3 = x3 * tw[3 * t];
                    }
                } else {
                    if (n) {
                        const int t = n << tw_shift;
                        x1 = cmulc(tw[t], x1);
                        x2 = cmulc(tw[2 * t], x2);
                        x3 = cmulc(tw[3 * t], x3);
                    }
                    bfly4_inv(x0, x1, x2, x3);
                }
                base[0] = x0;
                base[S] = x1;
                base[2 * S] = x2;
                base[3 * S] = x3;
            } else {
                cplx<T> x0 = base[0], x1 = base[S];
                if (!INVERSE) {
                    cplx<T> d = x0 - x1;
                    x0 = x0 + x1;
                    x1 = n ? d * tw[n << tw_shift] : d;
                } else {
                    if (n) x1 = cmulc(tw[n << tw_shift], x1);
                    cplx<T> d = x0 - x1;
                    x0 = x0 + x1;
                    x1 = d;
                }
                base[0] = x0;
                base[S] = x1;
            }
        }
        __syncthreads();
    }
}

}  // namespace rt

// Team FFT for gfx950: a transform of length M = 2^m is run by a team of T = M / E threads, each holding E elements
// in registers, in K rounds of radix r_s <= E (decimation in frequency forward, its exact reverse inverse). A round
// reads the E elements of each thread, runs E / r_s in-register DFTs of length r_s with the twiddles of the round, and
// writes them back; rounds exchange data through LDS, or through global memory where a pass begins or ends, so a
// transform costs K - 1 LDS exchanges instead of one per radix-2/4 stage.
//
// The forward transform leaves its output in digit-reversed order (mixed radix) and the inverse takes that order
// back, so no permutation pass is run: the spectra are read through per-position tables (frequency fold index,
// divergence factor) that carry the true frequency. Forward round s of a block of length L = 2^ll[s] with stride
// S = L / r: for n < S, y_q = w_L^(q n) sum_p x_(n + p S) W_r^(p q), written to position n + q S; the inverse round
// multiplies by conj(w_L^(q n)) and applies the inverse length-r DFT, undoing the forward round up to the factor r.
#pragma once

#include <vector>

#include "common.hpp"

namespace rt {

constexpr int kMaxRounds = 8;

struct FftPlan {
    int M = 1;
    int log2M = 0;
    int bits = 0;                  // the radix limit 2^bits the plan was made with
    int K = 0;                     // rounds
    int log2E = 0;                 // log2 of the largest radix: elements per thread
    int lr[kMaxRounds] = {};       // log2 radix of each round
    int ll[kMaxRounds] = {};       // log2 of the block length L of each round
    int twoff[kMaxRounds] = {};    // offset of each round's twiddles w_L^(q n) at twoff + n r + q
    int twcount = 0;
};

// Plan of length M (a power of two) with radices of at most 2^maxbits, as even as possible, larger first.
FftPlan make_plan(int M, int maxbits);
// Frequency held by each output position of the forward transform (host).
std::vector<int> output_frequencies(const FftPlan &p);
// The twiddle table of the plan, twcount entries (host, long double accuracy rounded once).
std::vector<c128> plan_twiddles(const FftPlan &p);

inline int next_pow2(long n) {
    int m = 1;
    while (m < n) m <<= 1;
    return m;
}

__host__ __device__ inline int team_threads(const FftPlan &p) { return p.M >> p.log2E; }
// Padded LDS offset of a position: one spare element every E keeps the strided rounds off a single bank.
__host__ __device__ inline int padded(const FftPlan &p, int pos) { return pos + (pos >> p.log2E); }
__host__ __device__ inline int padded_length(const FftPlan &p) { return p.M + (p.M >> p.log2E); }

// ------------------------------------------------------------------------------------------------------------------
// In-register DFTs.
// ------------------------------------------------------------------------------------------------------------------
// v * (c - i s) forward, v * (c + i s) inverse.
template <typename T, bool INV>
__device__ __forceinline__ cplx<T> mulw(cplx<T> v, T c, T s) {
    if (!INV) return {v.x * c + v.y * s, v.y * c - v.x * s};
    return {v.x * c - v.y * s, v.y * c + v.x * s};
}

// v * W_16^j, W_16 = exp(-2 pi i / 16) (conjugated for the inverse), 0 <= j < 8.
template <typename T, bool INV>
__device__ __forceinline__ cplx<T> rot16(cplx<T> v, int j) {
    const T c1 = static_cast<T>(0.92387953251128675613);
    const T s1 = static_cast<T>(0.38268343236508977173);
    const T h = static_cast<T>(0.70710678118654752440);
    switch (j) {
    case 0: return v;
    case 1: return mulw<T, INV>(v, c1, s1);
    case 2: return mulw<T, INV>(v, h, h);
    case 3: return mulw<T, INV>(v, s1, c1);
    case 4: return INV ? cplx<T>{-v.y, v.x} : cplx<T>{v.y, -v.x};
    case 5: return mulw<T, INV>(v, -s1, c1);
    case 6: return mulw<T, INV>(v, -h, h);
    default: return mulw<T, INV>(v, -c1, s1);
    }
}

template <int R>
__host__ __device__ constexpr int bit_reverse(int i) {
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r = (r << 1) | (i & 1);
        i >>= 1;
    }
    return r;
}

// y_q = sum_p x_p W_R^(+-p q), natural order in and out, R <= 16 (radix-2 levels, decimation in frequency).
// HALF: x_p = 0 for p >= R / 2 (the zero-padded half of an embedding), which makes the first level a copy.
template <typename T, int R, bool INV, bool HALF>
__device__ __forceinline__ void dft(cplx<T> *x) {
    if constexpr (R > 1) {
#pragma unroll
        for (int h = R / 2; h >= 1; h >>= 1) {
#pragma unroll
            for (int g = 0; g < R; g += 2 * h) {
#pragma unroll
                for (int k = 0; k < h; ++k) {
                    const int j = k * (16 / (2 * h));
                    const cplx<T> a = x[g + k], b = x[g + k + h];
                    if (HALF && h == R / 2) {
                        x[g + k] = a;
                        x[g + k + h] = rot16<T, INV>(a, j);
                    } else {
                        x[g + k] = a + b;
                        x[g + k + h] = rot16<T, INV>(a - b, j);
                    }
                }
            }
        }
        cplx<T> y[R];
#pragma unroll
        for (int i = 0; i < R; ++i) y[bit_reverse<R>(i)] = x[i];
#pragma unroll
        for (int i = 0; i < R; ++i) x[i] = y[i];
    }
}

// ------------------------------------------------------------------------------------------------------------------
// One round of a team. Thread t holds the butterflies beta = t + u T (u < E / r); butterfly beta of round s covers the
// positions base + q S, base = (beta / S) L + beta mod S. load(pos) and store(pos, v) move elements; HALF skips the
// loads of the zero half (first forward round of an embedding); twid = false skips the twiddles (the last round,
// S = 1, has none).
// ------------------------------------------------------------------------------------------------------------------
template <typename T, int E, int R, bool INV, bool HALF, typename Load, typename Store>
__device__ __forceinline__ void team_round_r(const FftPlan &p, int s, int t, const cplx<T> *__restrict__ tw,
                                             bool twid, Load &&load, Store &&store) {
    constexpr int U = E / R;
    const int T_ = p.M >> p.log2E;
    const int ls = p.ll[s] - p.lr[s];
    const cplx<T> *tws = tw + p.twoff[s];
    // One butterfly at a time: its loads, DFT and stores are independent of the next one's, so the register budget
    // of a thread is one butterfly's (the waves of a CU hide the latency instead).
#pragma unroll 1
    for (int u = 0; u < U; ++u) {
        const int beta = t + u * T_;
        const int n = beta & ((1 << ls) - 1);
        const int base = ((beta >> ls) << p.ll[s]) + n;
        cplx<T> v[R];
#pragma unroll
        for (int q = 0; q < R; ++q)
            v[q] = (HALF && R > 1 && q >= R / 2) ? mk<T>(0, 0) : load(base + (q << ls));
        if (!INV) {
            dft<T, R, false, HALF>(v);
            if (twid) {
#pragma unroll
                for (int q = 1; q < R; ++q) v[q] = v[q] * tws[n * R + q];
            }
        } else {
            if (twid) {
#pragma unroll
                for (int q = 1; q < R; ++q) v[q] = cmulc(tws[n * R + q], v[q]);
            }
            dft<T, R, true, false>(v);
        }
#pragma unroll
        for (int q = 0; q < R; ++q) store(base + (q << ls), v[q]);
    }
}

template <typename T, int E, bool INV, bool HALF, typename Load, typename Store>
__device__ __forceinline__ void team_round(const FftPlan &p, int s, int t, const cplx<T> *__restrict__ tw, bool twid,
                                           Load &&load, Store &&store) {
    switch (p.lr[s]) {
    case 0: team_round_r<T, E, 1, INV, HALF>(p, s, t, tw, twid, load, store); break;
    case 1:
        if constexpr (E >= 2) team_round_r<T, E, 2, INV, HALF>(p, s, t, tw, twid, load, store);
        break;
    case 2:
        if constexpr (E >= 4) team_round_r<T, E, 4, INV, HALF>(p, s, t, tw, twid, load, store);
        break;
    case 3:
        if constexpr (E >= 8) team_round_r<T, E, 8, INV, HALF>(p, s, t, tw, twid, load, store);
        break;
    default:
        if constexpr (E >= 16) team_round_r<T, E, 16, INV, HALF>(p, s, t, tw, twid, load, store);
        break;
    }
}

// A whole transform of one column held in LDS at buf[padded(pos)], run by the T threads t < T of the block.
template <typename T, int E, bool INV>
__device__ void team_fft_lds(cplx<T> *buf, const FftPlan &p, const cplx<T> *__restrict__ tw, int t) {
    const bool member = t < team_threads(p);
    for (int k = 0; k < p.K; ++k) {
        const int s = INV ? p.K - 1 - k : k;
        if (member)
            team_round<T, E, INV, false>(
                p, s, t, tw, s < p.K - 1, [&](int pos) { return buf[padded(p, pos)]; },
                [&](int pos, cplx<T> v) { buf[padded(p, pos)] = v; });
        __syncthreads();
    }
}

}  // namespace rt

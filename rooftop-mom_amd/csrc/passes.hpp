// The three matvec passes of the Toeplitz product (product.hip), specialised per transform plan at compile time so
// that every position offset of a round is an immediate and each kernel carries only its own radix path.
//
//   A  forward x transform of `group` rooftop columns of one component: the first round reads the rooftop vector
//      (threads along x, coalesced), the last round writes the work array with the threads along the columns, so
//      the transposition costs no pass of its own;
//   B  per x position: forward y transforms of both components, the three-plane multiply in registers between the
//      last forward round and the first inverse round, inverse y transforms; contiguous work columns;
//   C  inverse x transform of `group` columns, the mirror of A, cut to the rooftop rows.
// Blocks are numbered XCD-major (blockIdx b runs on XCD b mod 8): consecutive logical blocks -- adjacent column
// groups in A/C, the right-hand sides of one x position in B -- share an L2, so their partial cache lines of the
// work array and their spectrum rows meet there.
#pragma once

#include <type_traits>

#include "rooftop.hpp"

namespace rt {

// a[pos] for pos < count, else zero, as a load from a valid element (the last one again) and a select: a load under
// a branch is waited for inside it, so the loads of a round would go out one at a time.
template <typename T>
__device__ __forceinline__ cplx<T> clamped_load(const cplx<T> *__restrict__ a, int pos, int count) {
    const cplx<T> v = a[min(pos, max(count - 1, 0))];
    return pos < count ? v : cplx<T>{0, 0};
}

// ------------------------------------------------------------------------------------------------------------------
// Compile-time plan: the same schedule as make_plan(1 << LOG2M, BITS) (fft_plan.cpp), whose tables it reads.
// ------------------------------------------------------------------------------------------------------------------
template <int LOG2M, int BITS>
struct CPlan {
    static constexpr int m = LOG2M;
    static constexpr int M = 1 << m;
    static constexpr int K = m == 0 ? 1 : (m + BITS - 1) / BITS;
    static constexpr int base_bits = m == 0 ? 0 : m / K;
    static constexpr int extra = m == 0 ? 0 : m % K;
    static constexpr int lr(int s) { return m == 0 ? 0 : base_bits + (s < extra ? 1 : 0); }
    static constexpr int ll(int s) {
        int l = m;
        for (int i = 0; i < s; ++i) l -= lr(i);
        return l;
    }
    static constexpr int twoff(int s) {
        int o = 0;
        for (int i = 0; i < s; ++i) o += 1 << ll(i);
        return o;
    }
    static constexpr int log2E = m == 0 ? 0 : base_bits + (extra > 0 ? 1 : 0);
    static constexpr int E = 1 << log2E;
    static constexpr int T = M >> log2E;   // threads of a team
    // LDS offset of position p: one element of padding after every E.
    static constexpr int pad(int p) { return p + (p >> log2E); }
};

// LDS layout of pass B's columns: position p at p + c1 (p >> a1) + c2 (p >> a2). Chosen per plan and element size
// from a model of the 64 LDS banks (8-byte accesses served 16 lanes at a time, 16-byte ones 8 at a time) over every
// round's access pattern: one element of padding per 128 bytes keeps both the unit-stride rounds and the stride-R
// last round conflict-free for most plans; 1024-point complex64 columns with radices up to 8 (rounds 8, 8, 4, 4) take
// two per 128 bytes and one more per 256.
struct BPad {
    int a1, c1, a2, c2;
};
__host__ __device__ constexpr BPad b_pad_rule(int m, int bits, int es) {
    if (m == 10 && bits == 3 && es == 8) return BPad{4, 2, 5, 1};
    return BPad{es == 8 ? 4 : 3, 1, 0, 0};
}
__host__ __device__ constexpr int b_pad(const BPad &r, int p) { return p + r.c1 * (p >> r.a1) + r.c2 * (p >> r.a2); }

template <int LOG2M, int BITS, int ES>
struct BPlan : CPlan<LOG2M, BITS> {
    // The rule's fields as immediates: b_pad(rule, p) takes the rule by reference, which makes device code read it
    // from memory and multiply at run time, so every LDS offset would need an address register of its own.
    static constexpr int A1 = b_pad_rule(LOG2M, BITS, ES).a1, C1 = b_pad_rule(LOG2M, BITS, ES).c1;
    static constexpr int A2 = b_pad_rule(LOG2M, BITS, ES).a2, C2 = b_pad_rule(LOG2M, BITS, ES).c2;
    static constexpr int pad(int p) { return p + C1 * (p >> A1) + C2 * (p >> A2); }
};

template <int I>
using ic = std::integral_constant<int, I>;

// The same pointer, opaque to the compiler: loads through it are not merged with earlier loads of the same
// addresses. The inverse rounds of pass B read the forward rounds' twiddles again through it instead of keeping
// them in registers across the multiply, which would cost occupancy.
// The pointer goes through the asm as a global-memory pointer, so the loads stay global loads (a flat load would
// also count against the LDS wait counter).
template <typename T>
__device__ __forceinline__ const T *opaque(const T *p) {
    using G = const T __attribute__((address_space(1))) *;
    G g = (G)p;
    asm volatile("" : "+s"(g));
    return (const T *)g;
}

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(ic<B>{});
        static_for<B + 1, E>(f);
    }
}
// E-1 down to B.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for_down(F &&f) {
    if constexpr (B < E) {
        f(ic<E - 1>{});
        static_for_down<B, E - 1>(f);
    }
}

// The twiddles w^q, q = 1 .. R - 1, of one butterfly from its table row (row[q] = w^q): complex128 reads w^1 and w^4
// and forms the others by products (each the product of at most four table entries, a few units in the last place);
// complex64 reads w^1, w^2, w^4 and w^8 and forms the others by one or two products. The twiddle reads were the larger
// share of the passes' vector-memory requests (one per element and round, against one data read per element and
// pass); a product costs four VALU instructions.
template <typename T, int R>
__device__ __forceinline__ void twiddle_row(const cplx<T> *__restrict__ row, cplx<T> *t) {
    if constexpr (sizeof(T) == sizeof(double) && R >= 4) {
        t[1] = row[1];
        t[2] = t[1] * t[1];
        t[3] = t[2] * t[1];
        if constexpr (R >= 8) {
            t[4] = row[4];
            t[5] = t[4] * t[1];
            t[6] = t[4] * t[2];
            t[7] = t[4] * t[3];
        }
        if constexpr (R >= 16) {
            t[8] = t[4] * t[4];
#pragma unroll
            for (int q = 9; q < 16; ++q) t[q] = t[8] * t[q - 8];
        }
    } else if constexpr (R >= 4) {   // complex64: w^1, w^2, w^4 and w^8 read, the others one product each
        t[1] = row[1];
        t[2] = row[2];
        t[3] = t[2] * t[1];
        if constexpr (R >= 8) {
            t[4] = row[4];
            t[5] = t[4] * t[1];
            t[6] = t[4] * t[2];
            t[7] = t[4] * t[3];
        }
        if constexpr (R >= 16) {
            t[8] = row[8];
#pragma unroll
            for (int q = 9; q < 16; ++q) t[q] = t[8] * t[q - 8];
        }
    } else {
#pragma unroll
        for (int q = 1; q < R; ++q) t[q] = row[q];
    }
}

// One round S of a team with a compile-time plan. Thread t holds butterflies beta = t + u T (u < E / r); butterfly
// beta covers positions base + q Str (Str = L / r). load(pos, ppos) / store(pos, ppos, v) receive the position and
// its padded LDS offset, both a per-butterfly value plus an immediate.
// HALF (forward): the upper half of the inputs is zero (the padding of an embedding); HALF (inverse): the upper half
// of the outputs is cut away, so it is neither stored nor computed.
template <typename T, class P, int S, bool INV, bool HALF, bool TWID, typename Load, typename Store>
__device__ __forceinline__ void cround(int t, const cplx<T> *__restrict__ tw, Load &&load, Store &&store) {
    constexpr int lr = P::lr(S), ll = P::ll(S), R = 1 << lr, U = P::E / R, ls = ll - lr;
    const cplx<T> *tws = tw + P::twoff(S);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int beta = t + u * P::T;
        const int n = beta & ((1 << ls) - 1);
        const int base = ((beta >> ls) << ll) + n;
        const int pb = P::pad(base);   // base and the offsets o below have disjoint bits: pad(base + o) = pb + pad(o)
        cplx<T> v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int o = q << ls;
            const int po = P::pad(o);
            v[q] = (!INV && HALF && R > 1 && q >= R / 2) ? mk<T>(0, 0) : load(base + o, pb + po);
        }
        if constexpr (!INV) {
            dft<T, R, false, HALF>(v);
            if constexpr (TWID) {
                cplx<T> w[R];
                twiddle_row<T, R>(tws + n * R, w);
#pragma unroll
                for (int q = 1; q < R; ++q) v[q] = v[q] * w[q];
            }
        } else {
            if constexpr (TWID) {
                cplx<T> w[R];
                twiddle_row<T, R>(tws + n * R, w);
#pragma unroll
                for (int q = 1; q < R; ++q) v[q] = cmulc(w[q], v[q]);
            }
            dft<T, R, true, false>(v);
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
            if (INV && HALF && R > 1 && q >= R / 2) continue;
            const int o = q << ls;
            const int po = P::pad(o);
            store(base + o, pb + po, v[q]);
        }
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Shared helpers.
// ------------------------------------------------------------------------------------------------------------------
// Work array of the passes, per right-hand side: [component][x position (plan order)][y column], the y columns of
// one x position contiguous (row length ldj >= the columns of either component).
__host__ __device__ inline int64_t work_index(const ProdDims &d, int r, int comp, int kx, int j) {
    return ((static_cast<int64_t>(r) * 2 + comp) * d.Mx + kx) * d.ldj + j;
}

// Logical block of hardware block b in a grid of 8 * ceil(nblocks / 8): XCD b mod 8 runs the logical blocks
// [(b mod 8) per, (b mod 8 + 1) per) in order. -1 for the padding blocks.
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
    const int per = (nblocks + 7) >> 3;
    const int lb = (b & 7) * per + (b >> 3);
    return lb < nblocks ? lb : -1;
}
inline unsigned xcd_grid(int64_t nblocks) { return static_cast<unsigned>((nblocks + 7) / 8 * 8); }

// Decodes a pass-A/C logical block into (right-hand side, component, first rooftop column).
__device__ inline void group_of(const ProdDims &d, int group, int blk, int &r, int &comp, int &j0) {
    const int gx = (d.cols_x + group - 1) / group, gy = (d.cols_y + group - 1) / group;
    const int per = gx + gy;
    r = blk / per;
    const int rem = blk - r * per;
    comp = rem < gx ? 0 : 1;
    j0 = (comp ? rem - gx : rem) * group;
}

inline int64_t blocks_ac(const ProdDims &d, int nrhs, int g) {
    if (g < 0) g = -g;   // columns in turn: -g columns per block
    return static_cast<int64_t>(nrhs) * ((d.cols_x + g - 1) / g + (d.cols_y + g - 1) / g);
}

// Block-size bound of each pass (threads = team size x columns or x positions per block): pass A/C at most 1024
// (512 for complex128 with radix 16, whose butterflies need more registers), pass B one team of at least 256.
inline int max_threads_ac(int es, const FftPlan &p) { return es == 16 && p.log2E >= 4 ? 512 : 1024; }
inline int max_threads_b(const FftPlan &p) { return team_threads(p) > 256 ? team_threads(p) : 256; }
// Whether pass B can run `tile` x positions per block in form `teams`: 1, one team per position with both columns in
// LDS; 2, a team per column (two rounds and whole waves per team, at most 1024 threads); 3, one team per position
// whose columns take turns in one column of LDS (at least two rounds).
inline bool b_threads_fit(const FftPlan &p, int tile, int teams) {
    const int T = team_threads(p);
    if (teams == 1) return T * tile <= max_threads_b(p);
    if (teams == 3) return p.K >= 2 && T * tile <= max_threads_b(p);
    const int per = 2 * T;
    return teams == 2 && p.K >= 2 && T % 64 == 0 && per <= 1024 && per * tile <= (per > 256 ? per : 256);
}

// Pass A/C LDS: [column][position (padded)] with a column stride that puts the columns of one position on
// different banks (64 banks of 4 bytes).
inline int column_stride(const FftPlan &p, int group, int es) {
    const int bank_elems = 256 / es;
    const int mp = padded_length(p);
    const int up = (mp + bank_elems - 1) / bank_elems * bank_elems;
    const int skew = bank_elems / group > 0 ? bank_elems / group : 1;
    return up + skew;
}
// group < 0: -group columns in turn through one column of LDS (pass_a1_k, pass_c1_k).
inline size_t lds_ac_bytes(const FftPlan &p, int group, int es) {
    if (group < 0) return static_cast<size_t>(padded_length(p)) * es;
    return static_cast<size_t>(column_stride(p, group, es)) * group * es;
}
// Whether pass A/C can run `group` columns per block: side by side (group > 0, a team per column) or in turn (-2, -4:
// one team, at least two rounds).
inline bool ac_threads_fit(const FftPlan &p, int group, int es) {
    if (group < 0) return (group == -2 || group == -4) && p.K >= 2 && team_threads(p) <= max_threads_ac(es, p);
    return group >= 1 && team_threads(p) * group <= max_threads_ac(es, p);
}
// Padded length of a pass-B column (the BPlan layout).
inline int b_padded_length(const FftPlan &p, int es) { return b_pad(b_pad_rule(p.log2M, p.bits, es), p.M - 1) + 1; }
inline size_t lds_b_bytes(const FftPlan &p, int tile, int es, int teams = 1) {
    return static_cast<size_t>(teams == 3 ? 1 : 2) * tile * b_padded_length(p, es) * es;
}
// Pass B in three launches (pass_b_split.inl) holds one column per block.
inline size_t lds_b_split_bytes(const FftPlan &p, int es) { return static_cast<size_t>(padded_length(p)) * es; }

// Raises a kernel's dynamic LDS limit above the 64 KiB default (once per kernel and size); fails beyond the device.
void set_lds(const void *kernel, size_t bytes);
size_t lds_bytes_max();

// Launchers (passes_a.hip, passes_b.hip, passes_c.hip, one translation unit per pass and precision).
template <typename TC>
void launch_pass_a(const Product &op, const PlanTables &t, const c128 *in, cplx<TC> *work, int nrhs, int group,
                   hipStream_t s, const int *skip, int skip_stride, const KernelTimer *timer);
template <typename TC>
void launch_pass_b(const Product &op, const PlanTables &tx, const PlanTables &ty, cplx<TC> *work,
                   const cplx<TC> *bspec, int nrhs, int tile, int teams, hipStream_t s, const int *skip,
                   int skip_stride, const KernelTimer *timer);
template <typename TC>
void launch_pass_b_split(const Product &op, const PlanTables &tx, const PlanTables &ty, cplx<TC> *work,
                         cplx<TC> *spec, const cplx<TC> *bspec, int nrhs, hipStream_t s, const int *skip,
                         int skip_stride);
template <typename TC>
void launch_pass_c(const Product &op, const PlanTables &t, const cplx<TC> *work, c128 *out, int nrhs, int group,
                   hipStream_t s, const int *skip, int skip_stride, const KernelTimer *timer);

// Calls F with the compile-time plan of (log2M, bits).
#define RT_PLAN_CASE(L, F)                                      \
    case L:                                                     \
        if (bits >= 4) { F(L, 4) } else { F(L, 3) }             \
        break;
#define RT_BY_PLAN(LOG2M, BITS, F)                                                                   \
    do {                                                                                             \
        const int bits = (BITS);                                                                     \
        switch (LOG2M) {                                                                             \
            RT_PLAN_CASE(0, F) RT_PLAN_CASE(1, F) RT_PLAN_CASE(2, F) RT_PLAN_CASE(3, F)              \
            RT_PLAN_CASE(4, F) RT_PLAN_CASE(5, F) RT_PLAN_CASE(6, F) RT_PLAN_CASE(7, F)              \
            RT_PLAN_CASE(8, F) RT_PLAN_CASE(9, F) RT_PLAN_CASE(10, F) RT_PLAN_CASE(11, F)            \
            RT_PLAN_CASE(12, F) RT_PLAN_CASE(13, F)                                                  \
        default: fail(RT_EINVAL, "rooftop:product", "transform length 2^%d is not supported", LOG2M); \
        }                                                                                            \
    } while (0)

}  // namespace rt

// Preconditioners on gfx950: the Jacobi diagonal and the near-field preconditioner on 3x3-cell neighbourhoods.
//
// The reference's preconditioners are circulant ones (src/+rooftopMoM/Preconditioner.m:41-51) and it passes Jacobi as
// a handle (tests/+unit/+solver/ConjugateGradientTest.m:55-59). The near-field preconditioner is this project's:
// a symmetrized sparse approximate inverse P = (M + M^T)/2 whose row i is the least-squares fit
//     min || Z[F_i, S_i] m - e_i ||,
// S_i the rooftops with their first face within one cell of that of i (the 3x3-cell neighbourhood, <= 18), F_i those
// within two cells (<= 50). Entries of Z come from the generators through the spatial-block formulas of
// +preconditioner/+utils/generateSpatialBlocks.m:19-61. Z depends on offsets only, so every rooftop whose
// neighbourhood has the same existing rooftops shares its row: rows are fitted once per class, and P is applied as
// an 18-point stencil over the two rooftop grids with the class coefficients.
#include <algorithm>
#include <vector>

#include "nearfield_classes.hpp"
#include "rooftop.hpp"

namespace rt {

constexpr int kSlots = kNearFieldSlots;   // 2 directions x 3 x 3 first-face offsets
constexpr int kFit = 50;      // 2 directions x 5 x 5

// Generator value at absolute offsets (the generators are even in each offset); zero outside.
__device__ __forceinline__ c128 gen_at(const c128 *g, int rows, int cols, int p, int q) {
    p = p < 0 ? -p : p;
    q = q < 0 ? -q : q;
    if (!g || p >= rows || q >= cols) return mk(0.0, 0.0);
    return g[p + static_cast<int64_t>(rows) * q];
}

struct GenView {
    const c128 *mxx, *myy, *e;
    int cx, cy;
    double a, b;
};

// Entry of Z between a rooftop of direction dr and one of direction dc whose first faces differ by (p, q)
// (generateSpatialBlocks.m:35-48).
__device__ c128 z_entry(const GenView &v, int dr, int dc, int p, int q) {
    auto E = [&](int x, int y) { return gen_at(v.e, v.cx, v.cy, x, y); };
    if (dr == 0 && dc == 0)
        return gen_at(v.mxx, v.cx - 1, v.cy, p, q) - (1.0 / (v.a * v.a)) * (2.0 * E(p, q) - E(p - 1, q) - E(p + 1, q));
    if (dr == 1 && dc == 1)
        return gen_at(v.myy, v.cx, v.cy - 1, p, q) - (1.0 / (v.b * v.b)) * (2.0 * E(p, q) - E(p, q - 1) - E(p, q + 1));
    if (dr == 1) { p = -p; q = -q; }   // YX(p, q) = XY(-p, -q)
    return (-1.0 / (v.a * v.b)) * (E(p, q) - E(p, q - 1) - E(p + 1, q) + E(p + 1, q - 1));
}

// Sums over the wave, the same in every lane (a butterfly of lane exchanges, in a fixed order).
__device__ __forceinline__ double wave_total(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ c128 wave_total(c128 v) { return mk(wave_total(v.x), wave_total(v.y)); }
__device__ __forceinline__ c128 lane_value(c128 v, int k) {
    return mk(__shfl(v.x, k, 64), __shfl(v.y, k, 64));
}

// One wave per class: the fit matrix of the class's representative neighbourhood with a row per lane (the m <= 50
// rooftops of its 5 x 5 first-face neighbourhood; columns: the n <= 18 of its inner 3 x 3, then the right-hand side
// e_self), complex Householder QR on the lanes' registers (a column's norm and each reflection's inner products as
// wave sums), and the back substitution in every lane; coefficients by slot (0 where no rooftop).
constexpr int kFitWaves = 4;
__global__ void __launch_bounds__(64 * kFitWaves) sai_fit(GenView v, const int *__restrict__ dir,
                                                          const unsigned long long *__restrict__ bits, int classes,
                                                          c128 *__restrict__ coef, int *__restrict__ bad) {
    const int c = blockIdx.x * kFitWaves + (threadIdx.x >> 6);
    if (c >= classes) return;   // uniform over the wave; no barrier follows
    const int lane = threadIdx.x & 63;
    const unsigned long long mask = bits[c];
    const int d0 = dir[c];
    // The lane's row: the lane-th set bit of the mask (bit d * 25 + (dx + 2) + 5 (dy + 2)); the columns: the set bits
    // of the inner 3 x 3, in the same order.
    const int m = __popcll(mask);
    int rd = 0, rx = 0, ry = 0;
    {
        int seen = 0;
        for (int bit = 0; bit < 50; ++bit) {
            if (!((mask >> bit) & 1ull)) continue;
            if (seen == lane) {
                rd = bit / 25;
                rx = bit % 5 - 2;
                ry = (bit % 25) / 5 - 2;
            }
            ++seen;
        }
    }
    const bool row = lane < m;
    int cd[kSlots], cx[kSlots], cy[kSlots], slot[kSlots];
    int n = 0, self = -1;
    {
        int seen = 0;
        for (int bit = 0; bit < 50; ++bit) {
            if (!((mask >> bit) & 1ull)) continue;
            const int d = bit / 25, dx = bit % 5 - 2, dy = (bit % 25) / 5 - 2;
            if (d == d0 && dx == 0 && dy == 0) self = seen;
            ++seen;
            if (dx < -1 || dx > 1 || dy < -1 || dy > 1) continue;
            cd[n] = d;
            cx[n] = dx;
            cy[n] = dy;
            slot[n] = d * 9 + (dx + 1) + 3 * (dy + 1);
            ++n;
        }
    }
    c128 A[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; ++j)
        A[j] = (row && j < n) ? z_entry(v, rd, cd[j], rx - cx[j], ry - cy[j]) : mk(0.0, 0.0);
    c128 rhs = mk(row && lane == self ? 1.0 : 0.0, 0.0);
    bool ok = self >= 0 && n > 0;
#pragma unroll
    for (int k = 0; k < kSlots; ++k) {
        if (k >= n || !ok) break;   // uniform over the wave
        const c128 xk = A[k];
        const double nrm = sqrt(wave_total(lane >= k ? abs2(xk) : 0.0));
        if (!(nrm > 0)) {
            ok = false;
            break;
        }
        const c128 x0 = lane_value(xk, k);
        const double ax0 = sqrt(abs2(x0));
        const c128 phase = ax0 > 0 ? mk(x0.x / ax0, x0.y / ax0) : mk(1.0, 0.0);
        // The reflector v = x + phase nrm e_k on the lanes >= k; tau = 1 / (nrm (nrm + |x0|)).
        const c128 vk = lane < k ? mk(0.0, 0.0) : (lane == k ? x0 + nrm * phase : xk);
        const double tau = 1.0 / (nrm * (nrm + ax0));
#pragma unroll
        for (int j = k + 1; j < kSlots; ++j) {
            if (j >= n) break;
            const c128 w = tau * wave_total(cmulc(vk, A[j]));
            A[j] = A[j] - vk * w;
        }
        const c128 w = tau * wave_total(cmulc(vk, rhs));
        rhs = rhs - vk * w;
        if (lane == k) A[k] = mk(0.0, 0.0) - nrm * phase;   // R(k, k)
    }
    // Back substitution with R in lanes 0 .. n-1: every lane runs it on the broadcast entries.
    c128 x[kSlots];
#pragma unroll
    for (int k = kSlots - 1; k >= 0; --k) {
        x[k] = mk(0.0, 0.0);
        if (k >= n || !ok) continue;
        c128 sum = lane_value(rhs, k);
#pragma unroll
        for (int j = k + 1; j < kSlots; ++j)
            if (j < n) sum = sum - lane_value(A[j], k) * x[j];
        const c128 r = lane_value(A[k], k);
        const double den = abs2(r);
        x[k] = (1.0 / den) * cmulc(r, sum);
        if (!(den > 0) || !isfinite(x[k].x) || !isfinite(x[k].y)) ok = false;
    }
    if (lane < kSlots) coef[static_cast<int64_t>(c) * kSlots + lane] = mk(0.0, 0.0);
    if (!ok) {
        if (lane == 0) atomicAdd(bad, 1);
        return;
    }
#pragma unroll
    for (int j = 0; j < kSlots; ++j)
        if (j < n && lane == 0) coef[static_cast<int64_t>(c) * kSlots + slot[j]] = x[j];
}

// z = P r with P = (M + M^T)/2 as an 18-point stencil. The stencil of a rooftop depends only on its class and its
// neighbours' classes, so the rooftops are grouped into stencil classes whose 18 symmetrized coefficients
// (M_ij + M_ji)/2 are tabulated once (stencil_table): the apply reads its stencil class, then the 18 coefficients and
// the 18 neighbours' values, which do not depend on each other.
// One block per tile of TX x kTY face positions and right-hand side: the tile's values with a one-cell halo are
// staged in LDS once (coalesced along x), and the 18 neighbours of every rooftop are read from there, summed in slot
// order. SPLIT: a thread per rooftop (the first two waves the x rooftops, the last two the y rooftops of a 16 x 8
// tile), twice the threads of a small problem in flight; otherwise a thread per face position of a 32 x 8 tile, both
// of its rooftops (less halo per rooftop, for batches that fill the chip anyway).
constexpr int kTY = 8;
constexpr int kApplyThreads = 256;
template <typename TC, bool SPLIT>
__global__ void __launch_bounds__(kApplyThreads) sai_apply_tiled(const c128 *__restrict__ in, c128 *__restrict__ out,
                                                              const int *__restrict__ scls,
                                                              const cplx<TC> *__restrict__ scoef, int rows_x,
                                                              int cols_x, int rows_y, int cols_y, int64_t nx, int64_t n,
                                                              int tiles_x, int table_count, int64_t cls_stride,
                                                              int64_t coef_stride, const int *__restrict__ skip,
                                                              int skip_stride) {
    const int r = blockIdx.y;
    if (skip && skip[static_cast<int64_t>(r) * skip_stride]) return;
    scls += r * cls_stride;     // a batch of problems: this right-hand side's classes and coefficient table
    scoef += r * coef_stride;
    constexpr int kTX = SPLIT ? kApplyThreads / (2 * kTY) : kApplyThreads / kTY;
    constexpr int LX = kTX + 2, LY = kTY + 2;
    __shared__ cplx<TC> tile[2][LY][LX];
    // With table_count > 0 the class coefficients [table_count][18] are staged in LDS too (dynamic), so that the
    // neighbour values, the rooftops' stencil classes and the coefficients are all loaded at once: one memory latency
    // before the stencil instead of three in a row (values, then classes, then the coefficients they select).
    extern __shared__ unsigned char coef_raw[];
    cplx<TC> *coef_lds = reinterpret_cast<cplx<TC> *>(coef_raw);
    const int X0 = (blockIdx.x % tiles_x) * kTX, Y0 = (blockIdx.x / tiles_x) * kTY;
    const int d0 = SPLIT ? static_cast<int>(threadIdx.x) / (kTX * kTY) : 0;   // the first (or only) component
    const int f = threadIdx.x - d0 * (kTX * kTY);
    const int tx = f % kTX, ty = f / kTX;
    const int fx = X0 + tx, fy = Y0 + ty;
    constexpr int kComps = SPLIT ? 1 : 2;
    bool mine[kComps];
    int64_t idx[kComps];
    int stc[kComps];
    // Every load below reads a valid address (clamped into range) and a select drops what lies outside, so that
    // all of them are in flight together: a load under a branch is waited for inside it, one at a time.
#pragma unroll
    for (int c = 0; c < kComps; ++c) {
        const int di = d0 + c;
        const int rows_i = di ? rows_y : rows_x, cols_i = di ? cols_y : cols_x;
        mine[c] = fx < rows_i && fy < cols_i;
        idx[c] = (di ? nx : 0) + fx + static_cast<int64_t>(rows_i) * fy;
        const int sv = scls[min(idx[c], n - 1)];
        stc[c] = mine[c] ? sv : -1;
    }
    constexpr int kStage = 16 * 1024 / static_cast<int>(sizeof(cplx<TC>)) / kApplyThreads;   // the LDS table's cap
    const int tc = table_count * kSlots;
    cplx<TC> cv[kStage];   // (the table holds at least one class: entry 0 is always valid)
#pragma unroll
    for (int u = 0; u < kStage; ++u)
        cv[u] = scoef[min(static_cast<int>(threadIdx.x) + u * kApplyThreads, max(tc - 1, 0))];
    const c128 *x = in + static_cast<int64_t>(r) * n;
    constexpr int kHalo = (2 * LY * LX + kApplyThreads - 1) / kApplyThreads;
    cplx<TC> hv[kHalo];
#pragma unroll
    for (int u = 0; u < kHalo; ++u) {
        const int e = min(static_cast<int>(threadIdx.x) + u * kApplyThreads, 2 * LY * LX - 1);
        const int d = e / (LY * LX), rem = e - d * (LY * LX), ly = rem / LX, lx = rem - ly * LX;
        const int gx = X0 + lx - 1, gy = Y0 + ly - 1;
        const int rows = d ? rows_y : rows_x, cols = d ? cols_y : cols_x;
        const bool in_grid = gx >= 0 && gy >= 0 && gx < rows && gy < cols;
        const int cx = min(max(gx, 0), max(rows - 1, 0)), cy = min(max(gy, 0), max(cols - 1, 0));
        const cplx<TC> v = cvt<TC>(x[min((d ? nx : 0) + cx + static_cast<int64_t>(rows) * cy, n - 1)]);
        hv[u] = in_grid ? v : mk<TC>(0, 0);
    }
    if (tc > 0) {
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            const int e = static_cast<int>(threadIdx.x) + u * kApplyThreads;
            if (e < tc) coef_lds[e] = cv[u];
        }
    }
#pragma unroll
    for (int u = 0; u < kHalo; ++u) {
        const int e = static_cast<int>(threadIdx.x) + u * kApplyThreads;
        if (e < 2 * LY * LX) {
            const int d = e / (LY * LX), rem = e - d * (LY * LX), ly = rem / LX, lx = rem - ly * LX;
            tile[d][ly][lx] = hv[u];
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kComps; ++c) {
        if (!mine[c]) continue;
        const int64_t i = idx[c];
        const int st = stc[c];
        c128 res = mk(0.0, 0.0);
        {   // every coefficient and neighbour read first, then the sum over the non-zero coefficients in slot order
            const cplx<TC> *cf = (table_count > 0 ? coef_lds : scoef) + static_cast<int64_t>(max(st, 0)) * kSlots;
            cplx<TC> cs[kSlots], vs[kSlots];
#pragma unroll
            for (int d = 0; d < 2; ++d)
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int s = d * 9 + (dx + 1) + 3 * (dy + 1);
                        cs[s] = cf[s];
                        vs[s] = tile[d][ty + 1 + dy][tx + 1 + dx];
                    }
            cplx<TC> acc = mk<TC>(0, 0);
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const cplx<TC> nx = acc + cs[s] * vs[s];
                const bool use = cs[s].x != 0 || cs[s].y != 0;
                acc = mk<TC>(use ? nx.x : acc.x, use ? nx.y : acc.y);
            }
            if (st >= 0) res = cvt<double>(acc);
        }
        out[static_cast<int64_t>(r) * n + i] = res;
    }
}

// scoef[t][s] = (coef[ci][s] + coef[cj][s']) / 2 for the stencil class t of tuple (ci, cj_0 .. cj_17), s' the slot
// of the rooftop seen from its neighbour; 0 for an absent neighbour (cj < 0). The halving is exact, so the apply
// equals the one that sums (coef + coef) x and halves the sum.
template <typename TC>
__global__ void stencil_table(const int *__restrict__ tuples, int count, const cplx<TC> *__restrict__ coef,
                              cplx<TC> *__restrict__ scoef) {
    const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (e >= static_cast<int64_t>(count) * kSlots) return;
    const int t = static_cast<int>(e / kSlots), s = static_cast<int>(e - static_cast<int64_t>(t) * kSlots);
    const int *tp = tuples + static_cast<int64_t>(t) * (kSlots + 2);
    const int ci = tp[0], di = tp[1], cj = tp[2 + s];
    if (cj < 0) {
        scoef[e] = mk<TC>(0, 0);
        return;
    }
    const int d = s / 9, dx = s % 3 - 1, dy = (s / 3) % 3 - 1;
    const int sm = di * 9 + (1 - dx) + 3 * (1 - dy);
    (void)d;
    const cplx<TC> sum = coef[static_cast<int64_t>(ci) * kSlots + s] + coef[static_cast<int64_t>(cj) * kSlots + sm];
    scoef[e] = static_cast<TC>(0.5) * sum;
}

// Jacobi: z = r / diag(Z); diag is XX(0,0) on x rooftops and YY(0,0) on y rooftops (generateSpatialBlocks.m:35-45).
__global__ void jacobi_diag(GenView v, c128 *__restrict__ inv) {
    if (threadIdx.x != 0) return;
    for (int d = 0; d < 2; ++d) {
        if (d == 0 && !v.mxx) { inv[0] = mk(0.0, 0.0); continue; }
        if (d == 1 && !v.myy) { inv[1] = mk(0.0, 0.0); continue; }
        const c128 z = z_entry(v, d, d, 0, 0);
        const double den = abs2(z);
        inv[d] = (1.0 / den) * conjg(z);
    }
}

template <typename TC>
__global__ void jacobi_apply(const c128 *__restrict__ in, c128 *__restrict__ out, const cplx<TC> *__restrict__ inv,
                             const uint8_t *__restrict__ exists, int64_t nx, int64_t n, int nrhs,
                             const int *__restrict__ skip, int skip_stride) {
    const int64_t total = n * nrhs;
    for (int64_t id = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; id < total;
         id += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if (skip && skip[(id / n) * skip_stride]) continue;
        const int64_t i = id % n;
        cplx<TC> v = inv[i < nx ? 0 : 1] * cvt<TC>(in[id]);
        if (exists && !exists[i]) v = mk<TC>(0, 0);
        out[id] = cvt<double>(v);
    }
}

__global__ void to_c64_pc(const c128 *__restrict__ a, c64 *__restrict__ b, int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        b[i] = cvt<float>(a[i]);
}

// ------------------------------------------------------------------------------------------------------------------
// Host side.
// ------------------------------------------------------------------------------------------------------------------
static GenView view_of(Problem *p) {
    GenView v;
    v.mxx = p->g.has_x ? p->mxx.ptr : nullptr;
    v.myy = p->g.has_y ? p->myy.ptr : nullptr;
    v.e = p->e.ptr;
    v.cx = p->g.cx;
    v.cy = p->g.cy;
    v.a = p->g.a;
    v.b = p->g.b;
    return v;
}

std::unique_ptr<Product> create_preconditioner(Problem *p, int kind, int precision) {
    Context *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    const GridInfo &g = p->g;
    p->ensure_generators();
    auto pc = std::make_unique<Product>();
    pc->problem = p;
    pc->precision = precision;
    pc->dims.n = g.n;
    pc->dims.nx_count = g.nx;
    pc->dims.rows_x = g.rows_x;
    pc->dims.rows_y = g.rows_y;
    pc->dims.cols_x = g.has_x ? g.cols_x : 0;
    pc->dims.cols_y = g.has_y ? g.cols_y : 0;
    const GenView v = view_of(p);
    DeviceArray<int> bad;
    bad.allocate(ctx, 1, "preconditioner status");
    bad.zero(s);
    if (kind == RT_PRECOND_JACOBI) {
        pc->type = kJacobi;
        pc->diag64.allocate(ctx, 2, "The Jacobi diagonal");
        hipLaunchKernelGGL(jacobi_diag, dim3(1), dim3(64), 0, s, v, pc->diag64.ptr);
        RT_LAUNCH_CHECK();
        c128 h[2];
        RT_HIP(hipMemcpyAsync(h, pc->diag64.ptr, sizeof h, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        for (int d = 0; d < 2; ++d) {
            const bool present = d ? g.has_y : g.has_x;
            if (present && !(std::isfinite(h[d].x) && std::isfinite(h[d].y) && (h[d].x != 0 || h[d].y != 0)))
                fail(RT_ESINGULAR, "rooftopMoM:preconditioner:singular",
                     "The preconditioner is singular: the diagonal of the matrix has a zero or non-finite entry.");
        }
        if (precision != RT_C128) {
            pc->diag32.allocate(ctx, 2, "The Jacobi diagonal");
            hipLaunchKernelGGL(to_c64_pc, dim3(1), dim3(64), 0, s, pc->diag64.ptr, pc->diag32.ptr, int64_t(2));
            RT_LAUNCH_CHECK();
        }
    } else if (kind == RT_PRECOND_NEARFIELD) {
        pc->type = kNearField;
        // Classes (5 x 5 windows: one least-squares fit each) and stencil classes (a class, a direction and the
        // neighbours' classes: one symmetrized stencil each), formed on the device.
        NearFieldClasses nc = classify_near_field(*p, s);
        const int classes = nc.classes, stencils = nc.stencils;
        pc->class_count = classes;
        pc->cls = std::move(nc.cls);
        pc->coef64.allocate(ctx, static_cast<size_t>(classes) * kSlots, "The near-field coefficients");
        if (classes > 0) {
            hipLaunchKernelGGL(sai_fit, dim3((classes + kFitWaves - 1) / kFitWaves), dim3(64 * kFitWaves), 0, s, v,
                               nc.rep_dir.ptr, nc.rep_bits.ptr, classes, pc->coef64.ptr, bad.ptr);
            RT_LAUNCH_CHECK();
            int h_bad = 0;
            RT_HIP(hipMemcpyAsync(&h_bad, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
            RT_HIP(hipStreamSynchronize(s));
            if (h_bad)
                fail(RT_ESINGULAR, "rooftopMoM:preconditioner:singular",
                     "The preconditioner is singular: %d of %d near-field fits are rank deficient or not finite.",
                     h_bad, classes);
        }
        if (precision != RT_C128 && classes > 0) {
            pc->coef32.allocate(ctx, static_cast<size_t>(classes) * kSlots, "The near-field coefficients");
            hipLaunchKernelGGL(to_c64_pc, dim3(blocks_for(static_cast<size_t>(classes) * kSlots, 256, 1024)), dim3(256),
                               0, s, pc->coef64.ptr, pc->coef32.ptr, static_cast<int64_t>(classes) * kSlots);
            RT_LAUNCH_CHECK();
        }
        pc->stencil_count = stencils;
        pc->scls = std::move(nc.scls);
        if (stencils > 0) {
            const unsigned nb = blocks_for(static_cast<size_t>(stencils) * kSlots, 256, 65535);
            if (precision != RT_C64) {
                pc->scoef64.allocate(ctx, static_cast<size_t>(stencils) * kSlots, "The near-field stencils");
                hipLaunchKernelGGL(stencil_table<double>, dim3(nb), dim3(256), 0, s, nc.tuples.ptr, stencils,
                                   pc->coef64.ptr, pc->scoef64.ptr);
                RT_LAUNCH_CHECK();
            }
            if (precision != RT_C128) {
                pc->scoef32.allocate(ctx, static_cast<size_t>(stencils) * kSlots, "The near-field stencils");
                hipLaunchKernelGGL(stencil_table<float>, dim3(nb), dim3(256), 0, s, nc.tuples.ptr, stencils,
                                   pc->coef32.ptr, pc->scoef32.ptr);
                RT_LAUNCH_CHECK();
            }
            RT_HIP(hipStreamSynchronize(s));
        }
    } else {
        fail(RT_EINVAL, "rooftop:preconditioner", "unknown preconditioner kind %d", kind);
    }
    if (precision == RT_C64) {
        RT_HIP(hipStreamSynchronize(s));
        pc->diag64.reset();
        pc->coef64.reset();
    }
    RT_HIP(hipStreamSynchronize(s));
    return pc;
}

void preconditioner_apply(Product &pc, const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s,
                          const int *skip, int skip_stride) {
    if (prec == RT_MIXED) prec = RT_C128;
    if (!pc.has(prec)) fail(RT_EINVAL, "rooftop:product", "the preconditioner does not hold the requested precision");
    const ProdDims &d = pc.dims;
    const int64_t total = d.n * nrhs;
    // Algorithmic bytes: the vectors in and out, the class (or diagonal) of every rooftop, the coefficient table.
    const int es = prec == RT_C128 ? 16 : 8;
    const double bytes = 32.0 * total +
                         (pc.type == kJacobi ? static_cast<double>(es) * d.n
                                             : 4.0 * d.n + 18.0 * es * pc.stencil_count);
    KernelTimer timer(pc.problem->ctx, prec == RT_C128 ? RT_TIMED_PRECOND_C128 : RT_TIMED_PRECOND_C64, bytes);
    if (pc.type == kJacobi) {
        const unsigned blocks = blocks_for(total, 256, 8192);
        if (prec == RT_C128)
            launch_timed(&timer, jacobi_apply<double>, dim3(blocks), dim3(256), 0, s, in, out, pc.diag64.ptr,
                         pc.problem->exists.ptr, d.nx_count, d.n, nrhs, skip, skip_stride);
        else
            launch_timed(&timer, jacobi_apply<float>, dim3(blocks), dim3(256), 0, s, in, out, pc.diag32.ptr,
                         pc.problem->exists.ptr, d.nx_count, d.n, nrhs, skip, skip_stride);
        timer.done();
        return;
    }
    if (pc.class_count == 0) fail(RT_EINVAL, "rooftop:preconditioner", "the grid has no rooftop to precondition");
    // Tiles cover the face positions 0 .. countX-1 by 0 .. countY-1 (the y rooftops' rows and the x rooftops' columns).
    const int cx = d.rows_y > 0 ? d.rows_y : d.rows_x + 1, cy = d.cols_x > 0 ? d.cols_x : d.cols_y + 1;
    // A thread per rooftop on 16 x 8 tiles while the 32 x 8 tiles would leave fewer than 2048 blocks.
    const bool split = static_cast<int64_t>((cx + 31) / 32) * ((cy + kTY - 1) / kTY) * nrhs < 2048;
    const int tx_w = split ? kApplyThreads / (2 * kTY) : kApplyThreads / kTY;
    const int tiles_x = (cx + tx_w - 1) / tx_w, tiles_y = (cy + kTY - 1) / kTY;
    const dim3 grid(static_cast<unsigned>(tiles_x) * tiles_y, nrhs);
    // The coefficient table goes to LDS when it is small (a plate has a few dozen stencil classes).
    const size_t table = static_cast<size_t>(pc.stencil_count) * kSlots * es;
    const int staged = table <= 16 * 1024 ? pc.stencil_count : 0;
    const uint32_t lds = staged ? static_cast<uint32_t>(table) : 0;
    auto launch = [&](auto kernel, const auto *coef) {
        launch_timed(&timer, kernel, grid, dim3(kApplyThreads), lds, s, in, out, pc.scls.ptr, coef, d.rows_x,
                     d.cols_x, d.rows_y, d.cols_y, d.nx_count, d.n, tiles_x, staged, pc.cls_stride,
                     pc.coef_stride, skip, skip_stride);
    };
    if (prec == RT_C128) {
        if (split) launch(sai_apply_tiled<double, true>, pc.scoef64.ptr);
        else launch(sai_apply_tiled<double, false>, pc.scoef64.ptr);
    } else {
        if (split) launch(sai_apply_tiled<float, true>, pc.scoef32.ptr);
        else launch(sai_apply_tiled<float, false>, pc.scoef32.ptr);
    }
    timer.done();
}

}  // namespace rt

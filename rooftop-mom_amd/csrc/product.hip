// The Toeplitz product of the rooftop Green's matrix on gfx950: the three generators embedded evenly into an
// Mx x My circulant (period >= 2n-1 per split direction), their spectra folded by their even symmetry, and the
// matvec as three FFT passes with the spectral multiply fused into the middle one.
//
// Reference: the products of src/+rooftopMoM/+assembler/+greensMatrix/FFTApplicator.m:14-87 and
// @SplitFFTApplicator/applyBothDirections_.m:14-85 with the block spectra of +utils/generateBlockSpectra.m:19-50.
// With Dx = (1 - sx)/a, Dy = (1 - sy)/b the four block spectra collapse into three planes:
//     out_x = Mxx X - Dx E (conj(Dx) X + conj(Dy) Y),   out_y = Myy Y - Dy E (conj(Dx) X + conj(Dy) Y),
// the x and y current spectra X, Y and the face-charge spectrum conj(Dx) X + conj(Dy) Y ("Jx, Jy and charge").
//
// The passes themselves are in passes.hpp and pass_{a,b,c}.inl; this file holds the tables, the spectra, the launch
// shapes and their autotuner.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <type_traits>

#include "circulant_host.hpp"
#include "passes.hpp"

namespace rt {

// ------------------------------------------------------------------------------------------------------------------
// Tables.
// ------------------------------------------------------------------------------------------------------------------
void DimTables::build(Context *ctx, int M, double size, hipStream_t s) {
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int v = 0; v < 2; ++v) {
        PlanTables &t = p[v];
        t.plan = make_plan(M, v == 0 ? 3 : 4);
        const std::vector<c128> h_tw = plan_twiddles(t.plan);
        std::vector<c64> h_tw32(h_tw.size());
        for (size_t i = 0; i < h_tw.size(); ++i) h_tw32[i] = cvt<float>(h_tw[i]);
        std::vector<c128> h_d(M);
        std::vector<c64> h_d32(M);
        std::vector<int> h_fold(M);
        const std::vector<int> freq = output_frequencies(t.plan);
        for (int q = 0; q < M; ++q) {
            const int k = freq[q];
            h_fold[q] = k <= M - k ? k : M - k;
            const long double ang = 2.0L * pi * static_cast<long double>(k) / static_cast<long double>(M);
            // D = (1 - exp(i ang)) / size; 1 - cos = 2 sin^2(ang/2) keeps low frequencies accurate.
            const long double sh = std::sin(ang / 2);
            h_d[q] = {static_cast<double>(2.0L * sh * sh / size), static_cast<double>(-std::sin(ang) / size)};
            h_d32[q] = cvt<float>(h_d[q]);
        }
        t.tw.allocate(ctx, h_tw.size(), "FFT twiddles");
        t.tw32.allocate(ctx, h_tw.size(), "FFT twiddles");
        t.fold.allocate(ctx, M, "FFT fold table");
        std::vector<int> h_order(M);
        for (int q = 0; q < M; ++q) h_order[q] = q;
        std::stable_sort(h_order.begin(), h_order.end(), [&](int u, int v) { return h_fold[u] < h_fold[v]; });
        t.border.allocate(ctx, M, "pass B position order");
        t.d.allocate(ctx, M, "divergence table");
        t.d32.allocate(ctx, M, "divergence table");
        RT_HIP(hipMemcpyAsync(t.tw.ptr, h_tw.data(), h_tw.size() * sizeof(c128), hipMemcpyHostToDevice, s));
        RT_HIP(hipMemcpyAsync(t.tw32.ptr, h_tw32.data(), h_tw.size() * sizeof(c64), hipMemcpyHostToDevice, s));
        RT_HIP(hipMemcpyAsync(t.fold.ptr, h_fold.data(), M * sizeof(int), hipMemcpyHostToDevice, s));
        RT_HIP(hipMemcpyAsync(t.border.ptr, h_order.data(), M * sizeof(int), hipMemcpyHostToDevice, s));
        RT_HIP(hipMemcpyAsync(t.d.ptr, h_d.data(), M * sizeof(c128), hipMemcpyHostToDevice, s));
        RT_HIP(hipMemcpyAsync(t.d32.ptr, h_d32.data(), M * sizeof(c64), hipMemcpyHostToDevice, s));
        RT_HIP(hipStreamSynchronize(s));
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Spectra: the even embedding of a generator (G(p mod Mx, q mod My) = g(|p|, |q|) for |p| < rows, |q| < cols), its
// forward 2D FFT, and the fold to [0, Mx/2] x [0, My/2] (an even real-index circulant has an even spectrum). One
// block (one team) per column; plans with radices up to 8.
// ------------------------------------------------------------------------------------------------------------------
template <int E>
__global__ void __launch_bounds__(1024) spectra_a(const c128 *__restrict__ g, int rows, int cols, int modifier,
                                                  c128 *__restrict__ out, int Mx, int My, FftPlan plan,
                                                  const c128 *__restrict__ tw) {
    extern __shared__ unsigned char smem_raw[];
    c128 *buf = reinterpret_cast<c128 *>(smem_raw);
    const int q = blockIdx.x;
    const int qq = min(q, My - q);
    for (int i = threadIdx.x; i < Mx; i += blockDim.x) {
        const int pp = min(i, Mx - i);
        c128 v = mk(0.0, 0.0);
        if (g && pp < rows && qq < cols) {
            v = g[static_cast<int64_t>(qq) * rows + pp];
            if (modifier == RT_MODIFIER_REAL) v.y = 0;
            if (modifier == RT_MODIFIER_IMAG) v = mk(v.y, 0.0);
        }
        buf[padded(plan, i)] = v;
    }
    __syncthreads();
    team_fft_lds<double, E, false>(buf, plan, tw, threadIdx.x);
    for (int i = threadIdx.x; i < Mx; i += blockDim.x) out[static_cast<int64_t>(q) * Mx + i] = buf[padded(plan, i)];
}

template <int E>
__global__ void __launch_bounds__(1024) spectra_b(const c128 *__restrict__ in, c128 *__restrict__ plane, int Mx,
                                                  int My, int Hy, FftPlan plan, const c128 *__restrict__ tw,
                                                  const int *__restrict__ freq_x, const int *__restrict__ freq_y,
                                                  double scale) {
    extern __shared__ unsigned char smem_raw[];
    c128 *buf = reinterpret_cast<c128 *>(smem_raw);
    const int px = blockIdx.x;
    for (int q = threadIdx.x; q < My; q += blockDim.x) buf[padded(plan, q)] = in[static_cast<int64_t>(q) * Mx + px];
    __syncthreads();
    team_fft_lds<double, E, false>(buf, plan, tw, threadIdx.x);
    const int kx = freq_x[px];
    if (2 * kx > Mx) return;
    for (int py = threadIdx.x; py < My; py += blockDim.x) {
        const int ky = freq_y[py];
        if (2 * ky <= My) plane[static_cast<int64_t>(kx) * Hy + ky] = scale * buf[padded(plan, py)];
    }
}


// ------------------------------------------------------------------------------------------------------------------
// Host side.
// ------------------------------------------------------------------------------------------------------------------
size_t lds_bytes_max() {
    static std::atomic<size_t> cached{0};
    size_t v = cached.load(std::memory_order_relaxed);
    if (!v) {
        int dev = 0, got = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&got, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || got <= 0)
            got = 64 * 1024;
        v = static_cast<size_t>(got);
        cached.store(v, std::memory_order_relaxed);
    }
    return v;
}

void set_lds(const void *kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    if (bytes > lds_bytes_max())
        fail(RT_EINVAL, "rooftop:product", "a pass needs %zu bytes of LDS, more than this device has", bytes);
    static std::mutex lock;
    static std::unordered_map<const void *, size_t> raised;
    std::lock_guard<std::mutex> g(lock);
    size_t &have = raised[kernel];
    if (have >= bytes) return;
    RT_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
    have = bytes;
}

// Whether pass B of a shape runs in one launch (both columns of `tile` x positions in the LDS of a block).
template <typename TC>
static bool b_fits(const Product &op, const PassShape &sh) {
    const FftPlan &py = op.ty.of(sh.bits).plan;
    return b_threads_fit(py, sh.tile, sh.teams) && lds_b_bytes(py, sh.tile, sizeof(cplx<TC>), sh.teams) <= lds_bytes_max();
}
// Whether pass B can run in three launches instead (one column per block).
template <typename TC>
static bool b_split_fits(const Product &op, int bits) {
    const FftPlan &py = op.ty.of(bits).plan;
    return team_threads(py) <= 1024 && lds_b_split_bytes(py, sizeof(cplx<TC>)) <= lds_bytes_max();
}

// Whether a shape fits the device: threads per block and LDS.
template <typename TC>
static bool shape_fits(const Product &op, const PassShape &sh) {
    const FftPlan &px = op.tx.of(sh.bits).plan;
    const int es = sizeof(cplx<TC>);
    const bool ac = ac_threads_fit(px, sh.group, es) && lds_ac_bytes(px, sh.group, es) <= lds_bytes_max() &&
                    (sh.group > 0 || op.embedding);
    const bool b = b_fits<TC>(op, sh) || (sh.tile == 1 && sh.teams == 1 && b_split_fits<TC>(op, sh.bits));
    return sh.tile >= 1 && sh.tile <= op.dims.Mx && ac && b;
}

// Default launch shape (the autotuner may replace it), from the tuner's choices measured over plates of 32 to 4096
// cells at one and eight right-hand sides: radices up to 8 (complex64 transforms up to 2048 points, complex128 up to
// 512), else up to 16; pass A/C groups of four columns where they fit a block, else four (complex64) or two
// (complex128) columns in turn through one column of LDS; pass B through one column of LDS for
// transforms of at least 128 points in complex64 and 1024 in complex128; pass B tiles of up to 256 threads while the
// blocks stay a few dozen. The plan and the pass B form depend on the transform lengths alone, so a product runs the
// same kernels (and the same arithmetic per right-hand side) whatever the batch size.
int default_bits(const ProdDims &d, bool c64) { return std::max(d.Mx, d.My) <= (c64 ? 2048 : 512) ? 3 : 4; }

template <typename TC>
static PassShape default_shape(const Product &op, int nrhs, int bits = 0) {
    const ProdDims &d = op.dims;
    const bool c64 = sizeof(TC) == sizeof(float);
    PassShape s;
    s.bits = bits ? bits : default_bits(d, c64);
    const bool plain = op.embedding && !op.four_plane;
    if (!plain) s.bits = 3;   // face-grid and four-plane passes: the radix-8 plans
    const FftPlan &px = op.tx.of(s.bits).plan, &py = op.ty.of(s.bits).plan;
    const int es = sizeof(cplx<TC>);
    s.teams = plain && py.K >= 2 && py.M >= (c64 ? 128 : 1024) ? 3 : 1;
    int g = 4;
    while (g > 1 && (team_threads(px) * g > max_threads_ac(es, px) || lds_ac_bytes(px, g, es) > lds_bytes_max()))
        g >>= 1;
    // Long rows whose side-by-side groups stay below four columns: four (complex64) or two (complex128) in turn.
    const int turn = c64 ? -4 : -2;
    if (g < 4 && plain && ac_threads_fit(px, turn, es) && lds_ac_bytes(px, turn, es) <= lds_bytes_max()) g = turn;
    s.group = g;
    int tile = 1;
    while (tile * 2 <= d.Mx && team_threads(py) * tile * 2 <= 256 &&
           lds_b_bytes(py, tile * 2, es, s.teams) <= 48 * 1024 && static_cast<int64_t>(d.Mx / (tile * 2)) * nrhs >= 16)
        tile *= 2;
    s.tile = tile;
    return s;
}

template <typename TC>
static PassShape shape_for(const Product &op, int nrhs) {
    const auto &tuned = sizeof(TC) == sizeof(float) ? op.tuned32 : op.tuned64;
    for (const auto &e : tuned)
        if (e.first == nrhs) return e.second;
    return default_shape<TC>(op, nrhs);
}

// Algorithmic HBM bytes of one launch of each pass: every vector element, work element and spectrum value once.
template <typename TC>
static void pass_bytes(const ProdDims &d, int nrhs, double *bytes, bool four_plane = false) {
    const double sc = sizeof(cplx<TC>);
    const double vec = 16.0 * d.n * nrhs;
    const double work_bytes = sc * (static_cast<double>(d.cols_x) + d.cols_y) * d.Mx * nrhs;
    bytes[0] = vec + work_bytes;
    bytes[1] = 2 * work_bytes + (four_plane ? 4.0 * sc * d.Mx * d.My : 3.0 * sc * d.Hx * d.My);
    bytes[2] = work_bytes + vec;
}

// bspec[(fx My + py) 3 + k] = plane k at (fx, fold_y(py)), in the precision of the pass.
template <typename TC>
__global__ void build_bspec(const c128 *__restrict__ planes, int Hx, int Hy, int My, const int *__restrict__ fold_y,
                            cplx<TC> *__restrict__ bspec) {
    const int64_t total = static_cast<int64_t>(Hx) * My;
    const int64_t plane = static_cast<int64_t>(Hx) * Hy;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t fx = i / My;
        const int py = static_cast<int>(i - fx * My);
        const int64_t f = fx * Hy + fold_y[py];
        for (int k = 0; k < 3; ++k) bspec[3 * i + k] = cvt<TC>(planes[k * plane + f]);
    }
}

template <typename TC>
static const cplx<TC> *bspec_of(Product &op, int bits, hipStream_t s) {
    const int v = bits >= 4 ? 1 : 0;
    const ProdDims &d = op.dims;
    const size_t count = static_cast<size_t>(3) * d.Hx * d.My;
    auto build = [&](auto &arr) {
        using Elem = typename std::remove_reference<decltype(*arr.ptr)>::type;
        if (!arr.ptr) {
            if (!op.planes64.ptr || d.spec_stride)
                fail(RT_EINVAL, "rooftop:product", "the pass-B spectra of this launch shape are not built");
            arr.allocate(op.problem->ctx, count, "The spectra of the product");
            hipLaunchKernelGGL((build_bspec<decltype(Elem{}.x)>), dim3(blocks_for(count / 3, 256, 8192)), dim3(256),
                               0, s, op.planes64.ptr, d.Hx, d.Hy, d.My, op.ty.p[v].fold.ptr, arr.ptr);
            RT_LAUNCH_CHECK();
        }
        return reinterpret_cast<const cplx<TC> *>(arr.ptr);
    };
    if (sizeof(TC) == sizeof(double)) return build(op.bspec64[v]);
    return build(op.bspec32[v]);
}

// The full-period y spectra of the three-launch pass B, grown to nrhs right-hand sides.
template <typename TC>
static cplx<TC> *spec_of(Product &op, int nrhs) {
    const size_t per = static_cast<size_t>(2) * op.dims.Mx * op.dims.My;
    if (sizeof(TC) == sizeof(double)) {
        if (op.spec_rhs64 < nrhs) {
            op.spec64.allocate(op.problem->ctx, per * nrhs, "the product's working memory");
            op.spec_rhs64 = nrhs;
        }
        return reinterpret_cast<cplx<TC> *>(op.spec64.ptr);
    }
    if (op.spec_rhs32 < nrhs) {
        op.spec32.allocate(op.problem->ctx, per * nrhs, "the product's working memory");
        op.spec_rhs32 = nrhs;
    }
    return reinterpret_cast<cplx<TC> *>(op.spec32.ptr);
}

// The three passes of nrhs right-hand sides on stream s; after_a (if set) is recorded once pass A is queued.
template <typename TC>
static void run_chunk(Product &op, const c128 *in, c128 *out, cplx<TC> *work, int nrhs, const PassShape &sh,
                      hipStream_t s, const int *skip, int skip_stride, hipEvent_t after_a) {
    const ProdDims &d = op.dims;
    Context *ctx = op.problem->ctx;
    const PlanTables &tx = op.tx.of(sh.bits), &ty = op.ty.of(sh.bits);
    const cplx<TC> *bspec = op.four_plane ? reinterpret_cast<const cplx<TC> *>(sizeof(TC) == sizeof(double)
                                                                                    ? static_cast<const void *>(op.b4spec64.ptr)
                                                                                    : static_cast<const void *>(op.b4spec32.ptr))
                                          : bspec_of<TC>(op, sh.bits, s);
    if (op.four_plane && !b_fits<TC>(op, sh))
        fail(RT_EINVAL, "rooftop:preconditioner:unsupported",
             "The blockwise circulant preconditioner needs the face grid's columns to fit one block's LDS.");
    double by[3];
    pass_bytes<TC>(d, nrhs, by, op.four_plane);
    const int slot0 = sizeof(TC) == sizeof(float) ? RT_TIMED_PASS_A_C64 : RT_TIMED_PASS_A_C128;
    KernelTimer ta(ctx, slot0, by[0]);
    launch_pass_a<TC>(op, tx, in, work, nrhs, sh.group, s, skip, skip_stride, &ta);
    ta.done();
    if (after_a) RT_HIP(hipEventRecord(after_a, s));
    if (b_fits<TC>(op, sh)) {
        KernelTimer tb(ctx, slot0 + 1, by[1]);
        launch_pass_b<TC>(op, tx, ty, work, bspec, nrhs, sh.tile, sh.teams, s, skip, skip_stride, &tb);
        tb.done();
    } else {
        // Three launches, timed together as pass B (events recorded around them on the stream).
        if (d.spec_stride || d.dx_stride)
            fail(RT_EINVAL, "rooftop:product", "a batch of problems needs y transforms that fit one block's LDS");
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (ctx->timing) {
            e0 = ctx->take_event();
            RT_HIP(hipEventRecord(e0, s));
        }
        launch_pass_b_split<TC>(op, tx, ty, work, spec_of<TC>(op, nrhs), bspec, nrhs, s, skip, skip_stride);
        if (e0) {
            e1 = ctx->take_event();
            RT_HIP(hipEventRecord(e1, s));
            ctx->pending.push_back(
                Context::TimedRecord{e0, e1, slot0 + 1, by[1] * ctx->timed_scale, by[1], ctx->timed_step});
        }
    }
    KernelTimer tc(ctx, slot0 + 2, by[2]);
    launch_pass_c<TC>(op, tx, work, out, nrhs, sh.group, s, skip, skip_stride, &tc);
    tc.done();
}

// The matvec of nrhs right-hand sides. With sh.chunks > 1 the right-hand sides run in chunks that alternate between
// the main stream s and the context's side stream, the side stream one pass A behind: the HBM-bound passes A and C of
// one chunk then overlap the VALU- and LDS-bound pass B of another. Chunks touch disjoint columns of the input,
// output and work arrays; s waits for the side stream before it goes on.
template <typename TC>
static void run_passes(Product &op, const c128 *in, c128 *out, cplx<TC> *work, int nrhs, const PassShape &sh,
                       hipStream_t s, const int *skip = nullptr, int skip_stride = 0) {
    const ProdDims &d = op.dims;
    const bool batch_of_problems = d.spec_stride || d.mask_stride || d.dx_stride;   // chunk offsets not supported
    const int G = batch_of_problems ? 1 : std::min(sh.chunks, nrhs);
    if (G <= 1 || !b_fits<TC>(op, sh)) {
        run_chunk<TC>(op, in, out, work, nrhs, sh, s, skip, skip_stride, nullptr);
        return;
    }
    Context *ctx = op.problem->ctx;
    hipStream_t side = ctx->side_stream();
    hipEvent_t fork = ctx->sync_event(0), first_a = ctx->sync_event(1), join = ctx->sync_event(2);
    RT_HIP(hipEventRecord(fork, s));
    RT_HIP(hipStreamWaitEvent(side, fork, 0));
    const int64_t n = op.dims.n, per = static_cast<int64_t>(2) * op.dims.ldj * op.dims.Mx;
    int r0 = 0;
    for (int c = 0; c < G; ++c) {
        const int cnt = nrhs / G + (c < nrhs % G ? 1 : 0);
        if (c == 1) RT_HIP(hipStreamWaitEvent(side, first_a, 0));
        run_chunk<TC>(op, in + r0 * n, out + r0 * n, work + r0 * per, cnt, sh, (c & 1) ? side : s,
                      skip ? skip + static_cast<int64_t>(r0) * skip_stride : nullptr, skip_stride,
                      c == 0 ? first_a : nullptr);
        r0 += cnt;
    }
    RT_HIP(hipEventRecord(join, side));
    RT_HIP(hipStreamWaitEvent(s, join, 0));
}

template <typename TC>
static cplx<TC> *work_of(Product &op, int nrhs) {
    const size_t per_rhs = static_cast<size_t>(2) * op.dims.ldj * op.dims.Mx;
    if (sizeof(TC) == sizeof(double)) {
        if (op.work_rhs64 < nrhs) {
            op.work64.allocate(op.problem->ctx, per_rhs * nrhs, "the product's working memory");
            op.work_rhs64 = nrhs;
        }
        return reinterpret_cast<cplx<TC> *>(op.work64.ptr);
    }
    if (op.work_rhs32 < nrhs) {
        op.work32.allocate(op.problem->ctx, per_rhs * nrhs, "the product's working memory");
        op.work_rhs32 = nrhs;
    }
    return reinterpret_cast<cplx<TC> *>(op.work32.ptr);
}

template <typename TC>
static void apply_as(Product &op, const c128 *in, c128 *out, int nrhs, hipStream_t s, const int *skip,
                     int skip_stride) {
    cplx<TC> *work = work_of<TC>(op, nrhs);
    const PassShape sh = shape_for<TC>(op, nrhs);
    if (!shape_fits<TC>(op, sh))
        fail(RT_EINVAL, "rooftop:product", "the FFT columns of this product (%d x %d) exceed the LDS of this device",
             op.dims.Mx, op.dims.My);
    run_passes<TC>(op, in, out, work, nrhs, sh, s, skip, skip_stride);
}

void operator_apply(Product &op, const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s, const int *skip,
                    int skip_stride) {
    if (prec == RT_MIXED) prec = RT_C128;
    if (!op.has(prec)) fail(RT_EINVAL, "rooftop:product", "the product does not hold the requested precision");
    if (prec == RT_C128) apply_as<double>(op, in, out, nrhs, s, skip, skip_stride);
    else apply_as<float>(op, in, out, nrhs, s, skip, skip_stride);
}

// The matvec of nrhs right-hand sides with a caller's working memory (work, work_bytes), allocating nothing: the
// launch shape the product uses for nrhs, every right-hand side on stream s (no chunks on the side stream, whose
// results are bitwise those of one chunk), no kernel timing. Products whose pass B needs its full-period spectrum
// buffer, working memory beyond work_bytes, or pass-B spectra of a plan not built yet are refused.
template <typename TC>
static void apply_noalloc_as(Product &op, const c128 *in, c128 *out, int nrhs, hipStream_t s, void *work,
                             size_t work_bytes) {
    const PassShape sh = shape_for<TC>(op, nrhs);
    if (!shape_fits<TC>(op, sh))
        fail(RT_EINVAL, "rooftop:product", "the FFT columns of this product (%d x %d) exceed the LDS of this device",
             op.dims.Mx, op.dims.My);
    if (!b_fits<TC>(op, sh))
        fail(RT_EINVAL, "rooftop:product:noAlloc",
             "the %d-point y transforms of this product need the full-period spectrum buffer of the three-launch pass "
             "B, which the allocation-free application does not have", op.dims.My);
    const size_t need = static_cast<size_t>(2) * op.dims.ldj * op.dims.Mx * nrhs * sizeof(cplx<TC>);
    if (need > work_bytes)
        fail(RT_EINVAL, "rooftop:product:noAlloc",
             "%d right-hand sides of this product need %zu bytes of working memory, more than the %zu bytes of the "
             "allocation-free application", nrhs, need, work_bytes);
    const cplx<TC> *bspec;
    if (op.four_plane) {
        bspec = reinterpret_cast<const cplx<TC> *>(sizeof(TC) == sizeof(double)
                                                       ? static_cast<const void *>(op.b4spec64.ptr)
                                                       : static_cast<const void *>(op.b4spec32.ptr));
    } else {
        const int v = sh.bits >= 4 ? 1 : 0;
        bspec = reinterpret_cast<const cplx<TC> *>(sizeof(TC) == sizeof(double)
                                                       ? static_cast<const void *>(op.bspec64[v].ptr)
                                                       : static_cast<const void *>(op.bspec32[v].ptr));
    }
    if (!bspec)
        fail(RT_EINVAL, "rooftop:product:noAlloc",
             "the pass-B spectra of this launch shape are not built yet: apply the product once with "
             "rt_product_apply_device for this many right-hand sides first");
    cplx<TC> *w = static_cast<cplx<TC> *>(work);
    const PlanTables &tx = op.tx.of(sh.bits), &ty = op.ty.of(sh.bits);
    launch_pass_a<TC>(op, tx, in, w, nrhs, sh.group, s, nullptr, 0, nullptr);
    launch_pass_b<TC>(op, tx, ty, w, bspec, nrhs, sh.tile, sh.teams, s, nullptr, 0, nullptr);
    launch_pass_c<TC>(op, tx, w, out, nrhs, sh.group, s, nullptr, 0, nullptr);
}

void operator_apply_noalloc(Product &op, const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s, void *work,
                            size_t work_bytes) {
    if (prec == RT_MIXED) prec = RT_C128;
    if (!op.has(prec)) fail(RT_EINVAL, "rooftop:product", "the product does not hold the requested precision");
    if (prec == RT_C128) apply_noalloc_as<double>(op, in, out, nrhs, s, work, work_bytes);
    else apply_noalloc_as<float>(op, in, out, nrhs, s, work, work_bytes);
}

// The spectra kernels are instantiated per largest radix E of their plan.
#define RT_BY_E(E_VALUE, CALL)                      \
    switch (E_VALUE) {                              \
    case 1: { constexpr int E = 1; CALL; } break;   \
    case 2: { constexpr int E = 2; CALL; } break;   \
    case 4: { constexpr int E = 4; CALL; } break;   \
    case 8: { constexpr int E = 8; CALL; } break;   \
    default: { constexpr int E = 16; CALL; } break; \
    }

ProdDims operator_dims(const GridInfo &g) {
    ProdDims d{};
    d.rows_x = g.rows_x;
    d.rows_y = g.rows_y;
    d.cols_x = g.has_x ? g.cols_x : 0;
    d.cols_y = g.has_y ? g.cols_y : 0;
    d.ny = d.cols_x > d.cols_y ? d.cols_x : d.cols_y;
    d.ldj = (d.ny + 7) / 8 * 8;
    d.nx_count = g.nx;
    d.n = g.n;
    d.Mx = g.cx > 1 ? next_pow2(2L * g.cx - 1) : 1;
    d.My = g.cy > 1 ? next_pow2(2L * g.cy - 1) : 1;
    d.Hx = d.Mx / 2 + 1;
    d.Hy = d.My / 2 + 1;
    return d;
}

std::unique_ptr<Product> create_operator(Problem *p, int precision, int modifier) {
    Context *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    const GridInfo &g = p->g;
    auto op = std::make_unique<Product>();
    op->type = kOperator;
    op->problem = p;
    op->precision = precision;
    op->modifier = modifier;
    op->dims = operator_dims(g);
    const ProdDims &d = op->dims;
    const uint64_t plane_elems = static_cast<uint64_t>(d.Hx) * d.Hy;
    const uint64_t temp = static_cast<uint64_t>(d.Mx) * d.My * sizeof(c128);
    ctx->require(temp, "The spectra of the product (transform buffer)");
    p->ensure_generators();
    op->tx.build(ctx, d.Mx, g.a, s);
    op->ty.build(ctx, d.My, g.b, s);
    op->planes64.allocate(ctx, 3 * plane_elems, "The spectra of the product");
    op->planes64.zero(s);
    {
        const PlanTables &tx = op->tx.of(3), &ty = op->ty.of(3);
        DeviceArray<c128> tmp;
        tmp.allocate(ctx, static_cast<size_t>(d.Mx) * d.My, "The spectra of the product (transform buffer)");
        std::vector<int> fx = output_frequencies(tx.plan), fy = output_frequencies(ty.plan);
        DeviceArray<int> dfx, dfy;
        dfx.allocate(ctx, d.Mx, "FFT frequency table");
        dfy.allocate(ctx, d.My, "FFT frequency table");
        RT_HIP(hipMemcpyAsync(dfx.ptr, fx.data(), d.Mx * sizeof(int), hipMemcpyHostToDevice, s));
        RT_HIP(hipMemcpyAsync(dfy.ptr, fy.data(), d.My * sizeof(int), hipMemcpyHostToDevice, s));
        const double scale = 1.0 / (static_cast<double>(d.Mx) * d.My);
        const c128 *gens[3] = {g.has_x ? p->mxx.ptr : nullptr, g.has_y ? p->myy.ptr : nullptr, p->e.ptr};
        const int rows[3] = {g.cx - 1, g.cx, g.cx};
        const int cols[3] = {g.cy, g.cy - 1, g.cy};
        const size_t la = static_cast<size_t>(padded_length(tx.plan)) * sizeof(c128);
        const size_t lb = static_cast<size_t>(padded_length(ty.plan)) * sizeof(c128);
        const int ex = 1 << tx.plan.log2E, ey = 1 << ty.plan.log2E;
        for (int k = 0; k < 3; ++k) {
            if (!gens[k]) continue;
            RT_BY_E(ex, ({
                        set_lds(reinterpret_cast<const void *>(&spectra_a<E>), la);
                        hipLaunchKernelGGL((spectra_a<E>), dim3(d.My), dim3(std::max(64, team_threads(tx.plan))), la,
                                           s, gens[k], rows[k], cols[k], modifier, tmp.ptr, d.Mx, d.My, tx.plan,
                                           tx.tw.ptr);
                    }));
            RT_LAUNCH_CHECK();
            RT_BY_E(ey, ({
                        set_lds(reinterpret_cast<const void *>(&spectra_b<E>), lb);
                        hipLaunchKernelGGL((spectra_b<E>), dim3(d.Mx), dim3(std::max(64, team_threads(ty.plan))), lb,
                                           s, tmp.ptr, op->planes64.ptr + k * plane_elems, d.Mx, d.My, d.Hy, ty.plan,
                                           ty.tw.ptr, dfx.ptr, dfy.ptr, scale);
                    }));
            RT_LAUNCH_CHECK();
        }
        RT_HIP(hipStreamSynchronize(s));
    }
    // The pass-B spectra of the default plans of the precisions held.
    if (op->has(RT_C128)) bspec_of<double>(*op, default_bits(op->dims, false), s);
    if (op->has(RT_C64)) bspec_of<float>(*op, default_bits(op->dims, true), s);
    RT_HIP(hipStreamSynchronize(s));
    return op;
}

// ------------------------------------------------------------------------------------------------------------------
// Circulant preconditioners (Preconditioner.m:7-70, +preconditioner/generateCirculant.m:16-24 with the partwise
// builder generatePartwiseSpectra.m:16-30): every generator, zero-padded to the countX x countY face grid, is folded
// onto that period by the window of a kernel (foldGenerator.m:27-61: Strang, T. Chan, R. Chan, Huckle); the block
// spectra of the three circulants follow as for the Green's matrix (generateBlockSpectra.m) and are inverted frequency
// by frequency (@Spectra/invert.m); the product with the inverse runs the matvec passes on the face grid.
// ------------------------------------------------------------------------------------------------------------------
// Weight of the offset o of a period P (the windows of +kernels/generate*.m).
__device__ double circulant_window(int o, int P, int kind, double bandwidth) {
    const int a = o < 0 ? -o : o;
    switch (kind) {
    case RT_CIRCULANT_STRANG: return 2 * a < P ? 1.0 : (2 * a == P ? 0.5 : 0.0);
    case RT_CIRCULANT_TCHAN: return static_cast<double>(P - a) / P;
    case RT_CIRCULANT_RCHAN: return 1.0;
    default: return fmax(0.0, 1.0 - a / bandwidth);   // Huckle
    }
}

// Value of an even generator (rows x cols, column-major, zero-padded) at the offset (|p|, |q|), after the modifier.
__device__ __forceinline__ c128 generator_at(const c128 *__restrict__ g, int rows, int cols, int p, int q,
                                              int modifier) {
    if (p < 0) p = -p;
    if (q < 0) q = -q;
    if (p >= rows || q >= cols) return mk(0.0, 0.0);
    c128 v = g[static_cast<int64_t>(q) * rows + p];
    if (modifier == RT_MODIFIER_REAL) v.y = 0;
    if (modifier == RT_MODIFIER_IMAG) v = mk(v.y, 0.0);
    return v;
}

// The two-sided spatial generator of a diagonal block at the offset (p, q) (generateSpatialBlocks.m:19-61):
// XX(p, q) = mXX(p, q) - (2 e(p, q) - e(p - 1, q) - e(p + 1, q)) / a^2, zero beyond |p| = P - 2; YY along q with b.
// block 1: XX from g = mXX ((P-1) x Q); block 2: YY from g = mYY (P x (Q-1)); e is P x Q.
__device__ __forceinline__ c128 spatial_block_at(int block, const c128 *__restrict__ g, const c128 *__restrict__ e,
                                                 int P, int Q, double size, int p, int q, int modifier) {
    if (block == 1) {
        if ((p < 0 ? -p : p) > P - 2) return mk(0.0, 0.0);
        const c128 d = 2.0 * generator_at(e, P, Q, p, q, modifier) - generator_at(e, P, Q, p - 1, q, modifier) -
                       generator_at(e, P, Q, p + 1, q, modifier);
        return generator_at(g, P - 1, Q, p, q, modifier) - (1.0 / (size * size)) * d;
    }
    if ((q < 0 ? -q : q) > Q - 2) return mk(0.0, 0.0);
    const c128 d = 2.0 * generator_at(e, P, Q, p, q, modifier) - generator_at(e, P, Q, p, q - 1, modifier) -
                   generator_at(e, P, Q, p, q + 1, modifier);
    return generator_at(g, P, Q - 1, p, q, modifier) - (1.0 / (size * size)) * d;
}

// c(kx, ky) for kx <= P/2, ky <= Q/2 (the rest follows from the symmetry c(k) = c(P - k) of the fold of an even
// generator): the sum over the offsets ox in {kx, kx - P} (only 0 for kx = 0), oy likewise, of
// w(ox) w(oy) t(ox, oy). block 0: t = the generator g (rows x cols, column-major) zero-padded to P x Q; blocks 1, 2:
// t = the spatial XX or YY block of g and the electric generator e (the hybrid builder). Output [Hy][Hx].
__global__ void circulant_fold(const c128 *__restrict__ g, int rows, int cols, int P, int Q, int kind, double bwx,
                               double bwy, int modifier, c128 *__restrict__ out, int Hx, int Hy, int block,
                               const c128 *__restrict__ e, double size) {
    const int64_t total = static_cast<int64_t>(Hx) * Hy;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int kx = static_cast<int>(i % Hx), ky = static_cast<int>(i / Hx);
        c128 sum = mk(0.0, 0.0);
        for (int sx = 0; sx < (kx == 0 ? 1 : 2); ++sx) {
            const int ox = sx ? kx - P : kx;
            const double wx = circulant_window(ox, P, kind, bwx);
            for (int sy = 0; sy < (ky == 0 ? 1 : 2); ++sy) {
                const int oy = sy ? ky - Q : ky;
                const double w = wx * circulant_window(oy, Q, kind, bwy);
                if (w == 0) continue;
                const c128 v = block == 0 ? generator_at(g, rows, cols, ox, oy, modifier)
                                          : spatial_block_at(block, g, e, P, Q, size, ox, oy, modifier);
                sum = sum + w * v;
            }
        }
        out[i] = sum;
    }
}

// In place on the folded planes [3][Hx][Hy] (the circulants' spectra A, B, E; with blocks_given the first two are the
// diagonal blocks XX, YY themselves, the hybrid builder's): the inverse of the 2x2 block
// [XX, XY; YX, YY] (XX = A - |Dx|^2 E, YY = B - |Dy|^2 E, XY = -Dx conj(Dy) E, YX = -conj(Dx) Dy E) written in the
// same three-plane form as the product, (A', B', E') = ((YY - |Dx|^2 E) / det, (XX - |Dy|^2 E) / det, -E / det), so
// that the product's passes apply it; 1 / XX (1 / YY) alone for a strip. Scaled by 1/(Mx My) for the inverse FFT.
// A zero or non-finite determinant raises *singular (the reference's allfinite test of the inverse).
__device__ __forceinline__ c128 reciprocal(c128 z) {
    const double m2 = z.x * z.x + z.y * z.y;
    return mk(z.x / m2, -z.y / m2);
}
__global__ void circulant_invert(c128 *__restrict__ planes, int Hx, int Hy, int Mx, int My, double a, double b,
                                 int has_x, int has_y, int blocks_given, double scale, int *__restrict__ singular) {
    const int64_t plane = static_cast<int64_t>(Hx) * Hy;
    const double pi = 3.141592653589793;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < plane;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int fx = static_cast<int>(i / Hy), fy = static_cast<int>(i - static_cast<int64_t>(fx) * Hy);
        const double sx = 2 * sin(pi * fx / Mx) / a, sy = 2 * sin(pi * fy / My) / b;
        const double dx2 = sx * sx, dy2 = sy * sy;
        const c128 A = planes[i], B = planes[plane + i], E = planes[2 * plane + i];
        c128 a2 = mk(0.0, 0.0), b2 = mk(0.0, 0.0), e2 = mk(0.0, 0.0), det;
        if (has_x && has_y) {
            const c128 xx = blocks_given ? A : A - dx2 * E, yy = blocks_given ? B : B - dy2 * E;
            det = xx * yy - (dx2 * dy2) * (E * E);
            const c128 inv = reciprocal(det);
            a2 = (yy - dx2 * E) * inv;
            b2 = (xx - dy2 * E) * inv;
            e2 = mk(0.0, 0.0) - E * inv;
        } else if (has_x) {
            det = blocks_given ? A : A - dx2 * E;
            a2 = reciprocal(det);
        } else {
            det = blocks_given ? B : B - dy2 * E;
            b2 = reciprocal(det);
        }
        const bool ok = (det.x != 0 || det.y != 0) && isfinite(a2.x) && isfinite(a2.y) && isfinite(b2.x) &&
                        isfinite(b2.y) && isfinite(e2.x) && isfinite(e2.y);
        if (!ok) atomicExch(singular, 1);
        planes[i] = scale * a2;
        planes[plane + i] = scale * b2;
        planes[2 * plane + i] = scale * e2;
    }
}

// pc->b4spec{64,32}[(px My + py) 4 + c] = block c of the inverse at (freq_x[px], freq_y[py]) from planes
// [4][Mx][My] (unfolded frequencies), in the precision of the pass.
template <typename TC>
__global__ void build_b4spec(const c128 *__restrict__ planes, int Mx, int My, const int *__restrict__ freq_x,
                             const int *__restrict__ freq_y, cplx<TC> *__restrict__ out) {
    const int64_t total = static_cast<int64_t>(Mx) * My;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int px = static_cast<int>(i / My), py = static_cast<int>(i - static_cast<int64_t>(px) * My);
        const int64_t f = static_cast<int64_t>(freq_x[px]) * My + freq_y[py];
        for (int c = 0; c < 4; ++c) out[4 * i + c] = cvt<TC>(planes[c * total + f]);
    }
}

// The generators on the host, after the modifier, for the spectra formed there (circulant_host.cpp).
static std::vector<cd> host_generator(const DeviceArray<c128> &g, size_t count, int modifier, hipStream_t s) {
    std::vector<c128> raw(count);
    RT_HIP(hipMemcpyAsync(raw.data(), g.ptr, sizeof(c128) * count, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    std::vector<cd> out(count);
    for (size_t i = 0; i < count; ++i) {
        cd v(raw[i].x, raw[i].y);
        if (modifier == RT_MODIFIER_REAL) v = cd(v.real(), 0.0);
        if (modifier == RT_MODIFIER_IMAG) v = cd(v.imag(), 0.0);
        out[i] = v;
    }
    return out;
}

std::unique_ptr<Product> create_circulant(Problem *p, int kernel, int builder, int bandwidth_x, int bandwidth_y,
                                          int modifier, int precision) {
    Context *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    const GridInfo &g = p->g;
    if (builder < RT_BUILDER_PARTWISE || builder > RT_BUILDER_HYBRID || kernel < RT_CIRCULANT_STRANG ||
        kernel > RT_CIRCULANT_SUPEROPTIMAL)
        fail(RT_EINVAL, "rooftop:preconditioner:unsupported",
             "The circulant preconditioner takes the partwise, blockwise or hybrid builder with the strang, tChan, "
             "rChan, huckle or superoptimal kernel.");
    const bool hybrid = builder == RT_BUILDER_HYBRID;
    const bool blockwise = builder == RT_BUILDER_BLOCKWISE;
    auto pow2 = [](int v) { return v >= 1 && (v & (v - 1)) == 0; };
    // On a face grid of power-of-two counts the passes run on the face grid itself; on any other the inverse's block
    // circulants are re-embedded as Toeplitz products (generator c(p mod cx, q mod cy) at the offsets |p| < cx,
    // |q| < cy) and applied by the operator's power-of-two embedding, whose inputs are zero and outputs cut beyond the
    // face grid, so the product is the circulant product.
    const bool face_grid = pow2(g.cx) && pow2(g.cy);
    const bool on_host = blockwise || kernel == RT_CIRCULANT_SUPEROPTIMAL || !face_grid;
    if (kernel == RT_CIRCULANT_HUCKLE && (bandwidth_x < 1 || bandwidth_y < 1))
        fail(RT_EINVAL, "rooftopMoM:preconditioner:invalidBandwidth", "The bandwidth is a pair of positive integers.");
    auto pc = std::make_unique<Product>();
    pc->type = kCirculant;
    pc->embedding = !face_grid;
    pc->problem = p;
    pc->precision = precision;
    pc->modifier = modifier;
    ProdDims &d = pc->dims;
    d.rows_x = g.rows_x;
    d.rows_y = g.rows_y;
    d.cols_x = g.has_x ? g.cols_x : 0;
    d.cols_y = g.has_y ? g.cols_y : 0;
    d.ny = d.cols_x > d.cols_y ? d.cols_x : d.cols_y;
    d.ldj = (d.ny + 7) / 8 * 8;
    d.nx_count = g.nx;
    d.n = g.n;
    d.Mx = face_grid ? g.cx : (g.cx > 1 ? next_pow2(2L * g.cx - 1) : 1);
    d.My = face_grid ? g.cy : (g.cy > 1 ? next_pow2(2L * g.cy - 1) : 1);
    d.Hx = d.Mx / 2 + 1;
    d.Hy = d.My / 2 + 1;
    const uint64_t plane_elems = static_cast<uint64_t>(d.Hx) * d.Hy;
    ctx->require(static_cast<uint64_t>(d.Mx) * d.My * sizeof(c128), "The circulant preconditioner (transform buffer)");
    p->ensure_generators();
    pc->tx.build(ctx, d.Mx, g.a, s);
    pc->ty.build(ctx, d.My, g.b, s);
    if (on_host) {
        // Spectra formed on the host: the generators (after the modifier), their or their spatial blocks'
        // circulants, then the inverse.
        HostGenerators hg;
        hg.cx = g.cx;
        hg.cy = g.cy;
        hg.a = g.a;
        hg.b = g.b;
        hg.has_x = g.has_x;
        hg.has_y = g.has_y;
        hg.e = host_generator(p->e, static_cast<size_t>(g.cx) * g.cy, modifier, s);
        if (g.has_x) hg.mxx = host_generator(p->mxx, static_cast<size_t>(g.cx - 1) * g.cy, modifier, s);
        if (g.has_y) hg.myy = host_generator(p->myy, static_cast<size_t>(g.cx) * (g.cy - 1), modifier, s);
        const double bwx = bandwidth_x, bwy = bandwidth_y;
        auto upload = [&](const std::vector<cd> &h, c128 *dst) {
            std::vector<c128> raw(h.size());
            for (size_t i = 0; i < h.size(); ++i) raw[i] = mk(h[i].real(), h[i].imag());
            RT_HIP(hipMemcpyAsync(dst, raw.data(), sizeof(c128) * raw.size(), hipMemcpyHostToDevice, s));
            RT_HIP(hipStreamSynchronize(s));
        };
        if (blockwise || !face_grid) {
            // The inverse blocks (XX', XY', YX', YY') per frequency of the cx x cy period, kept unfolded as four
            // planes for the four-plane multiply of pass B: on the face grid directly, or re-embedded.
            std::array<std::vector<cd>, 4> blocks = circulant_blocks(hg, kernel, builder, bwx, bwy);
            if (!invert_blocks(blocks, g.has_x, g.has_y))
                fail(RT_ESINGULAR, "rooftopMoM:preconditioner:singular",
                     "The circulant approximation is singular at some frequency. Use another kernel or builder.");
            if (!b_fits<double>(*pc, PassShape{3, 1, 1, 1, 1}))
                fail(RT_EINVAL, "rooftop:preconditioner:unsupported",
                     "The four-plane circulant product needs columns of %d that fit one block's LDS.", d.My);
            const int64_t full = static_cast<int64_t>(d.Mx) * d.My;
            const double scale = 1.0 / static_cast<double>(full);
            std::vector<cd> planes(4 * full, 0.0);   // [c][fx My + fy]
            for (int c = 0; c < 4; ++c) {
                if (blocks[c].empty()) continue;
                std::vector<cd> spec = std::move(blocks[c]);   // [fx + cx fy]
                if (!face_grid) {
                    fft2_host(spec, g.cx, g.cy, true);   // the block circulant's generator
                    std::vector<cd> emb(full, 0.0);
                    for (int q = -(g.cy - 1); q < g.cy; ++q)
                        for (int pp = -(g.cx - 1); pp < g.cx; ++pp)
                            emb[(pp + d.Mx) % d.Mx + static_cast<size_t>(d.Mx) * ((q + d.My) % d.My)] =
                                spec[(pp + g.cx) % g.cx + static_cast<size_t>(g.cx) * ((q + g.cy) % g.cy)];
                    fft2_host(emb, d.Mx, d.My, false);
                    spec = std::move(emb);
                }
                for (int fx = 0; fx < d.Mx; ++fx)
                    for (int fy = 0; fy < d.My; ++fy)
                        planes[c * full + static_cast<int64_t>(fx) * d.My + fy] = scale * spec[fx + static_cast<size_t>(d.Mx) * fy];
            }
            DeviceArray<c128> planes4;
            planes4.allocate(ctx, 4 * full, "The circulant preconditioner (blocks)");
            upload(planes, planes4.ptr);
            const PlanTables &tx = pc->tx.of(3), &ty = pc->ty.of(3);
            const std::vector<int> fx = output_frequencies(tx.plan), fy = output_frequencies(ty.plan);
            DeviceArray<int> dfx, dfy;
            dfx.allocate(ctx, d.Mx, "FFT frequency table");
            dfy.allocate(ctx, d.My, "FFT frequency table");
            RT_HIP(hipMemcpyAsync(dfx.ptr, fx.data(), d.Mx * sizeof(int), hipMemcpyHostToDevice, s));
            RT_HIP(hipMemcpyAsync(dfy.ptr, fy.data(), d.My * sizeof(int), hipMemcpyHostToDevice, s));
            pc->four_plane = true;
            if (pc->has(RT_C128)) {
                pc->b4spec64.allocate(ctx, 4 * full, "The circulant preconditioner");
                hipLaunchKernelGGL(build_b4spec<double>, dim3(blocks_for(full, 256, 4096)), dim3(256), 0, s,
                                   planes4.ptr, d.Mx, d.My, dfx.ptr, dfy.ptr, pc->b4spec64.ptr);
                RT_LAUNCH_CHECK();
            }
            if (pc->has(RT_C64)) {
                pc->b4spec32.allocate(ctx, 4 * full, "The circulant preconditioner");
                hipLaunchKernelGGL(build_b4spec<float>, dim3(blocks_for(full, 256, 4096)), dim3(256), 0, s,
                                   planes4.ptr, d.Mx, d.My, dfx.ptr, dfy.ptr, pc->b4spec32.ptr);
                RT_LAUNCH_CHECK();
            }
            RT_HIP(hipStreamSynchronize(s));
            return pc;
        }
        // The superoptimal kernel with the partwise or hybrid builder: the three-plane form as for the folds, the
        // spectra formed on the host (even for even generators: their [0, M/2] quarter is kept).
        pc->planes64.allocate(ctx, 3 * plane_elems, "The circulant preconditioner");
        pc->planes64.zero(s);
        std::vector<cd> folded(plane_elems);
        for (int k = 0; k < 3; ++k) {
            if ((k == 0 && !g.has_x) || (k == 1 && !g.has_y)) continue;
            TwoSided t;
            if (hybrid && k < 2) {
                t = spatial_block(k + 1, hg);
            } else {
                const std::vector<cd> &gen = k == 0 ? hg.mxx : (k == 1 ? hg.myy : hg.e);
                t = even_extension(gen, k == 0 ? g.cx - 1 : g.cx, k == 1 ? g.cy - 1 : g.cy, g.cx, g.cy);
            }
            const std::vector<cd> spec = superoptimal_spectrum(t);
            for (int fx = 0; fx < d.Hx; ++fx)
                for (int fy = 0; fy < d.Hy; ++fy)
                    folded[static_cast<size_t>(fx) * d.Hy + fy] = spec[fx + static_cast<size_t>(d.Mx) * fy];
            upload(folded, pc->planes64.ptr + k * plane_elems);
        }
    } else {
        pc->planes64.allocate(ctx, 3 * plane_elems, "The circulant preconditioner");
        pc->planes64.zero(s);
    }
    {
        const PlanTables &tx = pc->tx.of(3), &ty = pc->ty.of(3);
        DeviceArray<c128> tmp, folded;
        tmp.allocate(ctx, static_cast<size_t>(d.Mx) * d.My, "The circulant preconditioner (transform buffer)");
        folded.allocate(ctx, plane_elems, "The circulant preconditioner (fold)");
        std::vector<int> fx = output_frequencies(tx.plan), fy = output_frequencies(ty.plan);
        DeviceArray<int> dfx, dfy;
        dfx.allocate(ctx, d.Mx, "FFT frequency table");
        dfy.allocate(ctx, d.My, "FFT frequency table");
        RT_HIP(hipMemcpyAsync(dfx.ptr, fx.data(), d.Mx * sizeof(int), hipMemcpyHostToDevice, s));
        RT_HIP(hipMemcpyAsync(dfy.ptr, fy.data(), d.My * sizeof(int), hipMemcpyHostToDevice, s));
        const c128 *gens[3] = {g.has_x ? p->mxx.ptr : nullptr, g.has_y ? p->myy.ptr : nullptr, p->e.ptr};
        const int rows[3] = {g.cx - 1, g.cx, g.cx};
        const int cols[3] = {g.cy, g.cy - 1, g.cy};
        const size_t la = static_cast<size_t>(padded_length(tx.plan)) * sizeof(c128);
        const size_t lb = static_cast<size_t>(padded_length(ty.plan)) * sizeof(c128);
        const int ex = 1 << tx.plan.log2E, ey = 1 << ty.plan.log2E;
        for (int k = 0; k < 3 && !on_host; ++k) {
            if (!gens[k]) continue;
            // Partwise: the folds of the generators; hybrid (generateHybridSpectra.m): the folds of the spatial
            // diagonal blocks XX, YY (generateSpatialBlocks.m) and of the electric generator.
            const int block = hybrid && k < 2 ? k + 1 : 0;
            hipLaunchKernelGGL(circulant_fold, dim3(blocks_for(plane_elems, 256, 4096)), dim3(256), 0, s, gens[k],
                               rows[k], cols[k], d.Mx, d.My, kernel, static_cast<double>(bandwidth_x),
                               static_cast<double>(bandwidth_y), modifier, folded.ptr, d.Hx, d.Hy, block, p->e.ptr,
                               k == 0 ? g.a : g.b);
            RT_LAUNCH_CHECK();
            // The fold is even on the period, so the product's spectrum kernels transform it as the even extension
            // of its half [0, P/2] x [0, Q/2].
            RT_BY_E(ex, ({
                        set_lds(reinterpret_cast<const void *>(&spectra_a<E>), la);
                        hipLaunchKernelGGL((spectra_a<E>), dim3(d.My), dim3(std::max(64, team_threads(tx.plan))), la,
                                           s, folded.ptr, d.Hx, d.Hy, RT_MODIFIER_NONE, tmp.ptr, d.Mx, d.My, tx.plan,
                                           tx.tw.ptr);
                    }));
            RT_LAUNCH_CHECK();
            RT_BY_E(ey, ({
                        set_lds(reinterpret_cast<const void *>(&spectra_b<E>), lb);
                        hipLaunchKernelGGL((spectra_b<E>), dim3(d.Mx), dim3(std::max(64, team_threads(ty.plan))), lb,
                                           s, tmp.ptr, pc->planes64.ptr + k * plane_elems, d.Mx, d.My, d.Hy, ty.plan,
                                           ty.tw.ptr, dfx.ptr, dfy.ptr, 1.0);
                    }));
            RT_LAUNCH_CHECK();
        }
        DeviceArray<int> flag;
        flag.allocate(ctx, 1, "singularity flag");
        flag.zero(s);
        hipLaunchKernelGGL(circulant_invert, dim3(blocks_for(plane_elems, 256, 4096)), dim3(256), 0, s,
                           pc->planes64.ptr, d.Hx, d.Hy, d.Mx, d.My, g.a, g.b, g.has_x ? 1 : 0, g.has_y ? 1 : 0,
                           hybrid ? 1 : 0, 1.0 / (static_cast<double>(d.Mx) * d.My), flag.ptr);
        RT_LAUNCH_CHECK();
        int singular = 0;
        RT_HIP(hipMemcpyAsync(&singular, flag.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        if (singular)
            fail(RT_ESINGULAR, "rooftopMoM:preconditioner:singular",
                 "The circulant approximation is singular at some frequency. Use another kernel or builder.");
    }
    if (pc->has(RT_C128)) bspec_of<double>(*pc, 3, s);
    if (pc->has(RT_C64)) bspec_of<float>(*pc, 3, s);
    RT_HIP(hipStreamSynchronize(s));
    return pc;
}

// ------------------------------------------------------------------------------------------------------------------
// Measurement and tuning.
// ------------------------------------------------------------------------------------------------------------------
// Average per-launch milliseconds and algorithmic bytes of each pass over `reps` applications with shape sh (one
// untimed application first), through the context's kernel timing.
template <typename TC>
static void time_shape(Product &op, int nrhs, int reps, const PassShape &sh, const c128 *in, c128 *out, double *ms,
                       double *bytes, hipStream_t s, double *wall = nullptr) {
    Context *ctx = op.problem->ctx;
    cplx<TC> *work = work_of<TC>(op, nrhs);
    run_passes<TC>(op, in, out, work, nrhs, sh, s);
    RT_HIP(hipStreamSynchronize(s));
    const bool was = ctx->timing;
    ctx->drain_timing();
    const int slot0 = sizeof(TC) == sizeof(float) ? RT_TIMED_PASS_A_C64 : RT_TIMED_PASS_A_C128;
    int64_t n0[3];
    double ms0[3], by0[3];
    for (int k = 0; k < 3; ++k) {
        n0[k] = ctx->timed_launches[slot0 + k];
        ms0[k] = ctx->timed_ms[slot0 + k];
        by0[k] = ctx->timed_bytes[slot0 + k];
    }
    ctx->timing = true;
    for (int it = 0; it < reps; ++it) run_passes<TC>(op, in, out, work, nrhs, sh, s);
    ctx->drain_timing();
    ctx->timing = false;
    if (wall) {
        // The wall time of back-to-back applications, with the kernel timing off.
        hipEvent_t w0 = ctx->take_event(), w1 = ctx->take_event();
        RT_HIP(hipEventRecord(w0, s));
        for (int it = 0; it < reps; ++it) run_passes<TC>(op, in, out, work, nrhs, sh, s);
        RT_HIP(hipEventRecord(w1, s));
        RT_HIP(hipEventSynchronize(w1));
        float wall_ms = 0;
        RT_HIP(hipEventElapsedTime(&wall_ms, w0, w1));
        ctx->event_pool.push_back(w0);
        ctx->event_pool.push_back(w1);
        *wall = wall_ms / reps;
    }
    ctx->timing = was;
    // Algorithmic bytes of one application, whatever its chunks (each chunk's launch of pass B reads the spectra
    // again, which is traffic, not algorithmic bytes).
    double app[3];
    pass_bytes<TC>(op.dims, nrhs, app, op.four_plane);
    for (int k = 0; k < 3; ++k) {
        // Per application: the chunks of one application add up.
        ms[k] = (ctx->timed_ms[slot0 + k] - ms0[k]) / reps;
        bytes[k] = app[k];
        ctx->timed_launches[slot0 + k] = n0[k];
        ctx->timed_ms[slot0 + k] = ms0[k];
        ctx->timed_bytes[slot0 + k] = by0[k];
    }
}

template <typename TC>
static void profile_as(Product &op, int nrhs, int reps, double *ms, double *bytes, hipStream_t s) {
    Context *ctx = op.problem->ctx;
    DeviceArray<c128> in, out;
    in.allocate(ctx, static_cast<size_t>(op.dims.n) * nrhs, "profile input");
    out.allocate(ctx, static_cast<size_t>(op.dims.n) * nrhs, "profile output");
    in.zero(s);
    const PassShape sh = shape_for<TC>(op, nrhs);
    if (!shape_fits<TC>(op, sh))
        fail(RT_EINVAL, "rooftop:product", "the FFT columns of this product (%d x %d) exceed the LDS of this device",
             op.dims.Mx, op.dims.My);
    time_shape<TC>(op, nrhs, reps, sh, in.ptr, out.ptr, ms, bytes, s, &ms[3]);
}

void operator_profile(Product &op, int nrhs, int reps, int prec, double *ms, double *bytes, hipStream_t s) {
    if (op.type != kOperator) fail(RT_EINVAL, "rooftop:product", "only an operator has matvec passes to profile");
    if (prec == RT_MIXED) prec = RT_C128;
    if (!op.has(prec)) fail(RT_EINVAL, "rooftop:product", "the product does not hold the requested precision");
    if (prec == RT_C128) profile_as<double>(op, nrhs, reps, ms, bytes, s);
    else profile_as<float>(op, nrhs, reps, ms, bytes, s);
}

// The autotuner: for each plan (radices up to 8 or up to 16) it times the pass A + C column groups with the default
// tile, then the pass B tiles with the best group, and keeps the fastest plan; the shape is recorded for this batch
// size. Pass A/C time depends on the group only and pass B time on the tile only. The inputs are zeros: the passes
// do the same work on any data. Each candidate runs enough applications for about 2 ms of passes (at least `reps`),
// twice, and keeps the faster measurement of each pass: a single right-hand side's passes last a few microseconds,
// and a choice made on one short sample is decided by noise.
template <typename TC>
static PassShape tune_as(Product &op, int nrhs, int reps, double *best_ms, hipStream_t s) {
    Context *ctx = op.problem->ctx;
    DeviceArray<c128> in, out;
    in.allocate(ctx, static_cast<size_t>(op.dims.n) * nrhs, "tuning input");
    out.allocate(ctx, static_cast<size_t>(op.dims.n) * nrhs, "tuning output");
    in.zero(s);
    PassShape best = default_shape<TC>(op, nrhs);
    {
        double ms[3], by[3], wall = 0;
        time_shape<TC>(op, nrhs, 3, best, in.ptr, out.ptr, ms, by, s, &wall);
        reps = std::max(reps, std::min(200, static_cast<int>(std::ceil(2.0 / std::max(wall, 1e-3)))));
    }
    // Per-pass milliseconds of shape c (and the wall time of one application): the faster of two measurements.
    auto measure = [&](const PassShape &c, double *ms, double *wall = nullptr) {
        double by[3], ms2[3], w2 = 0;
        time_shape<TC>(op, nrhs, reps, c, in.ptr, out.ptr, ms, by, s, wall);
        time_shape<TC>(op, nrhs, reps, c, in.ptr, out.ptr, ms2, by, s, wall ? &w2 : nullptr);
        for (int q = 0; q < 3; ++q) ms[q] = std::min(ms[q], ms2[q]);
        if (wall) *wall = std::min(*wall, w2);
    };
    double best_total = 1e300;
    for (int bits = 3; bits <= 4; ++bits) {
        PassShape sh = default_shape<TC>(op, nrhs, bits);
        double ac = 1e300, b = 1e300;
        int g_best = 0, t_best = 0, k_best = 1;
        for (int g : {1, 2, 4, 8, 16, -2, -4}) {   // side by side, then in turn
            PassShape c = sh;
            c.group = g;
            if (!shape_fits<TC>(op, c)) continue;
            double ms[3];
            measure(c, ms);
            if (ms[0] + ms[2] < ac) {
                ac = ms[0] + ms[2];
                g_best = g;
            }
        }
        if (!g_best) continue;
        sh.group = g_best;
        for (int teams = 1; teams <= 3; ++teams)
            for (int tile = 1; tile <= 16; tile *= 2) {
                PassShape c = sh;
                c.tile = tile;
                c.teams = teams;
                if (!shape_fits<TC>(op, c)) continue;
                double ms[3];
                measure(c, ms);
                if (ms[1] < b) {
                    b = ms[1];
                    t_best = tile;
                    k_best = teams;
                }
            }
        if (!t_best) continue;
        if (ac + b < best_total) {
            best_total = ac + b;
            best = PassShape{bits, g_best, t_best, k_best};
        }
    }
    // Then the chunk count, by the wall time of whole applications (the chunks overlap).
    {
        double ms[3], wall = 0;
        measure(best, ms, &wall);
        best_total = wall;
        const PassShape one = best;
        for (int chunks = 2; chunks <= 8 && chunks <= nrhs; chunks *= 2) {
            PassShape c = one;
            c.chunks = chunks;
            if (!b_fits<TC>(op, c)) break;
            measure(c, ms, &wall);
            if (wall < best_total) {
                best_total = wall;
                best = c;
            }
        }
    }
    auto &tuned = sizeof(TC) == sizeof(float) ? op.tuned32 : op.tuned64;
    bool found = false;
    for (auto &e : tuned)
        if (e.first == nrhs) {
            e.second = best;
            found = true;
        }
    if (!found) tuned.emplace_back(nrhs, best);
    if (best_ms) *best_ms = best_total;
    return best;
}

void operator_tune(Product &op, int nrhs, int prec, int *shape, double *ms, hipStream_t s) {
    if (op.type != kOperator) fail(RT_EINVAL, "rooftop:product", "only an operator has matvec passes to tune");
    if (prec == RT_MIXED) prec = RT_C128;
    if (!op.has(prec)) fail(RT_EINVAL, "rooftop:product", "the product does not hold the requested precision");
    const PassShape sh = prec == RT_C128 ? tune_as<double>(op, nrhs, 5, ms, s) : tune_as<float>(op, nrhs, 5, ms, s);
    if (shape) {
        shape[0] = sh.bits;
        shape[1] = sh.group;
        shape[2] = sh.tile;
        shape[3] = sh.teams;
        shape[4] = sh.chunks;
    }
}

void operator_set_shape(Product &op, int nrhs, int prec, const int *shape) {
    if (op.type != kOperator) fail(RT_EINVAL, "rooftop:product", "only an operator has matvec passes");
    if (prec == RT_MIXED) prec = RT_C128;
    if (!op.has(prec)) fail(RT_EINVAL, "rooftop:product", "the product does not hold the requested precision");
    const PassShape sh{shape[0], shape[1], shape[2], shape[3], shape[4]};
    auto pow2 = [](int v) { return v >= 1 && (v & (v - 1)) == 0; };
    if ((sh.bits != 3 && sh.bits != 4) || !(pow2(sh.group) || sh.group == -2 || sh.group == -4) || !pow2(sh.tile) ||
        sh.group > 16 || sh.tile > 16 || sh.teams < 1 || sh.teams > 3 || sh.chunks < 1 || sh.chunks > 64)
        fail(RT_EINVAL, "rooftop:product",
             "a shape is {3 or 4, a power of two <= 16 or -2 or -4, a power of two <= 16, 1 .. 3, 1 .. 64}");
    const bool fits = prec == RT_C128 ? shape_fits<double>(op, sh) : shape_fits<float>(op, sh);
    if (!fits) fail(RT_EINVAL, "rooftop:product", "the shape does not fit the device for this product");
    auto &tuned = prec == RT_C128 ? op.tuned64 : op.tuned32;
    for (auto &e : tuned)
        if (e.first == nrhs) {
            e.second = sh;
            return;
        }
    tuned.emplace_back(nrhs, sh);
}

uint64_t Product::stored_bytes() const {
    return planes64.bytes() + bspec64[0].bytes() + bspec64[1].bytes() + bspec32[0].bytes() + bspec32[1].bytes() +
           b4spec64.bytes() + b4spec32.bytes() + diag64.bytes() + diag32.bytes() + coef64.bytes() + coef32.bytes() + cls.bytes() + scls.bytes() +
           scoef64.bytes() + scoef32.bytes() + dense64.bytes() + dense32.bytes();
}

uint64_t Product::working_bytes() const {
    if (type != kOperator && type != kCirculant) return 0;
    const uint64_t per = static_cast<uint64_t>(2) * dims.ldj * dims.Mx;
    return per * (has(RT_C128) ? sizeof(c128) : sizeof(c64));
}

void Product::apply(const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s, const int *skip, int skip_stride) {
    if (type == kOperator || type == kCirculant)
        operator_apply(*this, in, out, nrhs, prec, s, skip, skip_stride);
    else if (type == kDense || type == kCallback)
        dense_apply(*this, in, out, nrhs, prec, s, skip, skip_stride);
    else
        preconditioner_apply(*this, in, out, nrhs, prec, s, skip, skip_stride);
}

}  // namespace rt

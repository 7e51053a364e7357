// The Toeplitz product of the rooftop Green's matrix on gfx950: the three generators embedded evenly into an
// Mx x My circulant (period >= 2n-1 per split direction), their spectra folded by their even symmetry, and the
// matvec as three LDS-staged FFT passes with the spectral multiply fused into the middle one.
//
// Reference: the products of src/+rooftopMoM/+assembler/+greensMatrix/FFTApplicator.m:14-87 and
// @SplitFFTApplicator/applyBothDirections_.m:14-85 with the block spectra of +utils/generateBlockSpectra.m:19-50.
// With Dx = (1 - sx)/a, Dy = (1 - sy)/b the four block spectra collapse into three planes:
//     out_x = Mxx X - Dx E (conj(Dx) X + conj(Dy) Y),   out_y = Myy Y - Dy E (conj(Dx) X + conj(Dy) Y),
// the x and y current spectra X, Y and the face-charge spectrum conj(Dx) X + conj(Dy) Y ("Jx, Jy and charge").
#include <cmath>

#include "rooftop.hpp"

namespace rt {

// ------------------------------------------------------------------------------------------------------------------
// Tables.
// ------------------------------------------------------------------------------------------------------------------
void DimTables::build(Context *ctx, int M, double size, hipStream_t s) {
    plan = make_plan(M);
    std::vector<c128> h_tw(M), h_d(M);
    std::vector<c64> h_tw32(M), h_d32(M);
    std::vector<int> h_fold(M);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int t = 0; t < M; ++t) {
        long double ang = -2.0L * pi * static_cast<long double>(t) / static_cast<long double>(M);
        h_tw[t] = {static_cast<double>(std::cos(ang)), static_cast<double>(std::sin(ang))};
        h_tw32[t] = {static_cast<float>(h_tw[t].x), static_cast<float>(h_tw[t].y)};
    }
    std::vector<int> freq = output_frequencies(plan);
    for (int p = 0; p < M; ++p) {
        int k = freq[p];
        h_fold[p] = k <= M - k ? k : M - k;
        long double ang = 2.0L * pi * static_cast<long double>(k) / static_cast<long double>(M);
        // D = (1 - exp(i ang)) / size; 1 - cos = 2 sin^2(ang/2) keeps low frequencies accurate.
        long double sh = std::sin(ang / 2);
        h_d[p] = {static_cast<double>(2.0L * sh * sh / size), static_cast<double>(-std::sin(ang) / size)};
        h_d32[p] = {static_cast<float>(h_d[p].x), static_cast<float>(h_d[p].y)};
    }
    tw.allocate(ctx, M, "FFT twiddles");
    tw32.allocate(ctx, M, "FFT twiddles");
    fold.allocate(ctx, M, "FFT fold table");
    d.allocate(ctx, M, "divergence table");
    d32.allocate(ctx, M, "divergence table");
    RT_HIP(hipMemcpyAsync(tw.ptr, h_tw.data(), M * sizeof(c128), hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(tw32.ptr, h_tw32.data(), M * sizeof(c64), hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(fold.ptr, h_fold.data(), M * sizeof(int), hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(d.ptr, h_d.data(), M * sizeof(c128), hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(d32.ptr, h_d32.data(), M * sizeof(c64), hipMemcpyHostToDevice, s));
    RT_HIP(hipStreamSynchronize(s));
}

template <typename T> const cplx<T> *tw_of(const DimTables &t);
template <> const c128 *tw_of<double>(const DimTables &t) { return t.tw.ptr; }
template <> const c64 *tw_of<float>(const DimTables &t) { return t.tw32.ptr; }
template <typename T> const cplx<T> *d_of(const DimTables &t);
template <> const c128 *d_of<double>(const DimTables &t) { return t.d.ptr; }
template <> const c64 *d_of<float>(const DimTables &t) { return t.d32.ptr; }

// Work array of the passes, per right-hand side: [component][x position (digit-reversed)][y column], the y columns
// of one x position contiguous (row length ldj >= the columns of either component). Pass A writes it transposed in
// 64-byte row segments, pass B reads and writes whole contiguous columns, pass C reads it back in row segments.
__host__ __device__ inline int64_t work_index(const ProdDims &d, int r, int comp, int kx, int j) {
    return ((static_cast<int64_t>(r) * 2 + comp) * d.Mx + kx) * d.ldj + j;
}

// Decodes a pass-A/C block into (right-hand side, component, first y column).
__device__ inline void group_of(const ProdDims &d, int group, int blk, int &r, int &comp, int &j0) {
    const int gx = (d.cols_x + group - 1) / group, gy = (d.cols_y + group - 1) / group;
    const int per = gx + gy;
    r = blk / per;
    const int rem = blk - r * per;
    comp = rem < gx ? 0 : 1;
    j0 = (comp ? rem - gx : rem) * group;
}

// ------------------------------------------------------------------------------------------------------------------
// Pass A: forward FFT along x of `group` consecutive rooftop columns of one component (zero-padded from their rows
// to Mx), written transposed into the work array.
// ------------------------------------------------------------------------------------------------------------------
template <typename TC, typename TIO>
__global__ void __launch_bounds__(256) pass_a(const cplx<TIO> *__restrict__ in, cplx<TC> *__restrict__ work,
                                              const uint8_t *__restrict__ exists, ProdDims d, FftPlan plan,
                                              const cplx<TC> *__restrict__ tw, int group,
                                              const int *__restrict__ skip, int skip_stride) {
    extern __shared__ unsigned char smem_raw[];
    cplx<TC> *buf = reinterpret_cast<cplx<TC> *>(smem_raw);
    const int M = d.Mx;
    const int stride = M + 1;
    int r, comp, j0;
    group_of(d, group, blockIdx.x, r, comp, j0);
    if (skip && skip[r * skip_stride]) return;
    const int cols = comp ? d.cols_y : d.cols_x;
    const int rows = comp ? d.rows_y : d.rows_x;
    const int ncols = min(group, cols - j0);
    const int64_t off = static_cast<int64_t>(r) * d.n + (comp ? d.nx_count : 0);
    for (int e = threadIdx.x; e < ncols * M; e += blockDim.x) {
        const int c = e / M, i = e - c * M;
        cplx<TC> v = mk<TC>(0, 0);
        if (i < rows) {
            const int64_t idx = static_cast<int64_t>(j0 + c) * rows + i;
            if (!exists || exists[(comp ? d.nx_count : 0) + idx]) v = cvt<TC>(in[off + idx]);
        }
        buf[c * stride + i] = v;
    }
    __syncthreads();
    lds_fft<TC, false>(buf, ncols, stride, plan, tw, threadIdx.x, blockDim.x);
    for (int e = threadIdx.x; e < ncols * M; e += blockDim.x) {
        const int kx = e / ncols, c = e - kx * ncols;
        work[work_index(d, r, comp, kx, j0 + c)] = buf[c * stride + kx];
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Pass C: inverse FFT along x of `group` consecutive columns read back from the work array; the first rows of each
// column are the product (the cut of FFTApplicator.m:80-84).
// ------------------------------------------------------------------------------------------------------------------
template <typename TC, typename TIO>
__global__ void __launch_bounds__(256) pass_c(const cplx<TC> *__restrict__ work, cplx<TIO> *__restrict__ out,
                                              const uint8_t *__restrict__ exists, ProdDims d, FftPlan plan,
                                              const cplx<TC> *__restrict__ tw, int group,
                                              const int *__restrict__ skip, int skip_stride) {
    extern __shared__ unsigned char smem_raw[];
    cplx<TC> *buf = reinterpret_cast<cplx<TC> *>(smem_raw);
    const int M = d.Mx;
    const int stride = M + 1;
    int r, comp, j0;
    group_of(d, group, blockIdx.x, r, comp, j0);
    if (skip && skip[r * skip_stride]) return;
    const int cols = comp ? d.cols_y : d.cols_x;
    const int rows = comp ? d.rows_y : d.rows_x;
    const int ncols = min(group, cols - j0);
    for (int e = threadIdx.x; e < ncols * M; e += blockDim.x) {
        const int kx = e / ncols, c = e - kx * ncols;
        buf[c * stride + kx] = work[work_index(d, r, comp, kx, j0 + c)];
    }
    __syncthreads();
    lds_fft<TC, true>(buf, ncols, stride, plan, tw, threadIdx.x, blockDim.x);
    const int64_t off = static_cast<int64_t>(r) * d.n + (comp ? d.nx_count : 0);
    for (int e = threadIdx.x; e < ncols * rows; e += blockDim.x) {
        const int c = e / rows, i = e - c * rows;
        const int64_t idx = static_cast<int64_t>(j0 + c) * rows + i;
        cplx<TC> v = buf[c * stride + i];
        if (exists && !exists[(comp ? d.nx_count : 0) + idx]) v = mk<TC>(0, 0);
        out[off + idx] = cvt<TIO>(v);
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Pass B: for `tile` consecutive x positions, forward FFT along y of both components (contiguous columns,
// zero-padded to My), the three-plane spectral multiply, inverse FFT along y, and the cut back to the columns.
// LDS holds [component][tile][My (+1 pad)].
// ------------------------------------------------------------------------------------------------------------------
template <typename TC>
__global__ void __launch_bounds__(256) pass_b(cplx<TC> *__restrict__ work, const cplx<TC> *__restrict__ planes,
                                              const int *__restrict__ fold_x, const int *__restrict__ fold_y,
                                              const cplx<TC> *__restrict__ dx, const cplx<TC> *__restrict__ dy,
                                              ProdDims d, FftPlan plan, const cplx<TC> *__restrict__ tw, int tile,
                                              const int *__restrict__ skip, int skip_stride) {
    const int r = blockIdx.y;
    if (skip && skip[r * skip_stride]) return;
    extern __shared__ unsigned char smem_raw[];
    cplx<TC> *buf = reinterpret_cast<cplx<TC> *>(smem_raw);
    const int My = d.My;
    const int stride = My + 1;
    const int px0 = blockIdx.x * tile;
    const int ncols = 2 * tile;
    for (int e = threadIdx.x; e < ncols * My; e += blockDim.x) {
        const int col = e / My, j = e - col * My;
        const int comp = col / tile, t = col - comp * tile;
        const int cols = comp ? d.cols_y : d.cols_x;
        cplx<TC> v = mk<TC>(0, 0);
        if (j < cols) v = work[work_index(d, r, comp, px0 + t, j)];
        buf[col * stride + j] = v;
    }
    __syncthreads();
    lds_fft<TC, false>(buf, ncols, stride, plan, tw, threadIdx.x, blockDim.x);
    const int64_t plane = static_cast<int64_t>(d.Hx) * d.Hy;
    for (int e = threadIdx.x; e < tile * My; e += blockDim.x) {
        const int t = e / My, py = e - t * My;
        const int px = px0 + t;
        const int64_t f = static_cast<int64_t>(fold_x[px]) * d.Hy + fold_y[py];
        const cplx<TC> a = planes[f], b = planes[plane + f], el = planes[2 * plane + f];
        const cplx<TC> dxv = dx[px], dyv = dy[py];
        const cplx<TC> X = buf[t * stride + py];
        const cplx<TC> Y = buf[(tile + t) * stride + py];
        const cplx<TC> phi = el * (cmulc(dxv, X) + cmulc(dyv, Y));
        buf[t * stride + py] = a * X - dxv * phi;
        buf[(tile + t) * stride + py] = b * Y - dyv * phi;
    }
    __syncthreads();
    lds_fft<TC, true>(buf, ncols, stride, plan, tw, threadIdx.x, blockDim.x);
    for (int e = threadIdx.x; e < ncols * My; e += blockDim.x) {
        const int col = e / My, j = e - col * My;
        const int comp = col / tile, t = col - comp * tile;
        const int cols = comp ? d.cols_y : d.cols_x;
        if (j < cols) work[work_index(d, r, comp, px0 + t, j)] = buf[col * stride + j];
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Spectra: the even embedding of a generator (G(p mod Mx, q mod My) = g(|p|, |q|) for |p| < rows, |q| < cols), its
// forward 2D FFT, and the fold to [0, Mx/2] x [0, My/2] (an even real-index circulant has an even spectrum).
// ------------------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) spectra_a(const c128 *__restrict__ g, int rows, int cols, int modifier,
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
        buf[i] = v;
    }
    __syncthreads();
    lds_fft<double, false>(buf, 1, Mx, plan, tw, threadIdx.x, blockDim.x);
    for (int i = threadIdx.x; i < Mx; i += blockDim.x) out[static_cast<int64_t>(q) * Mx + i] = buf[i];
}

__global__ void __launch_bounds__(256) spectra_b(const c128 *__restrict__ in, c128 *__restrict__ plane, int Mx,
                                                 int My, int Hy, FftPlan plan, const c128 *__restrict__ tw,
                                                 const int *__restrict__ fold_x, const int *__restrict__ fold_y,
                                                 const int *__restrict__ freq_x, const int *__restrict__ freq_y,
                                                 double scale) {
    extern __shared__ unsigned char smem_raw[];
    c128 *buf = reinterpret_cast<c128 *>(smem_raw);
    const int px = blockIdx.x;
    for (int q = threadIdx.x; q < My; q += blockDim.x) buf[q] = in[static_cast<int64_t>(q) * Mx + px];
    __syncthreads();
    lds_fft<double, false>(buf, 1, My, plan, tw, threadIdx.x, blockDim.x);
    const int kx = freq_x[px];
    if (2 * kx > Mx) return;
    for (int py = threadIdx.x; py < My; py += blockDim.x) {
        const int ky = freq_y[py];
        if (2 * ky <= My) plane[static_cast<int64_t>(fold_x[px]) * Hy + fold_y[py]] = scale * buf[py];
    }
}

__global__ void to_c64(const c128 *__restrict__ a, c64 *__restrict__ b, int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        b[i] = cvt<float>(a[i]);
}

// ------------------------------------------------------------------------------------------------------------------
// Host side.
// ------------------------------------------------------------------------------------------------------------------
// Default launch shape (the autotuner may replace it). Pass A/C: `group` y columns per block, 64-byte row segments
// when enough blocks remain (>= 256), fewer columns otherwise. Pass B: one x position per block unless the grid
// would be very large.
template <typename TC>
static PassShape default_shape(const ProdDims &d, int nrhs) {
    PassShape s;
    const int es = sizeof(cplx<TC>);
    int group = 64 / es;
    const int lds_cols = static_cast<int>((64 * 1024) / ((d.Mx + 1) * es));
    auto blocks_ac = [&](int g) {
        return static_cast<int64_t>(nrhs) * ((d.cols_x + g - 1) / g + (d.cols_y + g - 1) / g);
    };
    while (group > 1 && (group > lds_cols || blocks_ac(group) < 256)) group >>= 1;
    s.cols_per_block = group;
    const int thr = group * d.Mx / 4;
    s.threads_ac = thr < 64 ? 64 : thr > 256 ? 256 : (thr + 63) / 64 * 64;
    int tile = 1;
    const int lds_tile = static_cast<int>((80 * 1024) / (2 * (d.My + 1) * es));
    while (tile * 2 <= 16 && tile * 2 <= lds_tile && tile * 2 <= d.Mx &&
           static_cast<int64_t>(d.Mx / (tile * 2)) * nrhs >= 8192)
        tile *= 2;
    s.tile = tile;
    const int tb = 2 * tile * d.My / 4;
    s.threads_b = tb < 64 ? 64 : tb > 256 ? 256 : (tb + 63) / 64 * 64;
    return s;
}

static size_t lds_bytes_max() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
        v = 64 * 1024;
    return static_cast<size_t>(v);
}

template <typename TC>
static void set_lds(const void *kernel, size_t bytes) {
    if (bytes > 64 * 1024) {
        if (bytes > lds_bytes_max())
            fail(RT_EINVAL, "rooftop:product", "an FFT column of %zu bytes exceeds the LDS of this device", bytes);
        RT_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
    }
}

// Algorithmic HBM bytes of one launch of each pass: every vector element, work element and spectrum value once.
template <typename TC>
static void pass_bytes(const ProdDims &d, int nrhs, double *bytes) {
    const double sc = sizeof(cplx<TC>);
    const double vec = 16.0 * d.n * nrhs;
    const double work_bytes = sc * (static_cast<double>(d.cols_x) + d.cols_y) * d.Mx * nrhs;
    bytes[0] = vec + work_bytes;
    bytes[1] = 2 * work_bytes + 3.0 * sc * d.Hx * d.Hy;
    bytes[2] = work_bytes + vec;
}

template <typename TC>
static void run_passes(Product &op, const c128 *in, c128 *out, cplx<TC> *work, int nrhs, const PassShape &sh,
                       const cplx<TC> *planes, hipStream_t s, const int *skip = nullptr, int skip_stride = 0) {
    const ProdDims &d = op.dims;
    Context *ctx = op.problem->ctx;
    const uint8_t *ex = op.problem->exists.ptr;
    const int g = sh.cols_per_block;
    const unsigned blocks_ac = static_cast<unsigned>(nrhs * ((d.cols_x + g - 1) / g + (d.cols_y + g - 1) / g));
    const size_t lds_ac = static_cast<size_t>(g) * (d.Mx + 1) * sizeof(cplx<TC>);
    const size_t lds_b = static_cast<size_t>(2) * sh.tile * (d.My + 1) * sizeof(cplx<TC>);
    set_lds<TC>(reinterpret_cast<const void *>(&pass_a<TC, double>), lds_ac);
    set_lds<TC>(reinterpret_cast<const void *>(&pass_c<TC, double>), lds_ac);
    set_lds<TC>(reinterpret_cast<const void *>(&pass_b<TC>), lds_b);
    Context::PassRecord rec{};
    const bool timed = ctx->timing;
    if (timed) {
        for (auto &e : rec.ev) e = ctx->take_event();
        rec.precision = sizeof(TC) == sizeof(float) ? 1 : 0;
        pass_bytes<TC>(d, nrhs, rec.bytes);
        RT_HIP(hipEventRecord(rec.ev[0], s));
    }
    hipLaunchKernelGGL((pass_a<TC, double>), dim3(blocks_ac), dim3(sh.threads_ac), lds_ac, s,
                       reinterpret_cast<const cplx<double> *>(in), work, ex, d, op.tx.plan, tw_of<TC>(op.tx), g, skip, skip_stride);
    RT_LAUNCH_CHECK();
    if (timed) RT_HIP(hipEventRecord(rec.ev[1], s));
    hipLaunchKernelGGL((pass_b<TC>), dim3(d.Mx / sh.tile, nrhs), dim3(sh.threads_b), lds_b, s, work, planes,
                       op.tx.fold.ptr, op.ty.fold.ptr, d_of<TC>(op.tx), d_of<TC>(op.ty), d, op.ty.plan,
                       tw_of<TC>(op.ty), sh.tile, skip, skip_stride);
    RT_LAUNCH_CHECK();
    if (timed) RT_HIP(hipEventRecord(rec.ev[2], s));
    hipLaunchKernelGGL((pass_c<TC, double>), dim3(blocks_ac), dim3(sh.threads_ac), lds_ac, s, work,
                       reinterpret_cast<cplx<double> *>(out), ex, d, op.tx.plan, tw_of<TC>(op.tx), g, skip, skip_stride);
    RT_LAUNCH_CHECK();
    if (timed) {
        RT_HIP(hipEventRecord(rec.ev[3], s));
        ctx->pending.push_back(rec);
        if (ctx->pending.size() > 4096) ctx->drain_timing();
    }
}

void operator_apply(Product &op, const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s, const int *skip,
                    int skip_stride) {
    if (prec == RT_MIXED) prec = RT_C128;
    if (!op.has(prec)) fail(RT_EINVAL, "rooftop:product", "the product does not hold the requested precision");
    const ProdDims &d = op.dims;
    const size_t per_rhs = static_cast<size_t>(2) * d.ldj * d.Mx;
    if (prec == RT_C128) {
        if (op.work_rhs64 < nrhs) {
            op.work64.allocate(op.problem->ctx, per_rhs * nrhs, "the product's working memory");
            op.work_rhs64 = nrhs;
            op.shape64 = default_shape<double>(d, nrhs);
        }
        PassShape sh = default_shape<double>(d, nrhs);
        run_passes<double>(op, in, out, op.work64.ptr, nrhs, sh, op.planes64.ptr, s, skip, skip_stride);
    } else {
        if (op.work_rhs32 < nrhs) {
            op.work32.allocate(op.problem->ctx, per_rhs * nrhs, "the product's working memory");
            op.work_rhs32 = nrhs;
        }
        PassShape sh = default_shape<float>(d, nrhs);
        run_passes<float>(op, in, out, op.work32.ptr, nrhs, sh, op.planes32.ptr, s, skip, skip_stride);
    }
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
    ProdDims &d = op->dims;
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
    const uint64_t plane_elems = static_cast<uint64_t>(d.Hx) * d.Hy;
    const uint64_t temp = static_cast<uint64_t>(d.Mx) * d.My * sizeof(c128);
    ctx->require(temp, "The spectra of the product (transform buffer)");
    p->ensure_generators();
    op->tx.build(ctx, d.Mx, g.a, s);
    op->ty.build(ctx, d.My, g.b, s);
    op->planes64.allocate(ctx, 3 * plane_elems, "The spectra of the product");
    op->planes64.zero(s);
    {
        DeviceArray<c128> tmp;
        tmp.allocate(ctx, static_cast<size_t>(d.Mx) * d.My, "The spectra of the product (transform buffer)");
        std::vector<int> fx = output_frequencies(op->tx.plan), fy = output_frequencies(op->ty.plan);
        DeviceArray<int> dfx, dfy;
        dfx.allocate(ctx, d.Mx, "FFT frequency table");
        dfy.allocate(ctx, d.My, "FFT frequency table");
        RT_HIP(hipMemcpyAsync(dfx.ptr, fx.data(), d.Mx * sizeof(int), hipMemcpyHostToDevice, s));
        RT_HIP(hipMemcpyAsync(dfy.ptr, fy.data(), d.My * sizeof(int), hipMemcpyHostToDevice, s));
        const double scale = 1.0 / (static_cast<double>(d.Mx) * d.My);
        const c128 *gens[3] = {g.has_x ? p->mxx.ptr : nullptr, g.has_y ? p->myy.ptr : nullptr, p->e.ptr};
        const int rows[3] = {g.cx - 1, g.cx, g.cx};
        const int cols[3] = {g.cy, g.cy - 1, g.cy};
        for (int k = 0; k < 3; ++k) {
            if (!gens[k]) continue;
            size_t lds_a = static_cast<size_t>(d.Mx) * sizeof(c128);
            size_t lds_b = static_cast<size_t>(d.My) * sizeof(c128);
            set_lds<double>(reinterpret_cast<const void *>(&spectra_a), lds_a);
            set_lds<double>(reinterpret_cast<const void *>(&spectra_b), lds_b);
            hipLaunchKernelGGL(spectra_a, dim3(d.My), dim3(256), lds_a, s, gens[k], rows[k], cols[k], modifier,
                               tmp.ptr, d.Mx, d.My, op->tx.plan, op->tx.tw.ptr);
            RT_LAUNCH_CHECK();
            hipLaunchKernelGGL(spectra_b, dim3(d.Mx), dim3(256), lds_b, s, tmp.ptr, op->planes64.ptr + k * plane_elems,
                               d.Mx, d.My, d.Hy, op->ty.plan, op->ty.tw.ptr, op->tx.fold.ptr, op->ty.fold.ptr,
                               dfx.ptr, dfy.ptr, scale);
            RT_LAUNCH_CHECK();
        }
        RT_HIP(hipStreamSynchronize(s));
    }
    if (precision == RT_C64 || precision == RT_MIXED) {
        op->planes32.allocate(ctx, 3 * plane_elems, "The spectra of the product");
        hipLaunchKernelGGL(to_c64, dim3(blocks_for(3 * plane_elems, 256, 4096)), dim3(256), 0, s,
                           op->planes64.ptr, op->planes32.ptr, static_cast<int64_t>(3 * plane_elems));
        RT_LAUNCH_CHECK();
        if (precision == RT_C64) {
            RT_HIP(hipStreamSynchronize(s));
            op->planes64.reset();
        }
    }
    RT_HIP(hipStreamSynchronize(s));
    return op;
}

// Times `reps` applications with the context's kernel timing (the first application warms up and is not counted).
template <typename TC>
static void profile_passes(Product &op, int nrhs, int reps, cplx<TC> *work, const cplx<TC> *planes, double *ms,
                           double *bytes, hipStream_t s) {
    const ProdDims &d = op.dims;
    Context *ctx = op.problem->ctx;
    DeviceArray<c128> in, out;
    in.allocate(ctx, static_cast<size_t>(d.n) * nrhs, "profile input");
    out.allocate(ctx, static_cast<size_t>(d.n) * nrhs, "profile output");
    in.zero(s);
    const PassShape sh = default_shape<TC>(d, nrhs);
    run_passes<TC>(op, in.ptr, out.ptr, work, nrhs, sh, planes, s);
    RT_HIP(hipStreamSynchronize(s));
    const bool was = ctx->timing;
    ctx->drain_timing();
    const int p = sizeof(TC) == sizeof(float) ? 1 : 0;
    const int64_t n0 = ctx->launches[p];
    double ms0[3], by0[3];
    for (int k = 0; k < 3; ++k) { ms0[k] = ctx->pass_ms[p][k]; by0[k] = ctx->pass_bytes[p][k]; }
    ctx->timing = true;
    for (int it = 0; it < reps; ++it) run_passes<TC>(op, in.ptr, out.ptr, work, nrhs, sh, planes, s);
    ctx->drain_timing();
    ctx->timing = was;
    const double cnt = static_cast<double>(ctx->launches[p] - n0);
    for (int k = 0; k < 3; ++k) {
        ms[k] = (ctx->pass_ms[p][k] - ms0[k]) / cnt;
        bytes[k] = (ctx->pass_bytes[p][k] - by0[k]) / cnt;
        ctx->pass_ms[p][k] = ms0[k];
        ctx->pass_bytes[p][k] = by0[k];
    }
    ctx->launches[p] = n0;
}

void operator_profile(Product &op, int nrhs, int reps, int prec, double *ms, double *bytes, hipStream_t s) {
    if (op.type != kOperator) fail(RT_EINVAL, "rooftop:product", "only an operator has matvec passes to profile");
    if (prec == RT_MIXED) prec = RT_C128;
    if (!op.has(prec)) fail(RT_EINVAL, "rooftop:product", "the product does not hold the requested precision");
    const size_t per_rhs = static_cast<size_t>(2) * op.dims.ldj * op.dims.Mx;
    if (prec == RT_C128) {
        if (op.work_rhs64 < nrhs) {
            op.work64.allocate(op.problem->ctx, per_rhs * nrhs, "the product's working memory");
            op.work_rhs64 = nrhs;
        }
        profile_passes<double>(op, nrhs, reps, op.work64.ptr, op.planes64.ptr, ms, bytes, s);
    } else {
        if (op.work_rhs32 < nrhs) {
            op.work32.allocate(op.problem->ctx, per_rhs * nrhs, "the product's working memory");
            op.work_rhs32 = nrhs;
        }
        profile_passes<float>(op, nrhs, reps, op.work32.ptr, op.planes32.ptr, ms, bytes, s);
    }
}

uint64_t Product::stored_bytes() const {
    return planes64.bytes() + planes32.bytes() + diag64.bytes() + diag32.bytes() + coef64.bytes() + coef32.bytes() +
           cls.bytes();
}

uint64_t Product::working_bytes() const {
    if (type != kOperator) return 0;
    const uint64_t per = static_cast<uint64_t>(2) * dims.ldj * dims.Mx;
    return per * (has(RT_C128) ? sizeof(c128) : sizeof(c64));
}

void Product::apply(const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s, const int *skip, int skip_stride) {
    if (type == kOperator)
        operator_apply(*this, in, out, nrhs, prec, s, skip, skip_stride);
    else
        preconditioner_apply(*this, in, out, nrhs, prec, s, skip, skip_stride);
}

}  // namespace rt

// The C ABI of librooftop (include/rooftop.h): argument checks with the reference's error identifiers, handle
// ownership, host<->device staging. Every entry point catches the library's errors and returns their status.
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "krylov.hpp"
#include "rooftop.hpp"

namespace rt {
const char *last_error_id();
const char *last_error_message();

GridInfo GridInfo::make(int cx, int cy, double a, double b) {
    GridInfo g;
    g.cx = cx;
    g.cy = cy;
    g.a = a;
    g.b = b;
    g.rows_x = cx - 1;
    g.cols_x = cy;
    g.rows_y = cx;
    g.cols_y = cy - 1;
    g.nx = static_cast<int64_t>(g.rows_x) * g.cols_x;
    g.ny = static_cast<int64_t>(g.rows_y) * g.cols_y;
    g.n = g.nx + g.ny;
    g.has_x = g.nx > 0;
    g.has_y = g.ny > 0;
    return g;
}
std::shared_ptr<Problem> make_problem(Context *ctx, int cx, int cy, double a, double b, double k, int near_order,
                                      int far_order, const uint8_t *mask) {
    if (cx < 1 || cy < 1) fail(RT_EINVAL, "MATLAB:validators:mustBePositive", "The face counts are positive.");
    if (cx == 1 && cy == 1)
        fail(RT_EINVAL, "rooftopMoM:mesh:noRooftops", "A single face has no interior edge, so it carries no rooftop.");
    if (!(a > 0 && b > 0 && std::isfinite(a) && std::isfinite(b)))
        fail(RT_EINVAL, "MATLAB:validators:mustBePositive", "The face sizes are positive and finite.");
    if (!(k > 0 && std::isfinite(k)))
        fail(RT_EINVAL, "rooftopMoM:problem:invalidWavenumber", "The wavenumber is positive and finite.");
    if (near_order < 1 || far_order < 1 || far_order > 16 || near_order > 64)
        fail(RT_EINVAL, "rooftopMoM:problem:invalidQuadratureOrder",
             "The quadrature orders are 1..64 (near) and 1..16 (far).");
    auto p = std::make_shared<Problem>();
    p->ctx = ctx;
    p->g = GridInfo::make(cx, cy, a, b);
    p->k = k;
    p->near_order = near_order;
    p->far_order = far_order;
    // Budget of the generators (estimateGenerators.m:11) checked up front, as Problem.m:210-219 does.
    ctx->require(16ull * (p->g.n + static_cast<uint64_t>(cx) * cy), "The assembly of the generators");
    if (mask) {
        p->face_mask.assign(mask, mask + static_cast<size_t>(cx) * cy);
        std::vector<uint8_t> ex(static_cast<size_t>(p->g.n));
        auto face = [&](int x, int y) { return mask[x + static_cast<size_t>(cx) * y] != 0; };
        for (int j = 0; j < p->g.cols_x; ++j)
            for (int i = 0; i < p->g.rows_x; ++i)
                ex[i + static_cast<size_t>(p->g.rows_x) * j] = face(i, j) && face(i + 1, j);
        for (int j = 0; j < p->g.cols_y; ++j)
            for (int i = 0; i < p->g.rows_y; ++i)
                ex[p->g.nx + i + static_cast<size_t>(p->g.rows_y) * j] = face(i, j) && face(i, j + 1);
        p->exists.allocate(ctx, ex.size(), "The rooftop mask");
        RT_HIP(hipMemcpyAsync(p->exists.ptr, ex.data(), ex.size(), hipMemcpyHostToDevice, ctx->stream));
        RT_HIP(hipStreamSynchronize(ctx->stream));
    }
    return p;
}
}  // namespace rt

using namespace rt;

struct rt_ctx {
    Context c;
};
// A product keeps its problem alive (shared ownership), so handles may be destroyed in any order.
struct rt_problem {
    std::shared_ptr<Problem> p;
};
struct rt_product {
    std::shared_ptr<Problem> owner;   // declared first: destroyed after the product
    std::unique_ptr<Product> p;
};

namespace {

template <typename F>
int guarded(F &&f) {
    try {
        f();
        set_last_error(RT_OK, "", "");
        return RT_OK;
    } catch (const Error &e) {
        set_last_error(e.status, e.id, e.what());
        return e.status;
    } catch (const std::bad_alloc &) {
        set_last_error(RT_ENOMEM, "rooftopMoM:resources:insufficientMemory", "host allocation failed");
        return RT_ENOMEM;
    } catch (const std::exception &e) {
        set_last_error(RT_EINVAL, "rooftop:internal", e.what());
        return RT_EINVAL;
    }
}

void need(bool ok, int status, const char *id, const char *msg) {
    if (!ok) fail(status, id, "%s", msg);
}

SolveOptions options_of(const rt_solve_opts *o) {
    SolveOptions s;
    if (!o) return s;
    s.method = o->method;
    s.tolerance = o->tolerance;
    s.max_iterations = o->max_iterations;
    s.restart = o->restart;
    s.precision = o->precision;
    s.refine = o->refine;
    s.inner_tolerance = o->inner_tolerance;
    s.max_outer = o->max_outer > 0 ? o->max_outer : 20;
    if (s.precision == RT_C64_FP64_BASIS) {
        s.precision = RT_C64;
        s.fp64_basis = true;
    }
    if (s.method != RT_METHOD_CG && s.method != RT_METHOD_GMRES && s.method != RT_METHOD_BLOCK_CG)
        fail(RT_EINVAL, "MATLAB:validators:mustBeMember", "The method is 'cg', 'gmres' or 'blockcg'.");
    if (!(s.tolerance > 0 && s.tolerance < 1))
        fail(RT_EINVAL, "MATLAB:validators:mustBeLessThan", "The tolerance is a real number in (0, 1).");
    if (s.max_iterations <= 0)
        fail(RT_EINVAL, "MATLAB:validators:mustBePositive", "The maximum iteration count is a positive integer.");
    if (s.precision != RT_C128 && s.precision != RT_C64)
        fail(RT_EINVAL, "rooftop:solver:invalidPrecision", "The inner precision is RT_C128 or RT_C64.");
    if (s.precision == RT_C64 && !s.refine)
        fail(RT_EINVAL, "rooftop:solver:invalidPrecision",
             "A complex64 inner loop needs the FP64 refinement loop (refine = 1).");
    return s;
}

void fill_output(const SolveResult &r, rt_output *out) {
    if (!out) return;
    out->converged = r.converged;
    out->iteration_count = r.iterations;
    out->residual_norm = r.residual;
    out->time_s = r.time;
    out->outer_count = r.outer;
    out->residual_norms_count = static_cast<int32_t>(r.norms.size());
    if (out->residual_norms && out->residual_norms_capacity > 0) {
        const size_t k = std::min<size_t>(r.norms.size(), static_cast<size_t>(out->residual_norms_capacity));
        std::memcpy(out->residual_norms, r.norms.data(), k * sizeof(double));
    }
}

// Solves nrhs systems in lockstep on device buffers (one batched Krylov loop); one result per system.
std::vector<SolveResult> solve_device(rt_ctx *ctx, rt_product *op, rt_product *pc, const rt_solve_opts *opts,
                                      const c128 *rhs, int64_t n, int64_t nrhs, const c128 *x0, c128 *x) {
    need(ctx && op && op->p, RT_EINVAL, "rooftopMoM:solver:invalidOperator",
         "An operator is a product created by rt_product_create.");

    need(n == op->p->dims.n, RT_ESIZE, "rooftopMoM:solver:sizeMismatch",
         "The operator does not match the unknowns of the right-hand side.");
    need(!pc || (pc->p && pc->p->dims.n == n), RT_ESIZE, "rooftopMoM:solver:sizeMismatch",
         "The preconditioner does not match the unknowns of the right-hand side.");
    need(nrhs >= 1 && nrhs <= kMaxBatch, RT_EINVAL, "rooftopMoM:solver:invalidRightHandSide",
         "Between 1 and 1024 right-hand sides are solved at once.");
    need(op->p->problem->ctx == &ctx->c && (!pc || pc->p->problem->ctx == &ctx->c), RT_EINVAL,
         "rooftop:solver:contextMismatch", "The operator and the preconditioner must belong to the solving context.");
    SolveOptions o = options_of(opts);
    need(op->p->has(o.precision), RT_EINVAL, "rooftop:solver:invalidPrecision",
         "The operator does not hold the inner precision (create it with RT_MIXED).");
    need(!pc || pc->p->has(o.precision), RT_EINVAL, "rooftop:solver:invalidPrecision",
         "The preconditioner does not hold the inner precision (create it with RT_MIXED).");
    ctx->c.activate();
    return solve_batch(&ctx->c, *op->p, pc ? pc->p.get() : nullptr, o, rhs, x0, x, static_cast<int>(nrhs));
}

void fill_outputs(const std::vector<SolveResult> &r, rt_output *out) {
    if (!out) return;
    for (size_t k = 0; k < r.size(); ++k) fill_output(r[k], out + k);
}

void check_finite(const c128 *v, int64_t count, const char *id = "rooftopMoM:solver:invalidRightHandSide",
                  const char *msg = "The right-hand side is a numeric vector of finite values.") {
    for (int64_t i = 0; i < count; ++i)
        if (!std::isfinite(v[i].x) || !std::isfinite(v[i].y)) fail(RT_EINVAL, id, "%s", msg);
}

}  // namespace

extern "C" {

int rt_ctx_create(int32_t device, uint64_t budget, rt_ctx **out) {
    return guarded([&] {
        need(out != nullptr, RT_EINVAL, "rooftop:ctx", "out is NULL");
        auto c = std::make_unique<rt_ctx>();
        c->c.open(device, budget);
        noalloc_setup(device);   // the device address of the allocation-free working memory
        *out = c.release();
    });
}

void rt_ctx_destroy(rt_ctx *ctx) {
    if (!ctx) return;
    ctx->c.close();
    delete ctx;
}

int rt_ctx_memory(rt_ctx *ctx, uint64_t *held, uint64_t *ceiling) {
    return guarded([&] {
        need(ctx, RT_EINVAL, "rooftop:ctx", "ctx is NULL");
        if (held) *held = ctx->c.held;
        if (ceiling) *ceiling = ctx->c.ceiling;
    });
}

int rt_ctx_memory_peak(rt_ctx *ctx, int32_t reset, uint64_t *peak) {
    return guarded([&] {
        need(ctx, RT_EINVAL, "rooftop:ctx", "ctx is NULL");
        std::lock_guard<std::mutex> g(ctx->c.lock);
        if (peak) *peak = ctx->c.peak;
        if (reset) ctx->c.peak = ctx->c.held;
    });
}

int rt_ctx_synchronize(rt_ctx *ctx) {
    return guarded([&] {
        need(ctx, RT_EINVAL, "rooftop:ctx", "ctx is NULL");
        ctx->c.activate();
        RT_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int rt_ctx_stream(rt_ctx *ctx, void **stream) {
    return guarded([&] {
        need(ctx && stream, RT_EINVAL, "rooftop:ctx", "ctx or stream is NULL");
        *stream = reinterpret_cast<void *>(ctx->c.stream);
    });
}

int rt_ctx_set_timing(rt_ctx *ctx, int32_t on) {
    return guarded([&] {
        need(ctx != nullptr, RT_EINVAL, "rooftop:ctx", "ctx is NULL");
        ctx->c.timing = on != 0;
    });
}

int rt_ctx_kernel_timing(rt_ctx *ctx, int64_t *launches, double *ms, double *bytes) {
    return guarded([&] {
        need(ctx && launches && ms && bytes, RT_EINVAL, "rooftop:ctx", "NULL argument");
        ctx->c.activate();
        RT_HIP(hipStreamSynchronize(ctx->c.stream));
        ctx->c.drain_timing();
        for (int k = 0; k < RT_TIMED_COUNT; ++k) {
            const double n = ctx->c.timed_launches[k] ? static_cast<double>(ctx->c.timed_launches[k]) : 1.0;
            launches[k] = ctx->c.timed_launches[k];
            ms[k] = ctx->c.timed_ms[k] / n;
            bytes[k] = ctx->c.timed_bytes[k] / n;
            ctx->c.timed_launches[k] = 0;
            ctx->c.timed_ms[k] = 0;
            ctx->c.timed_bytes[k] = 0;
        }
    });
}

int rt_bandwidth_probe(rt_ctx *ctx, uint64_t bytes, int32_t reps, double *ms) {
    return guarded([&] {
        need(ctx && ms, RT_EINVAL, "rooftop:ctx", "NULL argument");
        ctx->c.activate();
        *ms = bandwidth_probe(&ctx->c, bytes, reps);
    });
}

int rt_problem_create(rt_ctx *ctx, int32_t cx, int32_t cy, double a, double b, double k, int32_t near_order,
                      int32_t far_order, const uint8_t *mask, rt_problem **out) {
    return guarded([&] {
        need(ctx && out, RT_EINVAL, "rooftop:problem", "ctx or out is NULL");
        ctx->c.activate();
        auto h = std::make_unique<rt_problem>();
        h->p = make_problem(&ctx->c, cx, cy, a, b, k, near_order, far_order, mask);
        *out = h.release();
    });
}

void rt_problem_destroy(rt_problem *p) {
    if (!p) return;
    if (p->p && p->p->ctx) (void)hipStreamSynchronize(p->p->ctx->stream);
    delete p;
}

int rt_problem_count(rt_problem *p, int64_t *count) {
    return guarded([&] {
        need(p && count, RT_EINVAL, "rooftop:problem", "NULL argument");
        *count = p->p->g.n;
    });
}

int rt_plane_waves_device(rt_problem *p, int64_t count, const double *directions, const double *polarizations,
                          const double *intensities, void *b) {
    return guarded([&] {
        need(p && directions && polarizations && b && count >= 1, RT_EINVAL, "rooftop:excitation",
             "invalid plane-wave arguments");
        Problem &q = *p->p;
        q.ctx->activate();
        plane_waves_device(q, count, directions, polarizations, intensities, static_cast<c128 *>(b), q.ctx->stream);
    });
}

int rt_radar_cross_sections_device(rt_problem *p, int64_t count, const void *b, const void *x,
                                   const double *intensities, double *sigma) {
    return guarded([&] {
        need(p && b && x && sigma && count >= 1, RT_EINVAL, "rooftop:excitation", "invalid observable arguments");
        Problem &q = *p->p;
        q.ctx->activate();
        radar_cross_sections_device(q, count, static_cast<const c128 *>(b), static_cast<const c128 *>(x), intensities,
                                    sigma, q.ctx->stream);
    });
}

int rt_input_impedances_device(rt_problem *p, int64_t count, const int64_t *port, int64_t port_len,
                               const double *voltages, const void *x, double *zin) {
    return guarded([&] {
        need(p && port && x && zin && count >= 1, RT_EINVAL, "rooftop:excitation", "invalid observable arguments");
        Problem &q = *p->p;
        q.ctx->activate();
        input_impedances_device(q, count, port, port_len, voltages, static_cast<const c128 *>(x), zin,
                                q.ctx->stream);
    });
}

int rt_problem_generators(rt_problem *p, void *mxx, void *myy, void *e) {
    return guarded([&] {
        need(p != nullptr, RT_EINVAL, "rooftop:problem", "NULL problem");
        Problem &q = *p->p;
        q.ctx->activate();
        q.ensure_generators();
        if (mxx && q.g.has_x) RT_HIP(hipMemcpy(mxx, q.mxx.ptr, q.mxx.bytes(), hipMemcpyDeviceToHost));
        if (myy && q.g.has_y) RT_HIP(hipMemcpy(myy, q.myy.ptr, q.myy.bytes(), hipMemcpyDeviceToHost));
        if (e) RT_HIP(hipMemcpy(e, q.e.ptr, q.e.bytes(), hipMemcpyDeviceToHost));
    });
}

int rt_problem_set_generators(rt_problem *p, const void *mxx, const void *myy, const void *e) {
    return guarded([&] {
        need(p != nullptr && e != nullptr, RT_EINVAL, "rooftop:problem", "NULL problem or electric generator");
        Problem &q = *p->p;
        q.ctx->activate();
        need(!q.g.has_x || mxx, RT_EINVAL, "rooftop:problem", "the grid has x-directed rooftops: magneticXX needed");
        need(!q.g.has_y || myy, RT_EINVAL, "rooftop:problem", "the grid has y-directed rooftops: magneticYY needed");
        const int cx = q.g.cx, cy = q.g.cy;
        if (q.g.has_x) {
            q.mxx.allocate(q.ctx, static_cast<size_t>(cx - 1) * cy, "The generators");
            RT_HIP(hipMemcpy(q.mxx.ptr, mxx, q.mxx.bytes(), hipMemcpyHostToDevice));
        }
        if (q.g.has_y) {
            q.myy.allocate(q.ctx, static_cast<size_t>(cx) * (cy - 1), "The generators");
            RT_HIP(hipMemcpy(q.myy.ptr, myy, q.myy.bytes(), hipMemcpyHostToDevice));
        }
        q.e.allocate(q.ctx, static_cast<size_t>(cx) * cy, "The generators");
        RT_HIP(hipMemcpy(q.e.ptr, e, q.e.bytes(), hipMemcpyHostToDevice));
        q.assembled = true;
    });
}

int rt_product_create(rt_problem *p, int32_t kind, int32_t precision, int32_t modifier, rt_product **out) {
    return guarded([&] {
        need(p && out, RT_EINVAL, "rooftop:product", "NULL argument");
        need(kind == RT_PRODUCT_FFT || kind == RT_PRODUCT_SPLIT_FFT || kind == RT_PRODUCT_LEAN_SPLIT_FFT, RT_EINVAL,
             "MATLAB:validators:mustBeMember", "The product is 'fft', 'splitFFT' or 'leanSplitFFT'.");
        need(precision == RT_C128 || precision == RT_C64 || precision == RT_MIXED, RT_EINVAL,
             "rooftop:product:invalidPrecision", "The precision is RT_C128, RT_C64 or RT_MIXED.");
        need(modifier >= RT_MODIFIER_NONE && modifier <= RT_MODIFIER_IMAG, RT_EINVAL, "rooftop:product",
             "The modifier is none, real or imag.");
        p->p->ctx->activate();
        auto h = std::make_unique<rt_product>();
        h->owner = p->p;
        h->p = create_operator(p->p.get(), precision, modifier);
        *out = h.release();
    });
}

int rt_precond_create_circulant(rt_problem *p, int32_t kernel, int32_t builder, int32_t bandwidth_x,
                                int32_t bandwidth_y, int32_t modifier, int32_t precision, rt_product **out) {
    return guarded([&] {
        need(p && out, RT_EINVAL, "rooftop:preconditioner", "NULL argument");
        need(kernel >= RT_CIRCULANT_STRANG && kernel <= RT_CIRCULANT_SUPEROPTIMAL, RT_EINVAL,
             "MATLAB:validators:mustBeMember", "The kernel is 'strang', 'tChan', 'rChan', 'huckle' or 'superoptimal'.");
        need(builder >= RT_BUILDER_PARTWISE && builder <= RT_BUILDER_HYBRID, RT_EINVAL,
             "MATLAB:validators:mustBeMember", "The builder is 'partwise', 'blockwise' or 'hybrid'.");
        need(modifier >= RT_MODIFIER_NONE && modifier <= RT_MODIFIER_IMAG, RT_EINVAL, "rooftop:product",
             "The modifier is none, real or imag.");
        need(precision == RT_C128 || precision == RT_C64 || precision == RT_MIXED, RT_EINVAL,
             "rooftop:product:invalidPrecision", "The precision is RT_C128, RT_C64 or RT_MIXED.");
        p->p->ctx->activate();
        const GridInfo &g = p->p->g;
        auto h = std::make_unique<rt_product>();
        h->owner = p->p;
        h->p = create_circulant(p->p.get(), kernel, builder, bandwidth_x > 0 ? bandwidth_x : g.cx,
                                bandwidth_y > 0 ? bandwidth_y : g.cy, modifier, precision);
        *out = h.release();
    });
}

int rt_precond_create(rt_problem *p, int32_t kind, int32_t precision, rt_product **out) {
    return guarded([&] {
        need(p && out, RT_EINVAL, "rooftop:preconditioner", "NULL argument");
        need(kind == RT_PRECOND_JACOBI || kind == RT_PRECOND_NEARFIELD, RT_EINVAL, "MATLAB:validators:mustBeMember",
             "The preconditioner is 'jacobi' or 'nearField'.");
        need(precision == RT_C128 || precision == RT_C64 || precision == RT_MIXED, RT_EINVAL,
             "rooftop:product:invalidPrecision", "The precision is RT_C128, RT_C64 or RT_MIXED.");
        p->p->ctx->activate();
        auto h = std::make_unique<rt_product>();
        h->owner = p->p;
        h->p = create_preconditioner(p->p.get(), kind, precision);
        *out = h.release();
    });
}

int rt_product_create_dense(rt_ctx *ctx, int64_t n, const void *matrix, int32_t precision, rt_product **out) {
    return guarded([&] {
        need(ctx && matrix && out, RT_EINVAL, "rooftopMoM:solver:invalidOperator", "NULL argument");
        need(n >= 1 && n <= (int64_t(1) << 20), RT_ESIZE, "rooftopMoM:solver:sizeMismatch",
             "A dense operator is a square matrix of 1 .. 2^20 rows.");
        need(precision == RT_C128 || precision == RT_C64 || precision == RT_MIXED, RT_EINVAL,
             "rooftop:product:invalidPrecision", "The precision is RT_C128, RT_C64 or RT_MIXED.");
        check_finite(static_cast<const c128 *>(matrix), n * n, "rooftopMoM:solver:invalidOperator",
                     "The operator matrix holds finite values.");
        ctx->c.activate();
        auto h = std::make_unique<rt_product>();
        h->p = create_dense(&ctx->c, n, static_cast<const c128 *>(matrix), precision, h->owner);
        *out = h.release();
    });
}

int rt_product_create_callback(rt_ctx *ctx, int64_t n, rt_apply_fn apply, void *user, rt_product **out) {
    return guarded([&] {
        need(ctx && apply && out, RT_EINVAL, "rooftopMoM:solver:invalidOperator", "NULL argument");
        need(n >= 1, RT_ESIZE, "rooftopMoM:solver:sizeMismatch", "An operator has at least one unknown.");
        ctx->c.activate();
        auto h = std::make_unique<rt_product>();
        h->p = create_callback(&ctx->c, n, apply, user, h->owner);
        *out = h.release();
    });
}

void rt_product_destroy(rt_product *p) {
    if (!p) return;
    if (p->p && p->p->problem) (void)hipStreamSynchronize(p->p->problem->ctx->stream);
    delete p;
}

int rt_product_apply_device(rt_product *p, const void *in, void *out, int64_t n, int64_t nrhs, int32_t precision) {
    return guarded([&] {
        need(p && in && out, RT_EINVAL, "rooftop:product", "NULL argument");
        need(n == p->p->dims.n && nrhs >= 1 && nrhs <= 65535, RT_ESIZE, "rooftopMoM:product:sizeMismatch",
             "The input must be a numeric Nx1 vector (N rooftops) per right-hand side.");
        Context *c = p->p->problem->ctx;
        c->activate();
        p->p->apply(static_cast<const c128 *>(in), static_cast<c128 *>(out), static_cast<int>(nrhs), precision,
                    c->stream);
    });
}

int rt_product_apply_device_noalloc(rt_product *p, const void *in, void *out, int64_t n, int64_t nrhs,
                                    int32_t precision) {
    return guarded([&] {
        need(p && in && out, RT_EINVAL, "rooftop:product", "NULL argument");
        need(n == p->p->dims.n && nrhs >= 1 && nrhs <= 65535, RT_ESIZE, "rooftopMoM:product:sizeMismatch",
             "The input must be a numeric Nx1 vector (N rooftops) per right-hand side.");
        Context *c = p->p->problem->ctx;
        c->activate();
        product_apply_noalloc(*p->p, static_cast<const c128 *>(in), static_cast<c128 *>(out), static_cast<int>(nrhs),
                              precision, c->stream);
    });
}

int rt_product_apply(rt_product *p, const void *in, void *out, int64_t n, int64_t nrhs, int32_t precision) {
    return guarded([&] {
        need(p && in && out, RT_EINVAL, "rooftop:product", "NULL argument");
        need(n == p->p->dims.n && nrhs >= 1 && nrhs <= 65535, RT_ESIZE, "rooftopMoM:product:sizeMismatch",
             "The input must be a numeric Nx1 vector (N rooftops) per right-hand side.");
        Context *c = p->p->problem->ctx;
        c->activate();
        DeviceArray<c128> din, dout;
        din.allocate(c, static_cast<size_t>(n * nrhs), "the product's input");
        dout.allocate(c, static_cast<size_t>(n * nrhs), "the product's output");
        RT_HIP(hipMemcpyAsync(din.ptr, in, din.bytes(), hipMemcpyHostToDevice, c->stream));
        p->p->apply(din.ptr, dout.ptr, static_cast<int>(nrhs), precision, c->stream);
        RT_HIP(hipMemcpyAsync(out, dout.ptr, dout.bytes(), hipMemcpyDeviceToHost, c->stream));
        RT_HIP(hipStreamSynchronize(c->stream));
    });
}

int rt_product_memory(rt_product *p, uint64_t *stored, uint64_t *working) {
    return guarded([&] {
        need(p != nullptr, RT_EINVAL, "rooftop:product", "NULL product");
        if (stored) *stored = p->p->stored_bytes();
        if (working) *working = p->p->working_bytes();
    });
}

int rt_product_spectra(rt_product *p, int32_t *dims, void *planes) {
    return guarded([&] {
        need(p && p->p, RT_EINVAL, "rooftop:product", "NULL product");
        const Product &q = *p->p;
        need(q.type == kOperator, RT_EINVAL, "rooftop:product",
             "Only the Green's matrix products hold the three spectral planes.");
        if (dims) {
            dims[0] = q.dims.Mx;
            dims[1] = q.dims.My;
            dims[2] = q.dims.Hx;
            dims[3] = q.dims.Hy;
        }
        if (planes) {
            Context *c = q.problem->ctx;
            c->activate();
            RT_HIP(hipStreamSynchronize(c->stream));
            RT_HIP(hipMemcpy(planes, q.planes64.ptr, q.planes64.bytes(), hipMemcpyDeviceToHost));
        }
    });
}

int rt_product_estimate_memory(int32_t count_x, int32_t count_y, int32_t precision, uint64_t *stored,
                               uint64_t *working) {
    return guarded([&] {
        need(count_x >= 1 && count_y >= 1 && (count_x > 1 || count_y > 1), RT_EINVAL, "rooftopMoM:mesh:noRooftops",
             "The grid carries no rooftop.");
        need(precision == RT_C128 || precision == RT_C64 || precision == RT_MIXED, RT_EINVAL,
             "rooftop:product:invalidPrecision", "The precision is RT_C128, RT_C64 or RT_MIXED.");
        const ProdDims d = operator_dims(GridInfo::make(count_x, count_y, 1.0, 1.0));
        const bool c128 = precision != RT_C64, c64 = precision != RT_C128;
        const uint64_t hx = d.Hx, my = d.My, plane = static_cast<uint64_t>(d.Hx) * d.Hy;
        if (stored) *stored = 3 * plane * 16 + 3 * hx * my * ((c128 ? 16 : 0) + (c64 ? 8 : 0));
        if (working) *working = static_cast<uint64_t>(2) * d.ldj * d.Mx * (c128 ? 16 : 8);
    });
}

int rt_product_profile(rt_product *p, int64_t nrhs, int32_t reps, int32_t precision, double *ms, double *bytes) {
    return guarded([&] {
        need(p && ms && bytes && reps >= 1 && nrhs >= 1 && nrhs <= 65535, RT_EINVAL, "rooftop:product",
             "invalid profile arguments");
        need(p->p->type == kOperator || p->p->type == kCirculant, RT_EINVAL, "rooftop:product",
             "Only the FFT products have launch shapes to profile or tune.");
        Context *c = p->p->problem->ctx;
        c->activate();
        operator_profile(*p->p, static_cast<int>(nrhs), reps, precision, ms, bytes, c->stream);
    });
}

int rt_product_tune(rt_product *p, int64_t nrhs, int32_t precision, int32_t *shape, double *ms) {
    return guarded([&] {
        need(p && p->p && nrhs >= 1 && nrhs <= 65535, RT_EINVAL, "rooftop:product", "invalid tuning arguments");
        need(p->p->type == kOperator || p->p->type == kCirculant, RT_EINVAL, "rooftop:product",
             "Only the FFT products have launch shapes to profile or tune.");
        Context *c = p->p->problem->ctx;
        c->activate();
        int sh[5] = {0, 0, 0, 0, 0};
        double t = 0;
        operator_tune(*p->p, static_cast<int>(nrhs), precision, sh, &t, c->stream);
        if (shape)
            for (int k = 0; k < 5; ++k) shape[k] = sh[k];
        if (ms) *ms = t;
    });
}

int rt_product_set_shape(rt_product *p, int64_t nrhs, int32_t precision, const int32_t *shape) {
    return guarded([&] {
        need(p && p->p && shape && nrhs >= 1 && nrhs <= 65535, RT_EINVAL, "rooftop:product", "invalid shape arguments");
        need(p->p->type == kOperator || p->p->type == kCirculant, RT_EINVAL, "rooftop:product",
             "Only the FFT products have launch shapes to profile or tune.");
        Context *c = p->p->problem->ctx;
        c->activate();
        const int sh[5] = {shape[0], shape[1], shape[2], shape[3], shape[4]};
        operator_set_shape(*p->p, static_cast<int>(nrhs), precision, sh);
    });
}

int rt_solve_device(rt_ctx *ctx, rt_product *op, rt_product *pc, const rt_solve_opts *opts, const void *rhs,
                    int64_t n, int64_t nrhs, const void *x0, void *x, rt_output *out) {
    return guarded([&] {
        need(rhs && x, RT_EINVAL, "rooftopMoM:solver:invalidRightHandSide", "NULL right-hand side or solution");
        const auto r = solve_device(ctx, op, pc, opts, static_cast<const c128 *>(rhs), n, nrhs,
                                    static_cast<const c128 *>(x0), static_cast<c128 *>(x));
        fill_outputs(r, out);
    });
}

int rt_solve(rt_ctx *ctx, rt_product *op, rt_product *pc, const rt_solve_opts *opts, const void *rhs, int64_t n,
             int64_t nrhs, const void *x0, void *x, rt_output *out) {
    return guarded([&] {
        need(ctx && rhs && x, RT_EINVAL, "rooftopMoM:solver:invalidRightHandSide", "NULL argument");
        need(n >= 1 && nrhs >= 1, RT_ESIZE, "rooftopMoM:solver:sizeMismatch", "empty right-hand side");
        check_finite(static_cast<const c128 *>(rhs), n * nrhs);
        Context *c = &ctx->c;
        c->activate();
        const size_t count = static_cast<size_t>(n * nrhs);
        DeviceArray<c128> db, dx, dx0;
        db.allocate(c, count, "the right-hand sides");
        dx.allocate(c, count, "the solutions");
        RT_HIP(hipMemcpyAsync(db.ptr, rhs, db.bytes(), hipMemcpyHostToDevice, c->stream));
        if (x0) {
            dx0.allocate(c, count, "the initial guesses");
            RT_HIP(hipMemcpyAsync(dx0.ptr, x0, dx0.bytes(), hipMemcpyHostToDevice, c->stream));
        }
        const auto r = solve_device(ctx, op, pc, opts, db.ptr, n, nrhs, x0 ? dx0.ptr : nullptr, dx.ptr);
        RT_HIP(hipMemcpyAsync(x, dx.ptr, dx.bytes(), hipMemcpyDeviceToHost, c->stream));
        RT_HIP(hipStreamSynchronize(c->stream));
        fill_outputs(r, out);
    });
}

int rt_device_alloc(rt_ctx *ctx, uint64_t bytes, void **ptr) {
    return guarded([&] {
        need(ctx && ptr, RT_EINVAL, "rooftop:ctx", "NULL argument");
        ctx->c.activate();
        *ptr = ctx->c.alloc(bytes, "A caller buffer");
    });
}

int rt_device_free(rt_ctx *ctx, void *ptr) {
    return guarded([&] {
        need(ctx != nullptr, RT_EINVAL, "rooftop:ctx", "NULL ctx");
        ctx->c.activate();
        RT_HIP(hipStreamSynchronize(ctx->c.stream));
        ctx->c.release(ptr);
    });
}

int rt_copy_to_device(rt_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    return guarded([&] {
        need(ctx && dst && src, RT_EINVAL, "rooftop:ctx", "NULL argument");
        ctx->c.activate();
        RT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
        RT_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int rt_copy_to_host(rt_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    return guarded([&] {
        need(ctx && dst && src, RT_EINVAL, "rooftop:ctx", "NULL argument");
        ctx->c.activate();
        RT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
        RT_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

const char *rt_last_error_id(void) { return rt::last_error_id(); }
const char *rt_last_error_message(void) { return rt::last_error_message(); }
const char *rt_version(void) { return "rooftop-hip 0.1.0 (gfx950)"; }

}  // extern "C"

// A batch of problems on one grid as one product: the served requests of one embedding size (the same FFT periods
// Mx x My) padded to a common face grid with their masks, each with its own wavenumber, face sizes, generators,
// spectra and near field. The matvec passes serve all of them in one launch per pass: right-hand side r reads its own
// pass-B spectra, rooftop mask and divergence tables at the strides of ProdDims, the near field its own stencil
// classes and coefficient table. Each problem's arithmetic is that of its own product on the padded grid.
#include <algorithm>

#include "rooftop.hpp"

namespace rt {

namespace {

template <typename T>
void copy_device(T *dst, const T *src, size_t count, hipStream_t s) {
    if (count) RT_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToDevice, s));
}

}  // namespace

StackedProducts create_stacked(Context *ctx, const std::vector<std::shared_ptr<Problem>> &members, int held,
                               int preconditioner) {
    if (members.empty()) fail(RT_EINVAL, "rooftop:server", "an empty batch of problems");
    const int R = static_cast<int>(members.size());
    hipStream_t s = ctx->stream;
    const GridInfo &g = members[0]->g;
    for (const auto &m : members)
        if (m->g.cx != g.cx || m->g.cy != g.cy)
            fail(RT_EINVAL, "rooftop:server", "the problems of a batch share one face grid");
    StackedProducts out;
    out.problem = std::make_shared<Problem>();
    Problem &p = *out.problem;
    p.ctx = ctx;
    p.g = g;
    p.assembled = true;
    const int64_t n = g.n;
    p.exists.allocate(ctx, static_cast<size_t>(R) * n, "The rooftop masks of the batch");
    for (int i = 0; i < R; ++i) {
        if (members[i]->exists.ptr)
            copy_device(p.exists.ptr + static_cast<int64_t>(i) * n, members[i]->exists.ptr, n, s);
        else
            RT_HIP(hipMemsetAsync(p.exists.ptr + static_cast<int64_t>(i) * n, 1, n, s));
    }

    auto op = std::make_unique<Product>();
    op->type = kOperator;
    op->problem = &p;
    op->precision = held;
    op->dims = operator_dims(g);
    ProdDims &d = op->dims;
    d.spec_stride = static_cast<int64_t>(3) * d.Hx * d.My;
    d.mask_stride = n;
    d.dx_stride = d.Mx;
    d.dy_stride = d.My;
    // Twiddles, folds and position orders depend on the periods only; the divergence tables are the members' own.
    op->tx.build(ctx, d.Mx, 1.0, s);
    op->ty.build(ctx, d.My, 1.0, s);
    for (int v = 0; v < 2; ++v) {
        op->tx.p[v].d.allocate(ctx, static_cast<size_t>(R) * d.Mx, "divergence tables of the batch");
        op->tx.p[v].d32.allocate(ctx, static_cast<size_t>(R) * d.Mx, "divergence tables of the batch");
        op->ty.p[v].d.allocate(ctx, static_cast<size_t>(R) * d.My, "divergence tables of the batch");
        op->ty.p[v].d32.allocate(ctx, static_cast<size_t>(R) * d.My, "divergence tables of the batch");
    }
    // The pass-B spectra of the default plans, which every member built with its product.
    const bool c128 = held != RT_C64, c64 = held != RT_C128;
    const int v64 = default_bits(d, false) >= 4 ? 1 : 0, v32 = default_bits(d, true) >= 4 ? 1 : 0;
    if (c128) op->bspec64[v64].allocate(ctx, static_cast<size_t>(R) * d.spec_stride, "The spectra of the batch");
    if (c64) op->bspec32[v32].allocate(ctx, static_cast<size_t>(R) * d.spec_stride, "The spectra of the batch");
    std::vector<std::unique_ptr<Product>> pcs;
    for (int i = 0; i < R; ++i) {
        std::unique_ptr<Product> m = create_operator(members[i].get(), held, RT_MODIFIER_NONE);
        const ProdDims &md = m->dims;
        if (md.Mx != d.Mx || md.My != d.My || md.Hx != d.Hx)
            fail(RT_EINVAL, "rooftop:server", "the problems of a batch share one embedding");
        for (int v = 0; v < 2; ++v) {
            copy_device(op->tx.p[v].d.ptr + static_cast<int64_t>(i) * d.Mx, m->tx.p[v].d.ptr, d.Mx, s);
            copy_device(op->tx.p[v].d32.ptr + static_cast<int64_t>(i) * d.Mx, m->tx.p[v].d32.ptr, d.Mx, s);
            copy_device(op->ty.p[v].d.ptr + static_cast<int64_t>(i) * d.My, m->ty.p[v].d.ptr, d.My, s);
            copy_device(op->ty.p[v].d32.ptr + static_cast<int64_t>(i) * d.My, m->ty.p[v].d32.ptr, d.My, s);
        }
        if (c128) copy_device(op->bspec64[v64].ptr + i * d.spec_stride, m->bspec64[v64].ptr, d.spec_stride, s);
        if (c64) copy_device(op->bspec32[v32].ptr + i * d.spec_stride, m->bspec32[v32].ptr, d.spec_stride, s);
        if (preconditioner) pcs.push_back(create_preconditioner(members[i].get(), preconditioner, held));
        RT_HIP(hipStreamSynchronize(s));   // the copies have read the member's spectra before they are freed
    }
    out.op = std::move(op);
    if (!preconditioner) return out;
    if (preconditioner != RT_PRECOND_NEARFIELD)
        fail(RT_EINVAL, "rooftop:server", "a batch of problems takes the near-field preconditioner or none");
    auto pc = std::make_unique<Product>();
    pc->type = kNearField;
    pc->problem = &p;
    pc->precision = held;
    pc->dims = pcs[0]->dims;
    int classes = 0, stencils = 0;
    for (const auto &q : pcs) {
        classes = std::max(classes, q->class_count);
        stencils = std::max(stencils, q->stencil_count);
    }
    pc->class_count = classes;
    pc->stencil_count = stencils;
    pc->cls_stride = n;
    pc->coef_stride = static_cast<int64_t>(stencils) * kNearFieldSlots;
    pc->scls.allocate(ctx, static_cast<size_t>(R) * n, "The near-field classes of the batch");
    if (c128) {
        pc->scoef64.allocate(ctx, static_cast<size_t>(R) * pc->coef_stride, "The near-field coefficients of the batch");
        pc->scoef64.zero(s);
    }
    if (c64) {
        pc->scoef32.allocate(ctx, static_cast<size_t>(R) * pc->coef_stride, "The near-field coefficients of the batch");
        pc->scoef32.zero(s);
    }
    for (int i = 0; i < R; ++i) {
        const Product &q = *pcs[i];
        copy_device(pc->scls.ptr + static_cast<int64_t>(i) * n, q.scls.ptr, n, s);
        const size_t table = static_cast<size_t>(q.stencil_count) * kNearFieldSlots;
        if (c128) copy_device(pc->scoef64.ptr + i * pc->coef_stride, q.scoef64.ptr, table, s);
        if (c64) copy_device(pc->scoef32.ptr + i * pc->coef_stride, q.scoef32.ptr, table, s);
    }
    RT_HIP(hipStreamSynchronize(s));
    out.pc = std::move(pc);
    return out;
}

}  // namespace rt

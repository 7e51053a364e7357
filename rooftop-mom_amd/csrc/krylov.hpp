// Krylov solvers of librooftop (krylov.hip).
#pragma once

#include <functional>
#include <vector>

#include "rooftop.hpp"

namespace rt {

using LinOp = std::function<void(const c128 *, c128 *)>;

struct SolveOptions {
    int method = RT_METHOD_CG;
    double tolerance = 1e-6;
    int max_iterations = 1000;
    int restart = 50;
    int precision = RT_C128;
    int refine = 0;
    double inner_tolerance = 0;
    int max_outer = 20;
};

struct InnerResult {
    bool converged = false;
    int iterations = 0;
    std::vector<double> norms;
};

struct SolveResult {
    bool converged = false;
    int iterations = 0;
    int outer = 0;
    std::vector<double> norms;
    double residual = 0;
    double time = 0;
};

// Device scratch of one solve: reduction partials, device scalars and their pinned host mirror.
struct Workspace {
    Context *ctx;
    hipStream_t s;
    int64_t n;
    DeviceArray<c128> partial;
    DeviceArray<c128> scal;
    c128 *host = nullptr;
    Workspace(Context *c, int64_t n_);
    ~Workspace();
    Workspace(const Workspace &) = delete;
    Workspace &operator=(const Workspace &) = delete;
    int red_blocks() const;
    void dot(const c128 *x, const c128 *y, bool conj, int slot);
    void nrm2(const c128 *x, int slot);
    void fetch(int first, int count);
    void copy(const c128 *x, c128 *y);
    void scale(const c128 *x, c128 a, c128 *y);
    void sub(const c128 *b, const c128 *ax, c128 *r);
    void axpy(c128 a, const c128 *x, c128 *y);
    void zero(c128 *x);
};

SolveResult solve_one(Context *ctx, Product &op, Product *pc, const SolveOptions &o, const c128 *b, const c128 *x0,
                      c128 *x);

}  // namespace rt

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

// Device state of one Krylov loop, written by its finalize kernels and polled by the host.
struct LoopState {
    int done;         // converged, broke down or reached the iteration limit: every later kernel returns at once
    int converged;
    int iterations;   // iterations completed (global count)
    int steps;        // GMRES: steps of the current cycle completed
    int pad[4];
};

// Device scratch of one solve, taken from the context's buffer pool: reduction partials, device scalars, the loop
// state, the residual history and their pinned host mirror.
struct Workspace {
    Context *ctx;
    hipStream_t s;
    int64_t n;
    int blocks;          // blocks of every row-partitioned kernel
    int64_t chunk;       // rows per block (a multiple of 256)
    c128 *partial;       // [blocks][64]
    c128 *scal;          // 256 device scalars
    LoopState *state;
    double *history;     // residual norms by iteration
    int history_cap;
    c128 *host;          // pinned mirror (256 scalars)
    Workspace(Context *c, int64_t n_, int max_iter);
    void nrm2(const c128 *x, int slot);
    void dot(const c128 *x, const c128 *y, bool conj, int slot);
    void fetch(int first, int count);
    void copy(const c128 *x, c128 *y);
    void scale(const c128 *x, c128 a, c128 *y);
    void sub(const c128 *b, const c128 *ax, c128 *r);
    void axpy(c128 a, const c128 *x, c128 *y);
    void zero(c128 *x);
    LoopState read_state();
    void read_history(int first, int count, std::vector<double> &out);
};

SolveResult solve_one(Context *ctx, Product &op, Product *pc, const SolveOptions &o, const c128 *b, const c128 *x0,
                      c128 *x);

}  // namespace rt

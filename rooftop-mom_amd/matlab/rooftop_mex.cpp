// rooftop_mex.cpp -- the MATLAB gateway of librooftop (include/rooftop.h): one MEX function that dispatches on a
// command string, used by the classes in +rooftopMoM/+gpu beside the reference package rooftopMoM. Handles cross into
// MATLAB as uint64 scalars and are released by the classes' delete, as Product.delete releases its budget
// (src/+rooftopMoM/Product.m:19-23). Arrays use the interleaved complex API (mex -R2018a), the ABI's own layout:
// column-major (re, im) pairs, rooftop vectors x-directed first with x fastest (UniformRooftops.m:1-9,44-55).
//
// Build (MATLAB R2018a or later, librooftop.so built by rooftop-mom_amd/build.py):
//     mex -R2018a -I<repo>/include rooftop_mex.cpp -L<repo>/rooftop-mom_amd -lrooftop
//
// Commands (h = a uint64 handle; precision 0 = complex128, 1 = complex64, 2 = both held):
//     c = rooftop_mex('context', device, budgetBytes)
//     p = rooftop_mex('problem', c, countX, countY, sizeX, sizeY, k, [near far], mask)      mask: [] or logical
//     [mXX, mYY, e] = rooftop_mex('generators', p, countX, countY)
//     o = rooftop_mex('operator', p, kind, precision, modifier)       kind 0 fft, 1 splitFFT, 2 leanSplitFFT
//     o = rooftop_mex('precondition', p, kind, precision)             kind 1 jacobi, 2 nearField
//     o = rooftop_mex('circulant', p, kernel, builder, [bx by], modifier, precision)
//     o = rooftop_mex('matrix', c, A, precision)                      a dense N x N operator on the device
//     o = rooftop_mex('handle', c, n, fh)                             a function handle, called back per product
//     y = rooftop_mex('apply', o, X, precision)                       N x R
//     [stored, working] = rooftop_mex('memory', o)
//     [X, out] = rooftop_mex('solve', c, A, P, B, X0, opts)           opts: struct of rt_solve_opts fields;
//                                                                     out: struct array, one per column of B
//     rooftop_mex('free', kind, h)                                    kind 'product', 'problem' or 'context'
#include <cstdint>
#include <cstring>
#include <vector>

#include "mex.h"
#include "rooftop.h"

namespace {

void check(int status) {
    if (status != RT_OK) mexErrMsgIdAndTxt(rt_last_error_id(), "%s", rt_last_error_message());
}

template <typename T>
T *handle(const mxArray *a) {
    if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1)
        mexErrMsgIdAndTxt("rooftop:mex:invalidHandle", "A handle is a uint64 scalar.");
    return reinterpret_cast<T *>(static_cast<uintptr_t>(*mxGetUint64s(a)));
}

mxArray *to_handle(void *p) {
    mxArray *a = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
    *mxGetUint64s(a) = static_cast<uint64_t>(reinterpret_cast<uintptr_t>(p));
    return a;
}

int32_t scalar(const mxArray *a) { return static_cast<int32_t>(mxGetScalar(a)); }

// A complex double copy of a numeric array (MATLAB drops the imaginary part of real arrays).
mxArray *as_complex(const mxArray *a) {
    if (!mxIsDouble(a)) mexErrMsgIdAndTxt("rooftopMoM:product:sizeMismatch", "The input must be numeric double.");
    mxArray *c = mxDuplicateArray(a);
    if (!mxIsComplex(c) && !mxMakeArrayComplex(c)) mexErrMsgIdAndTxt("rooftop:mex:memory", "out of memory");
    return c;
}

double field(const mxArray *s, const char *name, double fallback) {
    const mxArray *f = mxGetField(s, 0, name);
    return f && !mxIsEmpty(f) ? mxGetScalar(f) : fallback;
}

// A MATLAB function handle as an rt_apply_fn: every column is passed to feval(fh, x) on the MATLAB thread (rt_solve
// and rt_product_apply call the callback on the thread that called them, which is MATLAB's). An error inside the
// handle must not unwind through the library: it is trapped, the output left zero, and raised by the command that
// called the library once the library has returned.
struct HandleOperator {
    mxArray *fh;
};
mxArray *g_callback_error = nullptr;   // the first MException of a handle during the current library call

void call_handle(void *user, const void *input, void *output, int64_t n, int64_t nrhs) {
    HandleOperator *h = static_cast<HandleOperator *>(user);
    const mxComplexDouble *in = static_cast<const mxComplexDouble *>(input);
    mxComplexDouble *out = static_cast<mxComplexDouble *>(output);
    for (int64_t r = 0; r < nrhs && !g_callback_error; ++r) {
        mxArray *x = mxCreateDoubleMatrix(static_cast<mwSize>(n), 1, mxCOMPLEX);
        std::memcpy(mxGetComplexDoubles(x), in + r * n, sizeof(mxComplexDouble) * n);
        mxArray *args[2] = {h->fh, x};
        mxArray *y = nullptr;
        g_callback_error = mexCallMATLABWithTrap(1, &y, 2, args, "feval");
        if (!g_callback_error && (!mxIsDouble(y) || static_cast<int64_t>(mxGetNumberOfElements(y)) != n)) {
            mxArray *id = mxCreateString("rooftopMoM:solver:sizeMismatch");
            mxArray *msg = mxCreateString("The operator must return a numeric vector of the unknowns' length.");
            mxArray *margs[2] = {id, msg};
            mexCallMATLAB(1, &g_callback_error, 2, margs, "MException");
            mxDestroyArray(id);
            mxDestroyArray(msg);
        }
        if (!g_callback_error) {
            mxArray *yc = as_complex(y);
            std::memcpy(out + r * n, mxGetComplexDoubles(yc), sizeof(mxComplexDouble) * n);
            mxDestroyArray(yc);
        }
        if (y) mxDestroyArray(y);
        mxDestroyArray(x);
    }
}

// Raises the error a handle trapped during the library call that just returned, if any.
void raise_callback_error() {
    if (!g_callback_error) return;
    mxArray *e = g_callback_error;
    g_callback_error = nullptr;
    mexCallMATLAB(0, nullptr, 1, &e, "throw");
}

std::vector<HandleOperator *> g_handles;   // kept for the life of the MEX file (persistent function handles)

}  // namespace

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs < 1 || !mxIsChar(prhs[0])) mexErrMsgIdAndTxt("rooftop:mex:usage", "rooftop_mex(command, ...)");
    char cmd[32];
    mxGetString(prhs[0], cmd, sizeof cmd);
    if (!std::strcmp(cmd, "context")) {
        rt_ctx *c = nullptr;
        check(rt_ctx_create(scalar(prhs[1]), static_cast<uint64_t>(mxGetScalar(prhs[2])), &c));
        plhs[0] = to_handle(c);
    } else if (!std::strcmp(cmd, "problem")) {
        const int32_t cx = scalar(prhs[2]), cy = scalar(prhs[3]);
        const double *orders = mxGetDoubles(prhs[7]);
        std::vector<uint8_t> mask;
        if (nrhs > 8 && !mxIsEmpty(prhs[8])) {
            if (static_cast<int64_t>(mxGetNumberOfElements(prhs[8])) != static_cast<int64_t>(cx) * cy)
                mexErrMsgIdAndTxt("rooftopMoM:solver:sizeMismatch", "The mask must hold countX x countY faces.");
            const mxLogical *m = mxGetLogicals(prhs[8]);
            mask.assign(m, m + static_cast<size_t>(cx) * cy);
        }
        rt_problem *p = nullptr;
        check(rt_problem_create(handle<rt_ctx>(prhs[1]), cx, cy, mxGetScalar(prhs[4]), mxGetScalar(prhs[5]),
                                mxGetScalar(prhs[6]), static_cast<int32_t>(orders[0]),
                                static_cast<int32_t>(orders[mxGetNumberOfElements(prhs[7]) - 1]),
                                mask.empty() ? nullptr : mask.data(), &p));
        plhs[0] = to_handle(p);
    } else if (!std::strcmp(cmd, "generators")) {
        rt_problem *p = handle<rt_problem>(prhs[1]);
        const int32_t cx = scalar(prhs[2]), cy = scalar(prhs[3]);
        plhs[0] = mxCreateDoubleMatrix(cx > 1 ? cx - 1 : 0, cx > 1 ? cy : 0, mxCOMPLEX);
        plhs[1] = mxCreateDoubleMatrix(cy > 1 ? cx : 0, cy > 1 ? cy - 1 : 0, mxCOMPLEX);
        plhs[2] = mxCreateDoubleMatrix(cx, cy, mxCOMPLEX);
        check(rt_problem_generators(p, cx > 1 ? mxGetComplexDoubles(plhs[0]) : nullptr,
                                    cy > 1 ? mxGetComplexDoubles(plhs[1]) : nullptr, mxGetComplexDoubles(plhs[2])));
    } else if (!std::strcmp(cmd, "operator")) {
        rt_product *o = nullptr;
        check(rt_product_create(handle<rt_problem>(prhs[1]), scalar(prhs[2]), scalar(prhs[3]), scalar(prhs[4]), &o));
        plhs[0] = to_handle(o);
    } else if (!std::strcmp(cmd, "precondition")) {
        rt_product *o = nullptr;
        check(rt_precond_create(handle<rt_problem>(prhs[1]), scalar(prhs[2]), scalar(prhs[3]), &o));
        plhs[0] = to_handle(o);
    } else if (!std::strcmp(cmd, "circulant")) {
        const double *bw = mxGetDoubles(prhs[4]);
        rt_product *o = nullptr;
        check(rt_precond_create_circulant(handle<rt_problem>(prhs[1]), scalar(prhs[2]), scalar(prhs[3]),
                                          static_cast<int32_t>(bw[0]), static_cast<int32_t>(bw[1]), scalar(prhs[5]),
                                          scalar(prhs[6]), &o));
        plhs[0] = to_handle(o);
    } else if (!std::strcmp(cmd, "matrix")) {
        mxArray *a = as_complex(prhs[2]);
        if (mxGetM(a) != mxGetN(a)) mexErrMsgIdAndTxt("rooftopMoM:solver:invalidOperator", "A square matrix.");
        rt_product *o = nullptr;
        check(rt_product_create_dense(handle<rt_ctx>(prhs[1]), static_cast<int64_t>(mxGetM(a)),
                                      mxGetComplexDoubles(a), scalar(prhs[3]), &o));
        mxDestroyArray(a);
        plhs[0] = to_handle(o);
    } else if (!std::strcmp(cmd, "handle")) {
        HandleOperator *h = new HandleOperator{mxDuplicateArray(prhs[3])};
        mexMakeArrayPersistent(h->fh);
        g_handles.push_back(h);
        rt_product *o = nullptr;
        check(rt_product_create_callback(handle<rt_ctx>(prhs[1]), static_cast<int64_t>(mxGetScalar(prhs[2])),
                                         call_handle, h, &o));
        plhs[0] = to_handle(o);
    } else if (!std::strcmp(cmd, "apply")) {
        mxArray *x = as_complex(prhs[2]);
        const mwSize n = mxGetM(x), r = mxGetN(x);
        plhs[0] = mxCreateDoubleMatrix(n, r, mxCOMPLEX);
        const int status = rt_product_apply(handle<rt_product>(prhs[1]), mxGetComplexDoubles(x),
                                            mxGetComplexDoubles(plhs[0]), static_cast<int64_t>(n),
                                            static_cast<int64_t>(r), scalar(prhs[3]));
        mxDestroyArray(x);
        raise_callback_error();
        check(status);
    } else if (!std::strcmp(cmd, "memory")) {
        uint64_t stored = 0, working = 0;
        check(rt_product_memory(handle<rt_product>(prhs[1]), &stored, &working));
        plhs[0] = mxCreateDoubleScalar(static_cast<double>(stored));
        if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(static_cast<double>(working));
    } else if (!std::strcmp(cmd, "solve")) {
        rt_ctx *c = handle<rt_ctx>(prhs[1]);
        rt_product *A = handle<rt_product>(prhs[2]);
        rt_product *P = mxIsEmpty(prhs[3]) ? nullptr : handle<rt_product>(prhs[3]);
        mxArray *b = as_complex(prhs[4]);
        mxArray *x0 = mxIsEmpty(prhs[5]) ? nullptr : as_complex(prhs[5]);
        const mxArray *s = prhs[6];
        rt_solve_opts o;
        o.method = static_cast<int32_t>(field(s, "method", RT_METHOD_CG));
        o.tolerance = field(s, "tolerance", 1e-6);
        o.max_iterations = static_cast<int32_t>(field(s, "maxIterationCount", 1000));
        const double restart = field(s, "restart", 50);
        o.restart = mxIsInf(restart) ? 0 : static_cast<int32_t>(restart);
        o.precision = static_cast<int32_t>(field(s, "precision", RT_C128));
        o.refine = static_cast<int32_t>(field(s, "refine", o.precision == RT_C64 ? 1 : 0));
        o.inner_tolerance = field(s, "innerTolerance", 0);
        o.max_outer = static_cast<int32_t>(field(s, "maxOuter", 0));
        const mwSize n = mxGetM(b), count = mxGetN(b);
        if (x0 && (mxGetM(x0) != n || mxGetN(x0) != count))
            mexErrMsgIdAndTxt("rooftopMoM:solver:sizeMismatch", "The initial guess must match the right-hand side.");
        plhs[0] = mxCreateDoubleMatrix(n, count, mxCOMPLEX);
        const int32_t cap = 2 * o.max_iterations + 64;
        std::vector<double> norms(static_cast<size_t>(cap) * count);
        std::vector<rt_output> out(count);
        for (mwSize r = 0; r < count; ++r) {
            std::memset(&out[r], 0, sizeof(rt_output));
            out[r].residual_norms = norms.data() + static_cast<size_t>(r) * cap;
            out[r].residual_norms_capacity = cap;
        }
        const int status = rt_solve(c, A, P, &o, mxGetComplexDoubles(b), static_cast<int64_t>(n),
                                    static_cast<int64_t>(count), x0 ? mxGetComplexDoubles(x0) : nullptr,
                                    mxGetComplexDoubles(plhs[0]), out.data());
        mxDestroyArray(b);
        if (x0) mxDestroyArray(x0);
        raise_callback_error();
        check(status);
        const char *names[] = {"converged", "iterationCount", "residualNorms", "residualNorm", "time", "outerCount"};
        plhs[1] = mxCreateStructMatrix(1, count, 6, names);
        for (mwSize r = 0; r < count; ++r) {
            const int32_t k = out[r].residual_norms_count < cap ? out[r].residual_norms_count : cap;
            mxArray *h = mxCreateDoubleMatrix(k, 1, mxREAL);
            std::memcpy(mxGetDoubles(h), out[r].residual_norms, sizeof(double) * k);
            mxSetField(plhs[1], r, "converged", mxCreateLogicalScalar(out[r].converged != 0));
            mxSetField(plhs[1], r, "iterationCount", mxCreateDoubleScalar(out[r].iteration_count));
            mxSetField(plhs[1], r, "residualNorms", h);
            mxSetField(plhs[1], r, "residualNorm", mxCreateDoubleScalar(out[r].residual_norm));
            mxSetField(plhs[1], r, "time", mxCreateDoubleScalar(out[r].time_s));
            mxSetField(plhs[1], r, "outerCount", mxCreateDoubleScalar(out[r].outer_count));
        }
    } else if (!std::strcmp(cmd, "free")) {
        char kind[16];
        mxGetString(prhs[1], kind, sizeof kind);
        if (!std::strcmp(kind, "product")) rt_product_destroy(handle<rt_product>(prhs[2]));
        else if (!std::strcmp(kind, "problem")) rt_problem_destroy(handle<rt_problem>(prhs[2]));
        else rt_ctx_destroy(handle<rt_ctx>(prhs[2]));
    } else {
        mexErrMsgIdAndTxt("rooftop:mex:unknownCommand", "Unknown command %s.", cmd);
    }
}

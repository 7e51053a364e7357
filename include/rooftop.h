/*
 * rooftop.h -- C ABI of librooftop, the MI355X (gfx950) iterative-solve path of the rooftop-basis method of moments.
 *
 * The entry points replace the reference MATLAB package rooftopMoM 3.5.0 on its hot path; each cites the reference
 * interface it stands in for. A MATLAB MEX gateway, a Python ctypes binding or any other FFI binds these symbols.
 *
 * Conventions
 *   - Host arrays are interleaved complex (double re, double im) unless a function says otherwise, column-major like
 *     MATLAB (mxGetComplexDoubles). A rooftop vector holds the x-directed rooftops first, then the y-directed ones,
 *     each on its own grid with the x index fastest (src/+rooftopMoM/+basisFunctions/UniformRooftops.m:1-9,44-55).
 *   - "_device" entry points take device pointers (complex128) and run on the context's HIP stream; they return when
 *     the work is enqueued unless stated.
 *   - Every function returns RT_OK (0) or a negative status. rt_last_error_id()/rt_last_error_message() (thread-local)
 *     give the reference's error identifier and a message for the last failure on the calling thread.
 *   - Handles are opaque; the library owns every device buffer behind them. Calls on one handle must not overlap.
 */
#ifndef ROOFTOP_H
#define ROOFTOP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes and the reference error identifiers they carry. */
#define RT_OK 0
#define RT_EINVAL (-1)     /* validator errors, e.g. rooftopMoM:solver:invalidRestart                   */
#define RT_ESIZE (-2)      /* rooftopMoM:product:sizeMismatch, rooftopMoM:solver:sizeMismatch          */
#define RT_ENOMEM (-3)     /* rooftopMoM:resources:insufficientMemory                                   */
#define RT_ESINGULAR (-4)  /* rooftopMoM:preconditioner:singular                                        */
#define RT_EHIP (-5)       /* a HIP runtime error; the message holds the HIP error string               */

/* Product kinds (rooftopMoM.Operator 'product', src/+rooftopMoM/Operator.m:21-26). All three compute the same
 * Toeplitz product; on the GPU they share one circulant embedding of period >= 2n-1 per split direction. */
#define RT_PRODUCT_FFT 0
#define RT_PRODUCT_SPLIT_FFT 1
#define RT_PRODUCT_LEAN_SPLIT_FFT 2

/* Arithmetic of a product or preconditioner: complex128, complex64, or both stored (mixed solves). */
#define RT_C128 0
#define RT_C64 1
#define RT_MIXED 2
/* Solver precision only: complex64 products as RT_C64, with the GMRES basis kept in complex128 (by default a
 * complex64 inner GMRES stores its basis in complex64; every Gram-Schmidt coefficient is FP64 either way). */
#define RT_C64_FP64_BASIS 3

/* Generator modifiers (rooftopMoM.Operator 'modifier', Generators.m:25-37). */
#define RT_MODIFIER_NONE 0
#define RT_MODIFIER_REAL 1
#define RT_MODIFIER_IMAG 2

/* Preconditioner kinds. The reference's preconditioners are circulant (Preconditioner.m:41-51); these two are the
 * north-star ones: Jacobi = diag(Z) (tests/+unit/+solver/ConjugateGradientTest.m:55-59) and the near-field
 * preconditioner on 3x3-cell neighbourhoods (symmetrized sparse approximate inverse). */
#define RT_PRECOND_JACOBI 1
#define RT_PRECOND_NEARFIELD 2

/* Circulant kernels and builders (Preconditioner.m:31-48, +preconditioner/+kernels, generate*Spectra.m). */
#define RT_CIRCULANT_STRANG 0
#define RT_CIRCULANT_TCHAN 1
#define RT_CIRCULANT_RCHAN 2
#define RT_CIRCULANT_HUCKLE 3
#define RT_CIRCULANT_SUPEROPTIMAL 4
#define RT_BUILDER_PARTWISE 0
#define RT_BUILDER_BLOCKWISE 1
#define RT_BUILDER_HYBRID 2

/* Methods (rooftopMoM.Solver 'method', Solver.m:27-32). */
#define RT_METHOD_CG 0
#define RT_METHOD_GMRES 1
/* Block COCG (no counterpart in the reference, which solves one right-hand side at a time): the nrhs <= 64
 * right-hand sides of rt_solve share one block Krylov space; a system that converges leaves the block. */
#define RT_METHOD_BLOCK_CG 2

typedef struct rt_ctx rt_ctx;
typedef struct rt_problem rt_problem;
typedef struct rt_product rt_product;

/* Solver options: rooftopMoM.Solver name-value inputs (Solver.m:17-38) plus the precision of the inner loop.
 *   precision       RT_C128, RT_C64 or RT_C64_FP64_BASIS: arithmetic of the products inside the Krylov loop.
 *   refine          0: the reference's loop as is (the true residual is reported, not enforced);
 *                   1: an FP64 outer loop on the true residual norm(b - A x)/norm(b) until it is below tolerance
 *                      (iterative refinement; required with precision RT_C64).
 *   inner_tolerance relative tolerance of each inner solve when refine = 1 (0 picks the tolerance itself, at least
 *                   1e-6 for complex64 GMRES and 1e-4 for complex64 COCG).
 *   max_outer       limit of outer refinement steps (0 picks 20). */
typedef struct rt_solve_opts {
    int32_t method;
    double tolerance;          /* default 1e-6 */
    int32_t max_iterations;    /* default 1000 */
    int32_t restart;           /* GMRES restart, <= 0 means none (Inf) */
    int32_t precision;
    int32_t refine;
    double inner_tolerance;
    int32_t max_outer;
} rt_solve_opts;

/* rooftopMoM.solver.Output (src/+rooftopMoM/+solver/Output.m:14-21) plus the refinement count. residual_norms, if
 * not NULL, receives up to residual_norms_capacity monitored norms (iteration_count + 1 of them per RHS). */
typedef struct rt_output {
    int32_t converged;
    int32_t iteration_count;
    double residual_norm;      /* true relative residual norm(b - A x)/norm(b), FP64 */
    double time_s;
    int32_t outer_count;
    int32_t residual_norms_count;
    int32_t residual_norms_capacity;
    double *residual_norms;
} rt_output;

/* Context: one device, one HIP stream, and the byte ceiling of the reference's Budget
 * (src/+rooftopMoM/+resources/Budget.m:82-106; 0 = the device's free memory). */
int rt_ctx_create(int32_t device, uint64_t hbm_budget_bytes, rt_ctx **out);
void rt_ctx_destroy(rt_ctx *ctx);
/* Resources.held()/ceiling() (src/+rooftopMoM/Resources.m:55-71), in device bytes. */
int rt_ctx_memory(rt_ctx *ctx, uint64_t *held, uint64_t *ceiling);
/* The most device bytes the context held at once since it was created or since the last call with reset != 0 (a
 * reset starts the next interval at the bytes held now): the peak that Resources.held() (Resources.m:55-71) would
 * have shown, for checking a byte estimate against what a solve actually holds. */
int rt_ctx_memory_peak(rt_ctx *ctx, int32_t reset, uint64_t *peak);
/* Blocks until the context's stream is idle. */
int rt_ctx_synchronize(rt_ctx *ctx);
/* Device address of the context's HIP stream (hipStream_t), for callers that time or order work on it. */
int rt_ctx_stream(rt_ctx *ctx, void **stream);

/* rooftopMoM.Problem(mesh, k, 'quadratureOrder', [near far]) (src/+rooftopMoM/Problem.m:25-92) on a uniform
 * countX x countY grid of sizeX x sizeY faces (rooftopMoM.Mesh, Mesh.m:10-30). The generators are assembled on the
 * GPU on the first call that needs them (Problem.generators, Problem.m:135-145). mask (countX*countY bytes,
 * column-major, NULL = full plate) keeps the rooftops whose two faces are both metal. */
int rt_problem_create(rt_ctx *ctx, int32_t count_x, int32_t count_y, double size_x, double size_y,
                      double wavenumber, int32_t near_order, int32_t far_order, const uint8_t *mask,
                      rt_problem **out);
void rt_problem_destroy(rt_problem *problem);
/* Problem.rooftopCount (Problem.m:123-125): the unknowns of the full grid (a mask only zeroes rooftops). */
int rt_problem_count(rt_problem *problem, int64_t *count);
/* Problem.generators (Problem.m:135-145) copied to the host: magneticXX (countX-1)xcountY, magneticYY
 * countX x(countY-1), electric countX x countY, complex128 column-major; a NULL pointer skips that generator. */
int rt_problem_generators(rt_problem *problem, void *magnetic_xx, void *magnetic_yy, void *electric);
/* Replaces the generators by given ones (as Generators(...) is built from stored data in
 * tests/+reference/+assembler/+greensMatrix/ApplicatorTest.m:9-14); a missing direction takes NULL. */
int rt_problem_set_generators(rt_problem *problem, const void *magnetic_xx, const void *magnetic_yy,
                              const void *electric);

/* rooftopMoM.Operator(problem, 'product', kind, 'modifier', m) (src/+rooftopMoM/Operator.m:7-54). */
int rt_product_create(rt_problem *problem, int32_t kind, int32_t precision, int32_t modifier, rt_product **out);
/* A preconditioner product (the role of rooftopMoM.Preconditioner, src/+rooftopMoM/Preconditioner.m:7-70);
 * RT_ESINGULAR mirrors rooftopMoM:preconditioner:singular. */
int rt_precond_create(rt_problem *problem, int32_t kind, int32_t precision, rt_product **out);
/* rooftopMoM.Preconditioner(problem, 'kernel', k, 'builder', b, 'bandwidth', [bx by], 'modifier', m)
 * (src/+rooftopMoM/Preconditioner.m:7-70, +preconditioner/generateCirculant.m:16-24): the inverse of a block circulant
 * approximation of the Green's matrix on the face grid, applied by FFTs of the face grid (FFTApplicator.m): the
 * partwise, blockwise and hybrid builders (generatePartwiseSpectra.m:16-30, generateBlockwiseSpectra.m,
 * generateHybridSpectra.m) with the strang, tChan, rChan, huckle and superoptimal kernels (+preconditioner/+kernels),
 * on any face grid (power-of-two counts run on the face grid, others through a re-embedded Toeplitz product); a
 * four-plane product whose columns do not fit one block's LDS returns RT_EINVAL with rooftop:preconditioner:unsupported.
 * bandwidth_x/y: Huckle's bandwidths in faces (<= 0: the grid's counts);
 * RT_ESINGULAR with rooftopMoM:preconditioner:singular when the approximation is singular at some frequency. */
int rt_precond_create_circulant(rt_problem *problem, int32_t kernel, int32_t builder, int32_t bandwidth_x,
                                int32_t bandwidth_y, int32_t modifier, int32_t precision, rt_product **out);
/* The operator forms of rooftopMoM.solver.utils.convertToHandle (src/+rooftopMoM/+solver/+utils/convertToHandle.m:10-16,
 * used by Solver.solve, Solver.m:146-154) for rt_solve, as operator or preconditioner:
 *   rt_product_create_dense     a matrix: n x n complex128, column-major (copied to the device; RT_C64 / RT_MIXED keep
 *                               a complex64 copy for complex64 products);
 *   rt_product_create_callback  a function handle or an object with apply: apply(user, input, output, n, nrhs) is
 *                               called on the calling thread of rt_solve / rt_product_apply with host arrays of nrhs
 *                               column-major vectors (output zeroed before the call), once per product application
 *                               of the Krylov loop; it computes in its own arithmetic whatever the loop's precision. */
typedef void (*rt_apply_fn)(void *user, const void *input, void *output, int64_t n, int64_t nrhs);
int rt_product_create_dense(rt_ctx *ctx, int64_t n, const void *matrix, int32_t precision, rt_product **out);
int rt_product_create_callback(rt_ctx *ctx, int64_t n, rt_apply_fn apply, void *user, rt_product **out);
void rt_product_destroy(rt_product *product);
/* Product.apply (src/+rooftopMoM/Product.m:25-35) on nrhs column-major host vectors of n rooftops each; n must be
 * the rooftop count (RT_ESIZE, rooftopMoM:product:sizeMismatch). Synchronous. precision selects the stored
 * arithmetic for an RT_MIXED product (RT_C128 or RT_C64); otherwise pass the product's own. */
int rt_product_apply(rt_product *product, const void *input, void *output, int64_t n, int64_t nrhs,
                     int32_t precision);
/* The same on device buffers (complex128), enqueued on the context's stream. */
int rt_product_apply_device(rt_product *product, const void *input, void *output, int64_t n, int64_t nrhs,
                            int32_t precision);
/* rt_product_apply_device without any device allocation (no counterpart in the reference; for callers that cannot
 * allocate: code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The matvec passes
 * run with their working memory in a file-scope device array of the library, 64 MiB per device (held from the first
 * context on the device until the process ends, outside every context's budget), instead of the product's own work
 * buffer; the caller passes the device input and output. Limits: the FFT products only (operators and circulant
 * preconditioners); nrhs times the working bytes of one right-hand side in the requested precision (complex64 half
 * of complex128; rt_product_memory reports the working bytes of the precision held) within the 64 MiB (RT_EINVAL,
 * rooftop:product:noAlloc, beyond it); y transforms that fit one block's LDS; and the product applied once by
 * rt_product_apply_device with the same batch size and precision before (which builds the spectra of its launch
 * shape). Every call on one device shares that array: calls that overlap on one device race. */
int rt_product_apply_device_noalloc(rt_product *product, const void *input, void *output, int64_t n, int64_t nrhs,
                                    int32_t precision);
/* Product.memory / workingMemory (Product.m:47-57): device bytes of the stored spectra or coefficients, and of the
 * buffers one application uses. */
int rt_product_memory(rt_product *product, uint64_t *stored, uint64_t *working);

/* Applicator.spectra() (src/+rooftopMoM/+assembler/+greensMatrix/Applicator.m:20-22) of a Green's matrix product:
 * dims (4 entries, may be NULL) receives the embedding periods Mx, My and the folded extents Hx = Mx/2 + 1,
 * Hy = My/2 + 1; planes (may be NULL, else 3 Hx Hy complex128) the stored spectra in this library's three-plane form
 * (magneticXX, magneticYY, electric; each the 2D DFT of the generator's even embedding of period Mx x My, scaled by
 * 1/(Mx My), folded to frequencies [0, Mx/2] x [0, My/2], column-major). RT_EINVAL for a preconditioner. */
int rt_product_spectra(rt_product *product, int32_t *dims, void *planes);
/* The static Applicator.estimateWorkingMemory(rooftops, sizeX, sizeY, content) (Applicator.m:44-46,
 * FFTApplicator.m:36-49) with the stored bytes beside it, before any spectra exist: the device bytes that
 * rt_product_create(precision) of a count_x x count_y grid stores (stored) and that one application of it uses for
 * one right-hand side (working), as rt_product_memory reports them once it exists. */
int rt_product_estimate_memory(int32_t count_x, int32_t count_y, int32_t precision, uint64_t *stored,
                               uint64_t *working);

/* Measurement (no counterpart in the reference): runs `reps` applications of the operator on nrhs device vectors
 * and times each of its three kernels (pass A: x transforms, pass B: y transforms + spectral multiply, pass C:
 * inverse x transforms) with HIP events on the streams they run on. ms (4 entries): ms[k] (k < 3) is the kernel time
 * of pass k in one application, summed over the application's right-hand-side chunks, and ms[3] the wall time of one
 * application (less than the sum when chunks overlap); bytes (3 entries): bytes[k] the algorithmic HBM bytes of pass
 * k in one application (vector and spectra traffic, each element once). */
int rt_product_profile(rt_product *product, int64_t nrhs, int32_t reps, int32_t precision, double *ms,
                       double *bytes);

/* Autotuner (no counterpart in the reference; the north star's FFT tile and radix schedule): times the launch shapes
 * of the three passes for nrhs right-hand sides in the given precision -- plan radices up to 8 or 16, pass A/C
 * rooftop columns per block, pass B x positions per block and teams per position, right-hand-side chunks -- and
 * keeps the fastest for every later application with that many right-hand sides. shape (may be NULL, else 5
 * entries) receives {largest radix bits, pass A/C columns per block (side by side in LDS; -2 or -4: that many in turn
 * through one column of LDS), pass B positions per block, pass B form (1: one team
 * transforms both columns of a position in LDS, 2: a team per column, 3: one team, the two columns in turn through
 * one column of LDS), chunks (the right-hand sides split into chunks
 * that alternate between two streams, so that one chunk's passes overlap another's)}; ms (may be NULL) the wall
 * time of one application with it. */
int rt_product_tune(rt_product *product, int64_t nrhs, int32_t precision, int32_t *shape, double *ms);
/* Sets the launch shape (5 entries, as rt_product_tune returns them) used for nrhs right-hand sides in the given
 * precision;
 * RT_EINVAL when the shape is malformed or does not fit the device for this product. */
int rt_product_set_shape(rt_product *product, int64_t nrhs, int32_t precision, const int32_t *shape);

/* Measurement (no counterpart in the reference): while timing is on, the timed kernels below record a pair of HIP
 * events around each launch on the stream they run on, with the launch's algorithmic HBM bytes (every element it
 * must read or write counted once). rt_ctx_kernel_timing drains them (synchronizing the stream) and reports, per
 * slot, the launches counted, the average duration of one launch in ms and its average algorithmic bytes; then it
 * resets the counters. The arrays hold RT_TIMED_COUNT entries. */
enum {
    RT_TIMED_PASS_A_C128 = 0,   /* matvec pass A (x transforms), complex128 */
    RT_TIMED_PASS_B_C128 = 1,   /* matvec pass B (y transforms + spectral multiply) */
    RT_TIMED_PASS_C_C128 = 2,   /* matvec pass C (inverse x transforms) */
    RT_TIMED_PASS_A_C64 = 3,
    RT_TIMED_PASS_B_C64 = 4,
    RT_TIMED_PASS_C_C64 = 5,
    RT_TIMED_GS_DOTS = 6,         /* GMRES Gram-Schmidt dot pass: [s z] = V^H [p w] and the norms of p and w */
    RT_TIMED_GS_UPDATE = 7,       /* update pass: q = (p - V s) / beta, p = (w - V z - z_j q) / (beta nu), and beside
                                     it the step's Hessenberg columns, rotations and stopping test */
    RT_TIMED_PRECOND_C128 = 8,    /* preconditioner application */
    RT_TIMED_PRECOND_C64 = 9,
    RT_TIMED_GS_SUMS = 10,        /* the dot pass's block sums and the step's coefficients (beta, z_j, nu) */
    RT_TIMED_COUNT = 11
};
int rt_ctx_set_timing(rt_ctx *ctx, int32_t on);
int rt_ctx_kernel_timing(rt_ctx *ctx, int64_t *launches, double *ms, double *bytes);
/* Measurement (no counterpart in the reference): a plain device copy kernel of `bytes` bytes (one read, one write
 * per byte), `reps` times after one warm-up; *ms receives the average duration. 2 bytes / ms is the streaming HBM
 * bandwidth of this device, and the copy calibrates the rocprofv3 FETCH_SIZE / WRITE_SIZE counters. */
int rt_bandwidth_probe(rt_ctx *ctx, uint64_t bytes, int32_t reps, double *ms);

/* rooftopMoM.Solver(...).solve(operator, rhs, 'preconditioner', P, 'initialGuess', x0) (Solver.m:42-92,
 * IterativeMethod.m:27-70). precond may be NULL; x0 may be NULL (zeros). The nrhs (1..1024) right-hand sides are
 * solved together into x (n x nrhs): one Krylov loop per right-hand side, all advanced in lockstep so that every
 * product and preconditioner launch serves all of them; each stops at its own iteration with the result it would
 * have alone. out (may be NULL) points to nrhs records, one per right-hand side. */
int rt_solve(rt_ctx *ctx, rt_product *op, rt_product *precond, const rt_solve_opts *opts, const void *rhs,
             int64_t n, int64_t nrhs, const void *x0, void *x, rt_output *out);
/* The same on device buffers; synchronous (the loop polls its scalars). */
int rt_solve_device(rt_ctx *ctx, rt_product *op, rt_product *precond, const rt_solve_opts *opts, const void *rhs,
                    int64_t n, int64_t nrhs, const void *x0, void *x, rt_output *out);

/* Serving (no counterpart in the reference): a library-owned request loop. Requests of
 * different grids, frequencies and masks are admitted first come, first served when a worker slot is free and their
 * estimated device bytes fit the server's HBM budget (the semantics of Budget.require/hold/release,
 * src/+rooftopMoM/+resources/Budget.m:59-106); each slot has its own HIP stream, so admitted requests run
 * concurrently, and a request holds its generators, spectra, preconditioner and Krylov state only while it runs. A
 * request's result is the one rt_solve gives on the same inputs, whatever runs beside it. */
typedef struct rt_server rt_server;
typedef struct rt_request {
    int32_t count_x, count_y;    /* faces */
    double size_x, size_y;       /* face sizes (m) */
    double wavenumber;           /* k (rad/m) */
    int32_t near_order, far_order;   /* quadrature orders, 0 = [16, 8] */
    const uint8_t *mask;         /* count_x * count_y faces (x fastest), 1 = metal; NULL = full plate */
    const void *rhs;             /* n complex128 values (copied at submission) */
    int64_t n;                   /* rooftops of the grid with this mask's layout (all edges of the full grid) */
    int32_t preconditioner;      /* 0, RT_PRECOND_JACOBI or RT_PRECOND_NEARFIELD */
    rt_solve_opts opts;
} rt_request;
/* slots: concurrent requests (1..64); hbm_budget_bytes 0 = 90 % of the device's free memory. */
int rt_server_create(int32_t device, uint64_t hbm_budget_bytes, int32_t slots, rt_server **out);
/* Waits for the running requests; queued ones end with an error for their waiters. */
void rt_server_destroy(rt_server *server);
/* Copies the request and returns at once; RT_ENOMEM when it can never fit the budget. */
int rt_server_submit(rt_server *server, const rt_request *request, int64_t *ticket);
/* Blocks until the request has run; x (may be NULL) receives the solution (n complex128), out the solver output and
 * times (may be NULL) {queue wait, service, total latency} in seconds. A ticket is waited for once. */
int rt_server_wait(rt_server *server, int64_t ticket, void *x, int64_t n, rt_output *out, double *times);
/* Batching by embedding size: queued requests whose FFT periods (next power of two >= 2 count - 1 per direction,
 * at most 256: larger grids keep the device busy alone and are solved alone), face aspect, quadrature orders,
 * preconditioner (near field or none) and solver options are equal join the request
 * a slot takes, up to max_requests (1..64; 1, the default, solves every request alone) and while the group's bytes
 * fit the budget. A group runs on the largest face counts among its requests (the others padded with empty faces),
 * each request with its own generators, spectra, near field and Krylov loop, all of them in one launch per pass of
 * the product and of every Krylov kernel; its result equals its solve alone up to the rounding of the padding. */
int rt_server_set_batching(rt_server *server, int32_t max_requests);
/* groups: the solves that held more than one request; grouped: the requests they held. */
int rt_server_batch_stats(rt_server *server, int64_t *groups, int64_t *grouped);
/* The device bytes the server reserves for a request while it runs (its admission size). */
int rt_request_bytes(const rt_request *request, uint64_t *bytes);
/* counts: {submitted, completed, queued, running, peak running}; bytes: {reserved, budget}. */
int rt_server_stats(rt_server *server, int64_t *counts, uint64_t *bytes);
/* A 4-connected blob of metal faces grown by random frontier growth from a uniformly drawn seed face until
 * round(fill * count_x * count_y) faces are metal (seeded, portable; the masks of the served-stream workload). */
int rt_blob_mask(int32_t count_x, int32_t count_y, uint64_t seed, double fill, uint8_t *mask);

/* Batched right-hand sides and observables on the device. No counterpart in the reference, which builds one
 * excitation and reads one result at a time on the host; they replace, for a batch already in HBM:
 *   rt_plane_waves_device          generatePlaneWave.m:18-42 (Excitation.planeWave, Excitation.m:29-72) for `count`
 *                                  incidences: directions and polarizations are host arrays [count][3] (normalised
 *                                  here; rooftopMoM:excitation:zeroVector / notPerpendicular as the reference),
 *                                  intensities host [count] or NULL (1); b the device batch, n x count complex128;
 *   rt_radar_cross_sections_device computeRadarCrossSection.m:18-19: sigma[r] = (k Z0)^2 |b_r.' x_r|^2 / (4 pi E_r^4)
 *                                  from the device batches b and x (n x count); sigma host [count];
 *   rt_input_impedances_device     computeInputImpedance.m:17: Z_in[r] = V_r / (edge lengths . x_r(port)); port host
 *                                  [port_len] distinct 0-based rooftop indices (rooftopMoM:excitation:invalidPort),
 *                                  voltages host [count] or NULL (1 V), x the device batch; zin host [count][2]
 *                                  (real, imaginary). */
int rt_plane_waves_device(rt_problem *problem, int64_t count, const double *directions, const double *polarizations,
                          const double *intensities, void *b);
int rt_radar_cross_sections_device(rt_problem *problem, int64_t count, const void *b, const void *x,
                                   const double *intensities, double *sigma);
int rt_input_impedances_device(rt_problem *problem, int64_t count, const int64_t *port, int64_t port_len,
                               const double *voltages, const void *x, double *zin);

/* Device buffers owned by the caller through the library (counted in the context's budget). */
int rt_device_alloc(rt_ctx *ctx, uint64_t bytes, void **ptr);
int rt_device_free(rt_ctx *ctx, void *ptr);
int rt_copy_to_device(rt_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int rt_copy_to_host(rt_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* Thread-local identifier (e.g. "rooftopMoM:product:sizeMismatch") and message of the last error. */
const char *rt_last_error_id(void);
const char *rt_last_error_message(void);
/* Library version string. */
const char *rt_version(void);

#ifdef __cplusplus
}
#endif

#endif /* ROOFTOP_H */

// Internal objects of librooftop: the problem (grid, wavenumber, generators on the device), the Toeplitz product
// (circulant embedding spectra and the three-pass FFT matvec) and the preconditioners.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"
#include "fft.hpp"

namespace rt {

// Grid of countX x countY faces of sizeX x sizeY and its rooftop layout (UniformRooftops.m:31-101): x-directed
// rooftops on a (countX-1) x countY grid first, then y-directed ones on countX x (countY-1), x index fastest.
struct GridInfo {
    int cx = 0, cy = 0;
    double a = 0, b = 0;
    int rows_x = 0, cols_x = 0, rows_y = 0, cols_y = 0;
    int64_t nx = 0, ny = 0, n = 0;
    bool has_x = false, has_y = false;
    static GridInfo make(int cx, int cy, double a, double b);
};

// Dimensions of one product, passed by value to the kernels.
struct ProdDims {
    int rows_x, rows_y;       // rooftop rows (x extent) of each component
    int cols_x, cols_y;       // rooftop columns (y extent), 0 for a missing component
    int ny;                   // columns of the work array per component (max of cols_x, cols_y)
    int ldj;                  // row length of the work array (ny rounded up to 8)
    int64_t nx_count;         // number of x-directed rooftops (offset of the y part of a vector)
    int64_t n;                // rooftops per right-hand side
    int Mx, My;               // embedding periods
    int Hx, Hy;               // folded spectrum extents M/2 + 1
    // A batch of problems on one grid (the served requests of one embedding size): right-hand side r reads its own
    // pass-B spectra at spec_stride r elements and its own rooftop mask at mask_stride r bytes (0: shared by all).
    int64_t spec_stride = 0;
    int64_t mask_stride = 0;
    int64_t dx_stride = 0, dy_stride = 0;   // and its own divergence tables (the face sizes differ)
};

struct Problem {
    Context *ctx = nullptr;
    GridInfo g;
    double k = 0;
    int near_order = 16, far_order = 8;
    std::vector<uint8_t> face_mask;          // countX*countY, empty for a full plate
    DeviceArray<uint8_t> exists;             // per rooftop (n), empty for a full plate
    DeviceArray<c128> mxx, myy, e;           // generators, column-major
    bool assembled = false;
    void ensure_generators();                // assembles on the GPU on first use (Problem.generators)
    void assemble();                         // generator.hip
};

// Per-dimension tables of a product: twiddles, fold index and divergence factor D(k) = (1 - exp(2 pi i k/M))/size
// at each digit-reversed output position.
// Tables of one transform direction for one plan.
struct PlanTables {
    FftPlan plan;
    DeviceArray<c128> tw;     // twiddles w_L^(q n) of the rounds
    DeviceArray<c64> tw32;
    DeviceArray<int> fold;    // fold index min(k, M - k) of the frequency k held at each position
    DeviceArray<int> border;  // the positions ordered by fold index: the two positions of a folded spectrum row
                              // (frequencies k and M - k) are neighbours, so pass B reads the row for both at once
    DeviceArray<c128> d;      // divergence factor (1 - exp(2 pi i k / M)) / size at each position
    DeviceArray<c64> d32;
};

// The tables of one direction for the plans with radices up to 8 and up to 16.
struct DimTables {
    PlanTables p[2];
    void build(Context *ctx, int M, double size, hipStream_t s);
    const PlanTables &of(int bits) const { return p[bits >= 4 ? 1 : 0]; }
};

// Launch shape of the three matvec passes (picked by the autotuner or the default heuristic).
struct PassShape {
    int bits = 4;    // largest radix of the plans: 2^bits elements per thread
    int group = 1;   // pass A / C: rooftop columns per block, side by side in LDS; -2, -4: in turn through one column
    int tile = 1;    // pass B: x positions per block
    int teams = 1;   // pass B form: 1, one team runs both columns of an x position; 2, a team per column; 3, one
                     // team whose columns take turns in one column of LDS
    int chunks = 1;  // right-hand-side chunks, alternating between the main and the side stream
};

enum ProductType { kOperator = 0, kJacobi = 1, kNearField = 2, kCirculant = 3, kDense = 4, kCallback = 5 };

struct Product {
    int type = kOperator;
    Problem *problem = nullptr;
    int precision = RT_C128;
    int modifier = RT_MODIFIER_NONE;
    // The FFT passes run on an even embedding of period >= 2n - 1 (the Toeplitz product: half of every column is
    // zero padding or cut away) or, false, on the face grid itself (a circulant preconditioner).
    bool embedding = true;
    ProdDims dims{};
    // Operator.
    DimTables tx, ty;
    DeviceArray<c128> planes64;      // [3][Hx][Hy]: magneticXX, magneticYY, electric, scaled by 1/(Mx My)
    // The same spectra as pass B reads them, per plan (radices up to 8, up to 16): [Hx][My][3], x folded, y in the
    // plan's position order, the three values of a position together. Built when a plan is first used.
    DeviceArray<c128> bspec64[2];
    DeviceArray<c64> bspec32[2];
    // A blockwise circulant preconditioner: the inverse blocks (XX, XY, YX, YY) per frequency, unfolded, as pass B
    // reads them: [x position][y position][4] in the radix-8 plans' orders.
    bool four_plane = false;
    DeviceArray<c128> b4spec64;
    DeviceArray<c64> b4spec32;
    DeviceArray<c128> spec64;        // [nrhs][2][Mx][My]: full y spectra of the three-launch pass B (large periods)
    DeviceArray<c64> spec32;
    int spec_rhs64 = 0, spec_rhs32 = 0;
    DeviceArray<c128> work64;        // [nrhs][2][Mx][ldj]
    DeviceArray<c64> work32;
    int work_rhs64 = 0, work_rhs32 = 0;
    std::vector<std::pair<int, PassShape>> tuned64, tuned32;   // shapes picked by the autotuner, per batch size
    // Preconditioners.
    DeviceArray<c128> diag64;                 // Jacobi: 1/diag(Z) per rooftop
    DeviceArray<c64> diag32;
    DeviceArray<c128> coef64;                 // near field: [classes][18] symmetrized-SAI row coefficients
    DeviceArray<c64> coef32;
    DeviceArray<int> cls;                     // near field: class of every rooftop (-1: no rooftop)
    int class_count = 0;
    DeviceArray<int> scls;                    // near field: stencil class of every rooftop (-1: no rooftop)
    DeviceArray<c128> scoef64;                // near field: [stencil classes][18] symmetrized coefficients
    DeviceArray<c64> scoef32;
    int stencil_count = 0;
    int64_t cls_stride = 0, coef_stride = 0;   // a batch of problems: per right-hand side classes and tables
    // A dense operator (n x n, column-major) and a host callback (convertToHandle.m:10-16).
    DeviceArray<c128> dense64;
    DeviceArray<c64> dense32;
    rt_apply_fn callback = nullptr;
    void *callback_user = nullptr;
    void *staging = nullptr;   // pinned host copies of a callback's input and output
    size_t staging_bytes = 0;

    Product() = default;
    Product(const Product &) = delete;
    Product &operator=(const Product &) = delete;
    ~Product() {
        if (staging) (void)hipHostFree(staging);
    }
    bool has(int prec) const { return precision == RT_MIXED || precision == prec; }
    // skip: optional device flags, one per right-hand side r at skip[r * skip_stride] (stride 0: one flag for all);
    // the kernels of a right-hand side whose flag is non-zero at run time return at once.
    void apply(const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s, const int *skip = nullptr,
               int skip_stride = 0);
    uint64_t stored_bytes() const;
    uint64_t working_bytes() const;
};

// abi.cpp: a problem on a grid of cx x cy faces of a x b, wavenumber k, optional face mask (cx * cy, x fastest).
std::shared_ptr<Problem> make_problem(Context *ctx, int cx, int cy, double a, double b, double k, int near_order,
                                      int far_order, const uint8_t *mask);
// product.hip
// The dimensions of the Green's matrix product of a grid: rooftop rows and columns, the power-of-two embedding
// periods M >= 2 count - 1 and the folded spectrum extents M / 2 + 1.
ProdDims operator_dims(const GridInfo &g);
// Radix bits (3: radices up to 8, 4: up to 16) of the default plan of a product's transforms in complex64 or
// complex128; its pass-B spectra are built with the product.
int default_bits(const ProdDims &d, bool c64);
std::unique_ptr<Product> create_operator(Problem *p, int precision, int modifier);
// The circulant preconditioner of the problem on its face grid (RT_CIRCULANT_* kernel, RT_BUILDER_* builder).
std::unique_ptr<Product> create_circulant(Problem *p, int kernel, int builder, int bandwidth_x, int bandwidth_y,
                                          int modifier, int precision);
void operator_apply(Product &op, const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s, const int *skip,
                    int skip_stride);
// The same with the caller's working memory (work_bytes at work), allocating nothing.
void operator_apply_noalloc(Product &op, const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s, void *work,
                            size_t work_bytes);
// noalloc.hip: the allocation-free application with the library's file-scope working memory; noalloc_setup reads
// that memory's address on the current device once (at context creation).
void noalloc_setup(int device);
void product_apply_noalloc(Product &op, const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s);
void operator_profile(Product &op, int nrhs, int reps, int prec, double *ms, double *bytes, hipStream_t s);
// Times the launch shapes of the passes for nrhs right-hand sides and keeps the fastest for that batch size;
// shape (may be NULL) receives {bits, group, tile}, ms the summed pass time of one application.
void operator_tune(Product &op, int nrhs, int prec, int *shape, double *ms, hipStream_t s);
// calibrate.hip: average ms of one plain device copy of `bytes`.
double bandwidth_probe(Context *ctx, uint64_t bytes, int reps);
// Sets the shape {bits, group, tile} used for nrhs right-hand sides (a saved tuning result).
void operator_set_shape(Product &op, int nrhs, int prec, const int *shape);
// dense.hip: a dense n x n operator from a host matrix (complex128, column-major), a host-callback operator, and
// their application; owner receives the product's bare problem (the context and n).
std::unique_ptr<Product> create_dense(Context *ctx, int64_t n, const c128 *host_matrix, int precision,
                                      std::shared_ptr<Problem> &owner);
std::unique_ptr<Product> create_callback(Context *ctx, int64_t n, rt_apply_fn fn, void *user,
                                         std::shared_ptr<Problem> &owner);
void dense_apply(Product &op, const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s, const int *skip,
                 int skip_stride);
// nearfield.hip
constexpr int kNearFieldSlots = 18;   // coefficients of a near-field stencil class
std::unique_ptr<Product> create_preconditioner(Problem *p, int kind, int precision);
// stack.hip: a batch of problems on one padded face grid (same embedding) as one operator and one near-field
// preconditioner (preconditioner 0: none), each right-hand side reading its own problem's data.
struct StackedProducts {
    std::shared_ptr<Problem> problem;   // the common grid and the stacked masks
    std::unique_ptr<Product> op, pc;
};
StackedProducts create_stacked(Context *ctx, const std::vector<std::shared_ptr<Problem>> &members, int held,
                               int preconditioner);
void preconditioner_apply(Product &pc, const c128 *in, c128 *out, int nrhs, int prec, hipStream_t s,
                          const int *skip, int skip_stride);

// Gauss-Legendre rule of n points on [0, 1] (generator.hip).
void gauss_legendre(int n, std::vector<double> &pts, std::vector<double> &wts);

// Batched right-hand sides and observables on the device (excitation.hip).
void plane_waves_device(Problem &p, int64_t count, const double *directions, const double *polarizations,
                        const double *intensities, c128 *b, hipStream_t s);
void radar_cross_sections_device(Problem &p, int64_t count, const c128 *b, const c128 *x, const double *intensities,
                                 double *sigma, hipStream_t s);
void input_impedances_device(Problem &p, int64_t count, const int64_t *port, int64_t port_len, const double *voltages,
                             const c128 *x, double *zin, hipStream_t s);

}  // namespace rt

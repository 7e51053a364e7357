// Krylov solvers of librooftop (krylov.hip): COCG and GMRES(m) on R right-hand sides at once.
#pragma once

#include <functional>
#include <vector>

#include "rooftop.hpp"

namespace rt {

// A batched linear map on R systems (vectors of n, system-major); skip: per-system stop flags (stride kStateInts).
using LinOp = std::function<void(const c128 *, c128 *, const int *)>;

struct SolveOptions {
    int method = RT_METHOD_CG;
    double tolerance = 1e-6;
    int max_iterations = 1000;
    int restart = 50;
    int precision = RT_C128;
    int refine = 0;
    double inner_tolerance = 0;
    int max_outer = 20;
    bool fp64_basis = false;   // a complex64 inner GMRES keeps its basis in complex128 (RT_C64_FP64_BASIS)
};

struct SolveResult {
    bool converged = false;
    int iterations = 0;
    int outer = 0;
    std::vector<double> norms;
    double residual = 0;
    double time = 0;
};

// Device state of one system of a Krylov loop, written by its finalize kernels and polled by the host.
struct LoopState {
    int done;         // every later kernel of this system returns at once (finished, or the iteration limit)
    int converged;
    int finished;     // converged or broke down: stays done across GMRES cycles
    int iterations;   // iterations completed
    int steps;        // GMRES: steps of the current cycle completed
    int best;         // COCG: iteration of the best monitored norm
    int arrivals;     // GMRES: blocks of the running step's block sums done so far (zero between launches)
    int live;         // GMRES: the iteration whose step runs for this system (set by its block sums)
    float monitored;  // GMRES: monitored residual of the last finalized iteration (the host's poll schedule)
};
constexpr int kStateInts = sizeof(LoopState) / sizeof(int);
constexpr int kScal = 512;   // device scalar slots per system
constexpr int kMaxBatch = 1024;   // systems of one batched solve

// Device scratch of one batched solve, from the context's buffer pool.
struct Workspace {
    Context *ctx;
    hipStream_t s;
    int64_t n;
    int R;
    int blocks;          // row blocks per system of every reduction
    int64_t chunk;       // rows per block (a multiple of 256)
    bool gs_tile_update; // GMRES: the Gram-Schmidt update pass runs on gs_blocks x gs_chunk (else blocks x chunk)
    int gs_blocks;       // its row blocks per system
    int64_t gs_chunk;    // their rows per block (a multiple of 64)
    int dot_blocks;      // row blocks per system of the GMRES Gram-Schmidt dot pass
    int dot_per;         // their 64-row tiles per block
    int dot_strided;     // 1: block b takes the tiles b, b + dot_blocks, ...; 0: a run of dot_per tiles
    c128 *partial;       // [R][blocks][128]
    c128 *scal;          // [R][kScal]
    LoopState *state;    // [R]
    double *history;     // [R][cap]
    int cap;
    unsigned char *host; // pinned: [0, read_bytes) takes device-to-host reads, the rest stages host-to-device writes
    size_t host_bytes;
    size_t read_bytes;
    size_t stage_off = 0;   // next free byte of the staging part (every copy from it before is still in flight)
    Workspace(Context *c, int64_t n_, int R_, int max_iter);
    const int *skip() const { return &state->done; }
    void nrm2(const c128 *x, int slot, const LoopState *st);
    void dot(const c128 *x, const c128 *y, bool conj, int slot, const LoopState *st);
    std::vector<double> fetch_re(int slot);                  // real part of one slot of every system
    // Real parts of several slots of every system and the loop states, read with one synchronization.
    std::vector<std::vector<double>> fetch_slots(const std::vector<int> &slots, std::vector<LoopState> *states);
    unsigned char *stage(size_t bytes);   // host bytes for one asynchronous host-to-device copy
    void put_re(int slot, const std::vector<double> &v);     // sets one slot of every system (real)
    void copy(const c128 *x, c128 *y, const LoopState *st);
    void sub(const c128 *b, const c128 *ax, c128 *r, const LoopState *st);
    void add(const c128 *x, c128 *y, const LoopState *st);   // y += x
    void zero(c128 *x);
    void reset_state(const std::vector<char> &finished);
    std::vector<LoopState> read_state();
    void set_scale(const std::vector<LoopState> &st);   // share of systems running, for the kernel timers
    std::vector<double> read_history(int r, int first, int count);
};

// Solves R systems A x_r = b_r (b, x0, x: R vectors of n, system-major). One result per system.
std::vector<SolveResult> solve_batch(Context *ctx, Product &op, Product *pc, const SolveOptions &o, const c128 *b,
                                     const c128 *x0, c128 *x, int R);

}  // namespace rt

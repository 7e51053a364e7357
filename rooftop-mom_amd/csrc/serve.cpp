// The serving loop (no counterpart in the reference, whose Solver.solve serves one caller at a time): requests of
// different grids, frequencies and masks are admitted by an HBM budget and solved by a pool of worker slots, each
// with its own HIP stream and device buffers, so the requests on the device run concurrently and a request that
// arrives while others run starts as soon as a slot and its bytes are free (first come, first served). A request
// keeps its generators, spectra, near field and Krylov state resident while it runs and returns all of it when it
// ends. The solve of a request is exactly the one of rt_solve on the same inputs: the kernels of one request read
// only its own buffers and reduce in a fixed order, so its result does not depend on what runs beside it.
//
// The budget follows the reference's Budget.require/hold/release (src/+rooftopMoM/+resources/Budget.m:59-106): a
// request's bytes are estimated before it starts (the byte models of estimateProduct.m:13-37, estimateSolver.m:19-26
// and estimateGenerators.m:11, re-derived for this library's device buffers), held while it runs and released at its
// end; a request larger than the whole budget is refused at submission with rooftopMoM:resources:insufficientMemory.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>

#include "krylov.hpp"
#include "rooftop.hpp"

namespace rt {

using Clock = std::chrono::steady_clock;

struct ServeJob {
    int64_t ticket = 0;
    int cx = 0, cy = 0;
    double a = 0, b = 0, k = 0;
    int near_order = 16, far_order = 8;
    std::vector<uint8_t> mask;
    std::vector<c128> rhs;
    int preconditioner = 0;
    SolveOptions opts;
    uint64_t bytes = 0;
    Clock::time_point submitted, started, finished;
    bool done = false;
    int status = RT_OK;
    std::string error_id, error_message;
    std::vector<c128> x;
    SolveResult result;
};

// Device bytes a request holds at its peak: generators, mask, spectra (master and the pass-B copies of both
// precisions), FFT tables, the products' working arrays, the near field, the Krylov vectors (a GMRES(m) basis or the
// COCG vectors) and the right-hand side and solution, with a margin for the one-off transform buffer of the spectra.
uint64_t estimate_request_bytes(int cx, int cy, const SolveOptions &o, int preconditioner) {
    const GridInfo g = GridInfo::make(cx, cy, 1.0, 1.0);
    const uint64_t n = static_cast<uint64_t>(g.n), cells = static_cast<uint64_t>(cx) * cy;
    const uint64_t Mx = cx > 1 ? next_pow2(2L * cx - 1) : 1, My = cy > 1 ? next_pow2(2L * cy - 1) : 1;
    const uint64_t Hx = Mx / 2 + 1, Hy = My / 2 + 1;
    const uint64_t ldj = (std::max(g.cols_x, g.cols_y) + 7) / 8 * 8;
    const bool mixed = o.precision == RT_C64;
    uint64_t b = 0;
    b += 16 * (n + 3 * cells);                                    // generators
    b += n;                                                       // mask
    b += 3 * Hx * Hy * 16;                                        // spectra (FP64 master)
    b += 3 * Hx * My * (16 + (mixed ? 8 : 0));                    // pass-B spectra
    b += 2 * ldj * Mx * (16 + (mixed ? 8 : 0));                   // working arrays
    b += (Mx + My) * 4 * 64;                                      // FFT tables
    if (preconditioner == RT_PRECOND_NEARFIELD) b += 8 * n + 4096 * 18 * 24;   // classes, stencil classes, tables
    if (preconditioner == RT_PRECOND_JACOBI) b += 24 * n;
    // GMRES: a basis of m + 1 columns (m = min(restart or the iteration limit, n, the iteration limit)), the pending
    // and work vectors, the Hessenberg matrices and the reduction partials; COCG: its vectors.
    const uint64_t maxit = static_cast<uint64_t>(std::max(o.max_iterations, 1));
    const uint64_t m = std::min<uint64_t>({o.restart > 0 ? static_cast<uint64_t>(o.restart) : maxit, n, maxit});
    uint64_t vectors = o.method == RT_METHOD_GMRES ? m + 1 + 6 : 8;
    if (o.refine) vectors += 3;
    b += (vectors + 2) * n * 16;                                  // Krylov vectors, right-hand side, solution
    if (o.method == RT_METHOD_GMRES) b += (2 * m * (m + 1) + 6 * (m + 1)) * 16 + (m + 1) * 2048 * 16;
    const uint64_t spectra_transform = Mx * My * 16;              // freed before the solve starts
    // The near-field classification's keys, sorts and scans (and the face mask), freed before the solve starts.
    const uint64_t classify = preconditioner == RT_PRECOND_NEARFIELD ? 48 * n + cells : 0;
    return b + std::max(spectra_transform, classify) + (b >> 3) + (16u << 20);
}

struct Server {
    int device = 0;
    uint64_t budget = 0;
    // One context per worker slot, opened by rt_server_create on the calling thread (a slot that cannot open its
    // context fails the creation instead of leaving queued requests without a worker).
    std::vector<std::unique_ptr<Context>> contexts;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<ServeJob>> queue;
    std::map<int64_t, std::shared_ptr<ServeJob>> jobs;
    uint64_t reserved = 0;
    int running = 0, peak_running = 0;
    int64_t next_ticket = 1, submitted = 0, completed = 0;
    bool stopping = false;
    std::vector<std::thread> workers;
    // Requests of one embedding size solved together (rt_server_set_batching): up to max_group per solve.
    int max_group = 1;
    int64_t groups = 0, grouped = 0;   // solves of more than one request, and the requests they held

    void run(int slot);
    void solve(Context &ctx, ServeJob &job);
    void solve_group(Context &ctx, const std::vector<std::shared_ptr<ServeJob>> &jobs);
};

// Whether two requests can share one batched solve: the same FFT periods (embedding size) of at most kBatchPeriod
// points, the same face aspect and quadrature orders, the near-field preconditioner or none on both, and equal solver
// options. Their grids, masks, wavenumbers, face sizes and right-hand sides may differ. Larger grids are solved alone:
// one of them already keeps the device's memory busy (its kernels are bandwidth-bound), so a group of them would
// only run one after another in one slot, padded to the largest, instead of side by side in several (measured on the
// served stream: four grids of 542 to 772 cells a side took 666 ms as a group and at most 389 ms alone).
constexpr long kBatchPeriod = 256;

static bool batchable(const ServeJob &a, const ServeJob &b) {
    auto period = [](int c) { return c > 1 ? next_pow2(2L * c - 1) : 1L; };
    const SolveOptions &p = a.opts, &q = b.opts;
    if (period(a.cx) > kBatchPeriod || period(a.cy) > kBatchPeriod) return false;
    return period(a.cx) == period(b.cx) && period(a.cy) == period(b.cy) && a.cx > 1 && a.cy > 1 && b.cx > 1 &&
           b.cy > 1 && a.b * b.a == b.b * a.a && a.near_order == b.near_order && a.far_order == b.far_order &&
           a.preconditioner == b.preconditioner && a.preconditioner != RT_PRECOND_JACOBI &&
           p.method == q.method && p.tolerance == q.tolerance && p.max_iterations == q.max_iterations &&
           p.restart == q.restart && p.precision == q.precision && p.refine == q.refine &&
           p.inner_tolerance == q.inner_tolerance && p.max_outer == q.max_outer && p.fp64_basis == q.fp64_basis;
}

// The device bytes of a group: every request as if on the group's padded grid, and the stacked copies of the spectra
// and near field beside the members' own while the batch product is built.
static uint64_t group_bytes(const std::vector<std::shared_ptr<ServeJob>> &jobs) {
    if (jobs.size() == 1) return jobs[0]->bytes;
    int cx = 0, cy = 0;
    for (const auto &j : jobs) {
        cx = std::max(cx, j->cx);
        cy = std::max(cy, j->cy);
    }
    const uint64_t one = estimate_request_bytes(cx, cy, jobs[0]->opts, jobs[0]->preconditioner);
    return (one + one / 2) * jobs.size();
}

void Server::solve(Context &ctx, ServeJob &job) {
    const SolveOptions &o = job.opts;
    auto p = make_problem(&ctx, job.cx, job.cy, job.a, job.b, job.k, job.near_order, job.far_order,
                          job.mask.empty() ? nullptr : job.mask.data());
    const int held = o.precision == RT_C64 ? RT_MIXED : RT_C128;
    auto op = create_operator(p.get(), held, RT_MODIFIER_NONE);
    std::unique_ptr<Product> pc;
    if (job.preconditioner) pc = create_preconditioner(p.get(), job.preconditioner, held);
    const int64_t n = p->g.n;
    DeviceArray<c128> b, x;
    b.allocate(&ctx, n, "the right-hand side");
    x.allocate(&ctx, n, "the solution");
    RT_HIP(hipMemcpyAsync(b.ptr, job.rhs.data(), n * sizeof(c128), hipMemcpyHostToDevice, ctx.stream));
    std::vector<SolveResult> r = solve_batch(&ctx, *op, pc.get(), o, b.ptr, nullptr, x.ptr, 1);
    job.x.resize(n);
    RT_HIP(hipMemcpyAsync(job.x.data(), x.ptr, n * sizeof(c128), hipMemcpyDeviceToHost, ctx.stream));
    RT_HIP(hipStreamSynchronize(ctx.stream));
    job.result = r[0];
}

void Server::solve_group(Context &ctx, const std::vector<std::shared_ptr<ServeJob>> &jobs) {
    // Every request on the common face grid (the largest counts of the group), its mask padded with empty faces, so
    // that one product serves them all; its right-hand side and solution are moved between the two rooftop layouts.
    const ServeJob &first = *jobs[0];
    const SolveOptions &o = first.opts;
    int cx = 0, cy = 0;
    for (const auto &j : jobs) {
        cx = std::max(cx, j->cx);
        cy = std::max(cy, j->cy);
    }
    const GridInfo pad = GridInfo::make(cx, cy, 1.0, 1.0);
    const int R = static_cast<int>(jobs.size());
    const int held = o.precision == RT_C64 ? RT_MIXED : RT_C128;
    std::vector<std::shared_ptr<Problem>> problems;
    for (const auto &j : jobs) {
        std::vector<uint8_t> mask(static_cast<size_t>(cx) * cy, 0);
        for (int y = 0; y < j->cy; ++y)
            for (int x = 0; x < j->cx; ++x)
                mask[x + static_cast<size_t>(cx) * y] = j->mask.empty() ? 1 : j->mask[x + static_cast<size_t>(j->cx) * y];
        problems.push_back(make_problem(&ctx, cx, cy, j->a, j->b, j->k, j->near_order, j->far_order, mask.data()));
    }
    StackedProducts batch = create_stacked(&ctx, problems, held, first.preconditioner);
    problems.clear();   // the batch holds copies of what the solve reads
    // Index of rooftop i of a request's own grid in the padded layout (x-directed first, x fastest).
    auto padded = [&](const ServeJob &j, int64_t i) -> int64_t {
        const GridInfo own = GridInfo::make(j.cx, j.cy, 1.0, 1.0);
        if (i < own.nx) return i % own.rows_x + static_cast<int64_t>(pad.rows_x) * (i / own.rows_x);
        const int64_t k = i - own.nx;
        return pad.nx + k % own.rows_y + static_cast<int64_t>(pad.rows_y) * (k / own.rows_y);
    };
    const int64_t n = pad.n;
    std::vector<c128> host(static_cast<size_t>(R) * n, mk(0.0, 0.0));
    for (int r = 0; r < R; ++r)
        for (int64_t i = 0; i < static_cast<int64_t>(jobs[r]->rhs.size()); ++i)
            host[r * n + padded(*jobs[r], i)] = jobs[r]->rhs[i];
    DeviceArray<c128> b, x;
    b.allocate(&ctx, host.size(), "the right-hand sides");
    x.allocate(&ctx, host.size(), "the solutions");
    RT_HIP(hipMemcpyAsync(b.ptr, host.data(), host.size() * sizeof(c128), hipMemcpyHostToDevice, ctx.stream));
    std::vector<SolveResult> res = solve_batch(&ctx, *batch.op, batch.pc.get(), o, b.ptr, nullptr, x.ptr, R);
    RT_HIP(hipMemcpyAsync(host.data(), x.ptr, host.size() * sizeof(c128), hipMemcpyDeviceToHost, ctx.stream));
    RT_HIP(hipStreamSynchronize(ctx.stream));
    for (int r = 0; r < R; ++r) {
        ServeJob &j = *jobs[r];
        j.x.resize(j.rhs.size());
        for (int64_t i = 0; i < static_cast<int64_t>(j.x.size()); ++i) j.x[i] = host[r * n + padded(j, i)];
        j.result = res[r];
    }
}

void Server::run(int slot) {
    Context &ctx = *contexts[slot];
    try {
        ctx.activate();   // the device of this worker thread
    } catch (const Error &) {
    }
    for (;;) {
        std::vector<std::shared_ptr<ServeJob>> group;
        uint64_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(m);
            cv_work.wait(lk, [&] {
                return stopping || (!queue.empty() && (reserved + queue.front()->bytes <= budget || running == 0));
            });
            if (stopping) break;
            group.push_back(queue.front());
            queue.pop_front();
            bytes = group_bytes(group);
            // Queued requests of the same embedding size join it while the group fits the budget (first come, first
            // served among them).
            for (auto it = queue.begin(); it != queue.end() && static_cast<int>(group.size()) < max_group;) {
                if (!batchable(*group[0], **it)) {
                    ++it;
                    continue;
                }
                group.push_back(*it);
                const uint64_t more = group_bytes(group);
                if (reserved + more > budget) {
                    group.pop_back();
                    break;
                }
                bytes = more;
                it = queue.erase(it);
            }
            reserved += bytes;
            running += static_cast<int>(group.size());
            peak_running = std::max(peak_running, running);
            if (group.size() > 1) {
                ++groups;
                grouped += static_cast<int64_t>(group.size());
            }
        }
        const Clock::time_point started = Clock::now();
        for (auto &job : group) job->started = started;
        int status = RT_OK;
        std::string error_id, error_message;
        try {
            if (group.size() == 1) solve(ctx, *group[0]);
            else solve_group(ctx, group);
        } catch (const Error &e) {
            status = e.status;
            error_id = e.id;
            error_message = e.what();
        } catch (const std::exception &e) {
            status = RT_EINVAL;
            error_id = "rooftop:server";
            error_message = e.what();
        }
        try {
            ctx.free_pool();
        } catch (const Error &) {
        }
        {
            std::lock_guard<std::mutex> lk(m);
            const Clock::time_point finished = Clock::now();
            for (auto &job : group) {
                job->finished = finished;
                job->status = status;
                job->error_id = error_id;
                job->error_message = error_message;
                job->done = true;
            }
            reserved -= bytes;
            running -= static_cast<int>(group.size());
            completed += static_cast<int64_t>(group.size());
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

}  // namespace rt

using namespace rt;

// The server state is shared with the callers blocked in rt_server_wait: rt_server_destroy drops the handle's
// reference after the workers have ended, and a waiter it woke still holds its own until it returns.
struct rt_server {
    std::shared_ptr<Server> s;
};

extern "C" {

int rt_blob_mask(int32_t cx, int32_t cy, uint64_t seed, double fill, uint8_t *mask) {
    try {
        if (cx < 1 || cy < 1 || !mask || !(fill > 0 && fill <= 1))
            fail(RT_EINVAL, "rooftop:mask", "positive counts, a mask buffer and a fill in (0, 1] are needed");
        const int64_t cells = static_cast<int64_t>(cx) * cy;
        std::fill(mask, mask + cells, 0);
        // splitmix64: a seeded, portable generator.
        uint64_t state = seed;
        auto next = [&]() {
            uint64_t z = (state += 0x9e3779b97f4a7c15ull);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            return z ^ (z >> 31);
        };
        auto uniform = [&](uint64_t bound) { return static_cast<int64_t>(next() % bound); };
        const int64_t target = std::max<int64_t>(1, static_cast<int64_t>(fill * static_cast<double>(cells) + 0.5));
        std::vector<int64_t> frontier;
        std::vector<uint8_t> queued(cells, 0);
        auto push = [&](int64_t c) {
            if (!mask[c] && !queued[c]) {
                queued[c] = 1;
                frontier.push_back(c);
            }
        };
        frontier.push_back(uniform(cells));
        queued[frontier[0]] = 1;
        for (int64_t grown = 0; grown < target && !frontier.empty(); ++grown) {
            const int64_t pick = uniform(frontier.size());
            const int64_t c = frontier[pick];
            frontier[pick] = frontier.back();
            frontier.pop_back();
            mask[c] = 1;
            const int x = static_cast<int>(c % cx), y = static_cast<int>(c / cx);
            if (x > 0) push(c - 1);
            if (x + 1 < cx) push(c + 1);
            if (y > 0) push(c - cx);
            if (y + 1 < cy) push(c + cx);
        }
        set_last_error(RT_OK, "", "");
        return RT_OK;
    } catch (const Error &e) {
        set_last_error(e.status, e.id, e.what());
        return e.status;
    }
}

int rt_server_create(int32_t device, uint64_t budget, int32_t slots, rt_server **out) {
    try {
        if (!out || slots < 1 || slots > 64) fail(RT_EINVAL, "rooftop:server", "1..64 slots and an output are needed");
        int count = 0;
        RT_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) fail(RT_EINVAL, "rooftop:server", "no such device");
        RT_HIP(hipSetDevice(device));
        size_t free_b = 0, total_b = 0;
        RT_HIP(hipMemGetInfo(&free_b, &total_b));
        auto h = std::make_unique<rt_server>();
        h->s = std::make_shared<Server>();
        Server &srv = *h->s;
        srv.device = device;
        srv.budget = budget ? budget : static_cast<uint64_t>(free_b) / 10 * 9;
        try {
            for (int i = 0; i < slots; ++i) {
                srv.contexts.push_back(std::make_unique<Context>());
                srv.contexts.back()->open(device, srv.budget);
            }
        } catch (...) {
            for (auto &c : srv.contexts) c->close();
            throw;
        }
        for (int i = 0; i < slots; ++i) srv.workers.emplace_back([p = &srv, i] { p->run(i); });
        *out = h.release();
        set_last_error(RT_OK, "", "");
        return RT_OK;
    } catch (const Error &e) {
        set_last_error(e.status, e.id, e.what());
        return e.status;
    }
}

void rt_server_destroy(rt_server *h) {
    if (!h) return;
    Server &srv = *h->s;
    {
        std::lock_guard<std::mutex> lk(srv.m);
        srv.stopping = true;
        for (auto &job : srv.queue) {   // never started: their waiters get an error
            job->done = true;
            job->status = RT_EINVAL;
            job->error_id = "rooftop:server";
            job->error_message = "the server was destroyed before the request started";
        }
        srv.queue.clear();
    }
    srv.cv_work.notify_all();
    srv.cv_done.notify_all();
    for (auto &t : srv.workers) t.join();
    for (auto &c : srv.contexts) c->close();
    delete h;   // a waiter still returning keeps the server state alive through its own reference
}

int rt_server_set_batching(rt_server *h, int32_t max_requests) {
    if (!h || max_requests < 1 || max_requests > 64) {
        set_last_error(RT_EINVAL, "rooftop:server", "1..64 requests per batched solve");
        return RT_EINVAL;
    }
    std::lock_guard<std::mutex> lk(h->s->m);
    h->s->max_group = max_requests;
    set_last_error(RT_OK, "", "");
    return RT_OK;
}

int rt_server_batch_stats(rt_server *h, int64_t *groups, int64_t *grouped) {
    if (!h) return RT_EINVAL;
    std::lock_guard<std::mutex> lk(h->s->m);
    if (groups) *groups = h->s->groups;
    if (grouped) *grouped = h->s->grouped;
    return RT_OK;
}

int rt_server_submit(rt_server *h, const rt_request *r, int64_t *ticket) {
    try {
        if (!h || !r || !ticket || !r->rhs) fail(RT_EINVAL, "rooftop:server", "NULL argument");
        const GridInfo g = GridInfo::make(r->count_x, r->count_y, 1.0, 1.0);
        if (r->count_x < 1 || r->count_y < 1 || g.n < 1)
            fail(RT_EINVAL, "rooftopMoM:mesh:noRooftops", "The grid carries no rooftop.");
        if (r->n != g.n)
            fail(RT_ESIZE, "rooftopMoM:solver:sizeMismatch", "The right-hand side does not match the grid's rooftops.");
        if (r->preconditioner != 0 && r->preconditioner != RT_PRECOND_JACOBI && r->preconditioner != RT_PRECOND_NEARFIELD)
            fail(RT_EINVAL, "MATLAB:validators:mustBeMember", "The preconditioner is none, jacobi or nearField.");
        auto job = std::make_shared<ServeJob>();
        job->cx = r->count_x;
        job->cy = r->count_y;
        job->a = r->size_x;
        job->b = r->size_y;
        job->k = r->wavenumber;
        job->near_order = r->near_order > 0 ? r->near_order : 16;
        job->far_order = r->far_order > 0 ? r->far_order : 8;
        if (r->mask) job->mask.assign(r->mask, r->mask + static_cast<size_t>(r->count_x) * r->count_y);
        const c128 *v = static_cast<const c128 *>(r->rhs);
        job->rhs.assign(v, v + g.n);
        SolveOptions o;
        o.method = r->opts.method;
        o.tolerance = r->opts.tolerance;
        o.max_iterations = r->opts.max_iterations;
        o.restart = r->opts.restart;
        o.precision = r->opts.precision;
        o.refine = r->opts.refine;
        o.inner_tolerance = r->opts.inner_tolerance;
        o.max_outer = r->opts.max_outer > 0 ? r->opts.max_outer : 20;
        if (o.precision == RT_C64_FP64_BASIS) {
            o.precision = RT_C64;
            o.fp64_basis = true;
        }
        if (o.method != RT_METHOD_CG && o.method != RT_METHOD_GMRES)
            fail(RT_EINVAL, "MATLAB:validators:mustBeMember", "The method is 'cg' or 'gmres'.");
        if (!(o.tolerance > 0 && o.tolerance < 1) || o.max_iterations < 1)
            fail(RT_EINVAL, "MATLAB:validators:mustBePositive", "Invalid tolerance or iteration limit.");
        if (o.precision != RT_C128 && o.precision != RT_C64)
            fail(RT_EINVAL, "rooftop:solver:invalidPrecision", "The inner precision is RT_C128 or RT_C64.");
        if (o.precision == RT_C64 && !o.refine)
            fail(RT_EINVAL, "rooftop:solver:invalidPrecision",
                 "A complex64 inner loop needs the FP64 refinement loop (refine = 1).");
        job->opts = o;
        job->preconditioner = r->preconditioner;
        job->bytes = estimate_request_bytes(job->cx, job->cy, o, job->preconditioner);
        Server &srv = *h->s;
        {
            std::lock_guard<std::mutex> lk(srv.m);
            if (job->bytes > srv.budget)
                fail(RT_ENOMEM, "rooftopMoM:resources:insufficientMemory",
                     "The request needs about %llu bytes, more than the server's budget of %llu.",
                     static_cast<unsigned long long>(job->bytes), static_cast<unsigned long long>(srv.budget));
            job->ticket = srv.next_ticket++;
            job->submitted = Clock::now();
            srv.jobs[job->ticket] = job;
            srv.queue.push_back(job);
            ++srv.submitted;
            *ticket = job->ticket;
        }
        srv.cv_work.notify_all();
        set_last_error(RT_OK, "", "");
        return RT_OK;
    } catch (const Error &e) {
        set_last_error(e.status, e.id, e.what());
        return e.status;
    }
}

int rt_request_bytes(const rt_request *r, uint64_t *bytes) {
    if (!r || !bytes || r->count_x < 1 || r->count_y < 1) return RT_EINVAL;
    SolveOptions o;
    o.method = r->opts.method;
    o.restart = r->opts.restart;
    o.max_iterations = r->opts.max_iterations > 0 ? r->opts.max_iterations : 1000;
    o.precision = r->opts.precision == RT_C64_FP64_BASIS ? RT_C64 : r->opts.precision;
    o.refine = r->opts.refine;
    *bytes = estimate_request_bytes(r->count_x, r->count_y, o, r->preconditioner);
    return RT_OK;
}

int rt_server_wait(rt_server *h, int64_t ticket, void *x, int64_t n, rt_output *out, double *times) {
    try {
        if (!h) fail(RT_EINVAL, "rooftop:server", "NULL server");
        std::shared_ptr<ServeJob> job;
        const std::shared_ptr<Server> srv = h->s;   // keeps the state alive while this call waits
        {
            std::unique_lock<std::mutex> lk(srv->m);
            auto it = srv->jobs.find(ticket);
            if (it == srv->jobs.end()) fail(RT_EINVAL, "rooftop:server", "unknown ticket %lld", (long long)ticket);
            job = it->second;
            srv->cv_done.wait(lk, [&] { return job->done; });
            srv->jobs.erase(ticket);
        }
        if (job->status != RT_OK) fail(job->status, job->error_id.c_str(), "%s", job->error_message.c_str());
        if (x) {
            if (n != static_cast<int64_t>(job->x.size()))
                fail(RT_ESIZE, "rooftopMoM:solver:sizeMismatch", "The solution buffer does not match the request.");
            std::memcpy(x, job->x.data(), job->x.size() * sizeof(c128));
        }
        if (out) {
            const SolveResult &r = job->result;
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
        if (times) {
            auto sec = [](Clock::duration d) { return std::chrono::duration<double>(d).count(); };
            times[0] = sec(job->started - job->submitted);
            times[1] = sec(job->finished - job->started);
            times[2] = sec(job->finished - job->submitted);
        }
        set_last_error(RT_OK, "", "");
        return RT_OK;
    } catch (const Error &e) {
        set_last_error(e.status, e.id, e.what());
        return e.status;
    }
}

int rt_server_stats(rt_server *h, int64_t *counts, uint64_t *bytes) {
    if (!h) return RT_EINVAL;
    Server &srv = *h->s;
    std::lock_guard<std::mutex> lk(srv.m);
    if (counts) {
        counts[0] = srv.submitted;
        counts[1] = srv.completed;
        counts[2] = static_cast<int64_t>(srv.queue.size());
        counts[3] = srv.running;
        counts[4] = srv.peak_running;
    }
    if (bytes) {
        bytes[0] = srv.reserved;
        bytes[1] = srv.budget;
    }
    return RT_OK;
}

}  // extern "C"

// Stencil classes of the near-field preconditioner on the device. A rooftop's sparse-approximate-inverse row depends
// only on which rooftops exist in the 5 x 5 first-face window of both directions around it (its class), and its
// symmetrized stencil on its class, its direction and the classes of its 18 first-face neighbours, which depend only
// on the 7 x 7 window (its stencil class). Both windows are functions of the face mask: a rooftop exists where both of
// its faces do, so the 7 x 7 window of either direction is fixed by the faces of the 8 x 8 face window that are
// endpoints of an existing rooftop in it (the corner face (+4, +4) is an endpoint of none). Every rooftop forms both
// keys (64 bits each), a radix sort groups equal keys, and the classes are numbered in key order with the lowest
// rooftop of each as its representative. The few distinct stencil tuples of the 7 x 7 windows' classes are then made
// unique on the host. Classes and stencil classes are those of a first-occurrence numbering of the rooftops' keys and
// tuples, in another order, so every rooftop gets the same coefficients.
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "nearfield_classes.hpp"

namespace rt {

namespace {

constexpr int kThreads = 256;

// Face window of rooftop i: bit (a + 3) + 8 (b + 3) for the face (fx + a, fy + b), a, b in [-3, 4], where (fx, fy) is
// the rooftop's first face; faces off the grid (or off the mask) are absent.
__device__ uint64_t face_window(const GridInfo g, const uint8_t *__restrict__ faces, int fx, int fy) {
    uint64_t w = 0;
#pragma unroll
    for (int b = -3; b <= 4; ++b) {
        const int y = fy + b;
#pragma unroll
        for (int a = -3; a <= 4; ++a) {
            const int x = fx + a;
            const bool in = x >= 0 && y >= 0 && x < g.cx && y < g.cy;
            const bool on = in && (!faces || faces[x + static_cast<int64_t>(g.cx) * y] != 0);
            w |= static_cast<uint64_t>(on) << ((a + 3) + 8 * (b + 3));
        }
    }
    return w;
}

// Keys of every rooftop: key5, the 5 x 5 window's rooftops of both directions (bit dd 25 + (dx + 2) + 5 (dy + 2)) with
// the direction at bit 60, as the host numbering formed it; key7, the faces of the window that are endpoints of an
// existing rooftop of the 7 x 7 window, with the direction at bit 63. 0 for a rooftop that does not exist (a real
// key holds the rooftop's own bit and faces).
__global__ void nf_keys(const GridInfo g, const uint8_t *__restrict__ faces, uint64_t *__restrict__ key5,
                        uint64_t *__restrict__ key7, int *__restrict__ index) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= g.n) return;
    const int d = i < g.nx ? 0 : 1;
    const int64_t k = d ? i - g.nx : i;
    const int rows = d ? g.rows_y : g.rows_x;
    const int fx = static_cast<int>(k % rows), fy = static_cast<int>(k / rows);
    const uint64_t F = face_window(g, faces, fx, fy);
    // Rooftops whose first face is at window offset (a, b), a, b in [-3, 3]: x-directed where faces (a, b) and
    // (a + 1, b) are present, y-directed where (a, b) and (a, b + 1) are.
    constexpr uint64_t kCols7 = 0x7f7f7f7f7f7f7full;   // a in [-3, 3], b in [-3, 3]
    const uint64_t rx = F & (F >> 1) & kCols7, ry = F & (F >> 8) & kCols7;
    const int own = 3 + 8 * 3;
    const bool exists = ((d ? ry : rx) >> own) & 1;
    index[i] = static_cast<int>(i);
    if (!exists) {
        key5[i] = 0;
        key7[i] = 0;
        return;
    }
    const uint64_t visible = rx | (rx << 1) | ry | (ry << 8);
    key7[i] = visible | (static_cast<uint64_t>(d) << 63);
    uint64_t bits = 0;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy)
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const int at = (dx + 3) + 8 * (dy + 3);
            bits |= ((rx >> at) & 1) << ((dx + 2) + 5 * (dy + 2));
            bits |= ((ry >> at) & 1) << (25 + (dx + 2) + 5 * (dy + 2));
        }
    key5[i] = bits | (static_cast<uint64_t>(d) << 60);
}

// flag[k] = 1 where sorted key k starts a run of a real key.
__global__ void nf_run_starts(const uint64_t *__restrict__ key, int64_t n, int *__restrict__ flag) {
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    flag[k] = key[k] != 0 && (k == 0 || key[k] != key[k - 1]) ? 1 : 0;
}

// id[index[k]] = the run number of sorted key k (-1 for the absent rooftops); rep[run] = the lowest rooftop of each
// run (the sort is stable, so a run's first element), and its key.
__global__ void nf_number(const uint64_t *__restrict__ key, const int *__restrict__ index,
                          const int *__restrict__ runs, int64_t n, int *__restrict__ id, int *__restrict__ rep,
                          uint64_t *__restrict__ rep_key) {
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    const bool real = key[k] != 0;
    const int run = runs[k] - 1;
    id[index[k]] = real ? run : -1;
    if (real && (k == 0 || key[k] != key[k - 1])) {
        rep[run] = index[k];
        if (rep_key) rep_key[run] = key[k];
    }
}

// Class representatives for the fit: direction and the 50 window bits.
__global__ void nf_class_reps(const uint64_t *__restrict__ rep_key, int classes, int *__restrict__ dir,
                              unsigned long long *__restrict__ bits) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= classes) return;
    dir[c] = static_cast<int>((rep_key[c] >> 60) & 1);
    bits[c] = rep_key[c] & ((1ull << 50) - 1);
}

// The tuple of each stencil class from its representative rooftop: (class, direction, the classes of the 18
// neighbours (direction dd, first-face offsets dx, dy in [-1, 1] at slot dd 9 + (dx + 1) + 3 (dy + 1)), -1 where
// there is none).
__global__ void nf_tuples(const GridInfo g, const int *__restrict__ cls, const int *__restrict__ rep, int stencils,
                          int *__restrict__ tuples) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= stencils) return;
    const int64_t i = rep[s];
    const int d = i < g.nx ? 0 : 1;
    const int64_t k = d ? i - g.nx : i;
    const int rows_i = d ? g.rows_y : g.rows_x;
    const int fx = static_cast<int>(k % rows_i), fy = static_cast<int>(k / rows_i);
    int *t = tuples + static_cast<int64_t>(s) * (kNearFieldSlots + 2);
    t[0] = cls[i];
    t[1] = d;
    for (int dd = 0; dd < 2; ++dd) {
        const int rows = dd ? g.rows_y : g.rows_x, cols = dd ? g.cols_y : g.cols_x;
        const bool present = dd ? g.has_y : g.has_x;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int gx = fx + dx, gy = fy + dy;
                int cj = -1;
                if (present && gx >= 0 && gy >= 0 && gx < rows && gy < cols)
                    cj = cls[(dd ? g.nx : 0) + gx + static_cast<int64_t>(rows) * gy];
                t[2 + dd * 9 + (dx + 1) + 3 * (dy + 1)] = cj;
            }
    }
}

// scls[i] = map[wclass[i]] (-1 stays -1).
__global__ void nf_remap(const int *__restrict__ wclass, const int *__restrict__ map, int64_t n, int *__restrict__ scls) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const int w = wclass[i];
    scls[i] = w < 0 ? -1 : map[w];
}

// Sorts the keys, numbers their runs into id (per rooftop) and rep (per run, its lowest rooftop) and returns the
// number of runs; rep_key (optional) receives each run's key.
int number_runs(Context *ctx, hipStream_t s, const uint64_t *key, const int *index, int64_t n, DeviceArray<int> &id,
                DeviceArray<int> &rep, DeviceArray<uint64_t> *rep_key) {
    DeviceArray<uint64_t> skey;
    DeviceArray<int> sindex, runs;
    skey.allocate(ctx, n, "near-field classification");
    sindex.allocate(ctx, n, "near-field classification");
    size_t bytes = 0;
    RT_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key, skey.ptr, index, sindex.ptr, static_cast<size_t>(n), 0, 64,
                                     s));
    {
        DeviceArray<unsigned char> tmp;
        tmp.allocate(ctx, bytes, "near-field classification");
        RT_HIP(rocprim::radix_sort_pairs(tmp.ptr, bytes, key, skey.ptr, index, sindex.ptr, static_cast<size_t>(n), 0,
                                         64, s));
    }
    DeviceArray<int> starts;
    starts.allocate(ctx, n, "near-field classification");
    runs.allocate(ctx, n, "near-field classification");
    const unsigned nb = blocks_for(n, kThreads);
    hipLaunchKernelGGL(nf_run_starts, dim3(nb), dim3(kThreads), 0, s, skey.ptr, n, starts.ptr);
    RT_LAUNCH_CHECK();
    bytes = 0;
    RT_HIP(rocprim::inclusive_scan(nullptr, bytes, starts.ptr, runs.ptr, static_cast<size_t>(n),
                                   rocprim::plus<int>(), s));
    {
        DeviceArray<unsigned char> tmp;
        tmp.allocate(ctx, bytes, "near-field classification");
        RT_HIP(rocprim::inclusive_scan(tmp.ptr, bytes, starts.ptr, runs.ptr, static_cast<size_t>(n),
                                       rocprim::plus<int>(), s));
    }
    starts.reset();
    int count = 0;
    RT_HIP(hipMemcpyAsync(&count, runs.ptr + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    rep.allocate(ctx, std::max(count, 1), "near-field classes");
    if (rep_key) rep_key->allocate(ctx, std::max(count, 1), "near-field classes");
    hipLaunchKernelGGL(nf_number, dim3(nb), dim3(kThreads), 0, s, skey.ptr, sindex.ptr, runs.ptr, n, id.ptr, rep.ptr,
                       rep_key ? rep_key->ptr : nullptr);
    RT_LAUNCH_CHECK();
    return count;
}

}  // namespace

NearFieldClasses classify_near_field(const Problem &p, hipStream_t s) {
    Context *ctx = p.ctx;
    const GridInfo &g = p.g;
    NearFieldClasses out;
    out.cls.allocate(ctx, g.n, "The near-field classes");
    out.scls.allocate(ctx, g.n, "The near-field stencil classes");
    if (g.n == 0) return out;
    DeviceArray<uint8_t> faces;
    if (!p.face_mask.empty()) {
        faces.allocate(ctx, p.face_mask.size(), "near-field classification");
        RT_HIP(hipMemcpyAsync(faces.ptr, p.face_mask.data(), p.face_mask.size(), hipMemcpyHostToDevice, s));
    }
    DeviceArray<uint64_t> key5, key7;
    DeviceArray<int> index;
    key5.allocate(ctx, g.n, "near-field classification");
    key7.allocate(ctx, g.n, "near-field classification");
    index.allocate(ctx, g.n, "near-field classification");
    hipLaunchKernelGGL(nf_keys, dim3(blocks_for(g.n, kThreads)), dim3(kThreads), 0, s, g, faces.ptr, key5.ptr,
                       key7.ptr, index.ptr);
    RT_LAUNCH_CHECK();
    faces.reset();
    {   // classes: the fit's representatives
        DeviceArray<int> rep;
        DeviceArray<uint64_t> rep_key;
        out.classes = number_runs(ctx, s, key5.ptr, index.ptr, g.n, out.cls, rep, &rep_key);
        key5.reset();
        out.rep_dir.allocate(ctx, std::max(out.classes, 1), "near-field fit");
        out.rep_bits.allocate(ctx, std::max(out.classes, 1), "near-field fit");
        if (out.classes > 0) {
            hipLaunchKernelGGL(nf_class_reps, dim3(blocks_for(out.classes, kThreads)), dim3(kThreads), 0, s,
                               rep_key.ptr, out.classes, out.rep_dir.ptr, out.rep_bits.ptr);
            RT_LAUNCH_CHECK();
        }
    }
    {   // stencil classes: the tuples of the 7 x 7 windows' classes, made unique on the host (two windows that differ
        // only around absent neighbours have the same tuple), numbered in the order of their first window class
        DeviceArray<int> rep, tuples7, wclass;
        wclass.allocate(ctx, g.n, "near-field classification");
        const int windows = number_runs(ctx, s, key7.ptr, index.ptr, g.n, wclass, rep, nullptr);
        key7.reset();
        index.reset();
        constexpr int W = kNearFieldSlots + 2;
        tuples7.allocate(ctx, static_cast<size_t>(std::max(windows, 1)) * W, "near-field classification");
        std::vector<int> host(static_cast<size_t>(windows) * W), map(std::max(windows, 1));
        if (windows > 0) {
            hipLaunchKernelGGL(nf_tuples, dim3(blocks_for(windows, kThreads)), dim3(kThreads), 0, s, g, out.cls.ptr,
                               rep.ptr, windows, tuples7.ptr);
            RT_LAUNCH_CHECK();
            RT_HIP(hipMemcpyAsync(host.data(), tuples7.ptr, host.size() * sizeof(int), hipMemcpyDeviceToHost, s));
            RT_HIP(hipStreamSynchronize(s));
        }
        std::vector<int> unique;
        std::unordered_map<std::string, int> seen;
        for (int w = 0; w < windows; ++w) {
            const std::string key(reinterpret_cast<const char *>(host.data() + static_cast<size_t>(w) * W),
                                  W * sizeof(int));
            auto it = seen.find(key);
            if (it == seen.end()) {
                it = seen.emplace(key, static_cast<int>(unique.size() / W)).first;
                unique.insert(unique.end(), host.begin() + static_cast<int64_t>(w) * W,
                              host.begin() + static_cast<int64_t>(w + 1) * W);
            }
            map[w] = it->second;
        }
        out.stencils = static_cast<int>(unique.size() / W);
        out.tuples.allocate(ctx, std::max<size_t>(unique.size(), W), "near-field stencil tuples");
        DeviceArray<int> dmap;
        dmap.allocate(ctx, map.size(), "near-field classification");
        if (out.stencils > 0)
            RT_HIP(hipMemcpyAsync(out.tuples.ptr, unique.data(), unique.size() * sizeof(int), hipMemcpyHostToDevice,
                                  s));
        RT_HIP(hipMemcpyAsync(dmap.ptr, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(nf_remap, dim3(blocks_for(g.n, kThreads)), dim3(kThreads), 0, s, wclass.ptr, dmap.ptr, g.n,
                           out.scls.ptr);
        RT_LAUNCH_CHECK();
        RT_HIP(hipStreamSynchronize(s));   // the host vectors are copied before they go out of scope
    }
    return out;
}

}  // namespace rt

// Batched right-hand sides and observables on the device: the plane wave of generatePlaneWave.m:18-42 for many
// incidences at once, written straight into a device batch, and the radar cross section
// (computeRadarCrossSection.m:18-19) and the input impedance (computeInputImpedance.m:17) of many solutions that are
// already in HBM -- the N x R host transfers of a batch stay off the path.
//
// The plane wave of a rooftop is the reference rooftop's quadrature sum times the translation phase of its cell,
// exp(-i k (dx a ix + dy b iy)) (the O(N) form of the reference): the sums, O(order^2) per incidence, are formed on the
// host in FP64 and the phases on the device, one FP64 sincos per rooftop.
#include <algorithm>
#include <cmath>
#include <complex>

#include "rooftop.hpp"

namespace rt {

namespace {

const double kZ0 = 376.730313668;   // +assembler/Constants.m:11

struct Wave {
    double kx, ky;   // phase per cell along x and y: -k dx a, -k dy b
    c128 ref_x, ref_y;
};

__global__ void k_plane_waves(const Wave *__restrict__ waves, int64_t n, int64_t nx, int rows_x, int rows_y,
                              const uint8_t *__restrict__ exists, c128 *__restrict__ b) {
    const int r = blockIdx.y;
    const Wave w = waves[r];
    c128 *br = b + static_cast<int64_t>(r) * n;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int64_t ix, iy;
        c128 ref;
        if (i < nx) {
            iy = i / rows_x;
            ix = i - iy * rows_x;
            ref = w.ref_x;
        } else {
            const int64_t j = i - nx;
            iy = j / rows_y;
            ix = j - iy * rows_y;
            ref = w.ref_y;
        }
        double s, c;
        sincos(w.kx * static_cast<double>(ix) + w.ky * static_cast<double>(iy), &s, &c);
        br[i] = (exists && !exists[i]) ? mk(0.0, 0.0) : ref * mk(c, s);
    }
}

// partial[r * gridDim.x + block] = sum over the block's rows of b_r(i) x_r(i) (unconjugated).
__global__ void __launch_bounds__(256) k_tdot(const c128 *__restrict__ b, const c128 *__restrict__ x, int64_t n,
                                              c128 *__restrict__ partial) {
    __shared__ double sh[2][4];
    const int r = blockIdx.y;
    const c128 *br = b + static_cast<int64_t>(r) * n, *xr = x + static_cast<int64_t>(r) * n;
    double re = 0, im = 0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const c128 p = br[i] * xr[i];
        re += p.x;
        im += p.y;
    }
    for (int o = 32; o > 0; o >>= 1) {
        re += __shfl_down(re, o, 64);
        im += __shfl_down(im, o, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][wave] = re;
        sh[1][wave] = im;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        partial[static_cast<int64_t>(r) * gridDim.x + blockIdx.x] =
            mk((sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]), (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]));
}

// out[r * count + p] = x_r(port[p]).
__global__ void k_gather(const c128 *__restrict__ x, int64_t n, const int64_t *__restrict__ port, int count,
                         c128 *__restrict__ out) {
    const int r = blockIdx.y;
    for (int p = threadIdx.x; p < count; p += blockDim.x) out[static_cast<int64_t>(r) * count + p] = x[r * n + port[p]];
}

void unit(const double *v, double *u, const char *name) {
    const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (!(len > 0) || !std::isfinite(len))
        fail(RT_EINVAL, "rooftopMoM:excitation:zeroVector", "The %s must not be zero.", name);
    for (int c = 0; c < 3; ++c) u[c] = v[c] / len;
}

}  // namespace

// The plane waves of `count` incidences (directions and polarizations [count][3], intensities [count] or null for
// 1) into the device batch b (n x count, column per incidence).
void plane_waves_device(Problem &p, int64_t count, const double *directions, const double *polarizations,
                        const double *intensities, c128 *b, hipStream_t s) {
    using cd = std::complex<double>;
    const GridInfo &g = p.g;
    std::vector<double> t, wt;
    gauss_legendre(p.far_order, t, wt);
    const int q = p.far_order;
    std::vector<Wave> waves(count);
    for (int64_t r = 0; r < count; ++r) {
        double d[3], e[3];
        unit(directions + 3 * r, d, "direction");
        unit(polarizations + 3 * r, e, "polarization");
        if (std::fabs(d[0] * e[0] + d[1] * e[1] + d[2] * e[2]) > 10 * 2.220446049250313e-16)
            fail(RT_EINVAL, "rooftopMoM:excitation:notPerpendicular",
                 "The polarization must be perpendicular to the direction of propagation.");
        const double amp = intensities ? intensities[r] : 1.0;
        const cd s0(0.0, -p.k * d[0] * g.a), s1(0.0, -p.k * d[1] * g.b);
        const cd e0 = std::exp(s0), e1 = std::exp(s1);
        cd sum_x = 0, sum_y = 0;
        for (int j = 0; j < q; ++j)       // v
            for (int i = 0; i < q; ++i) {  // u, fastest
                const cd phase = std::exp(s0 * t[i] + s1 * t[j]);
                const double w2 = wt[i] * wt[j];
                sum_x += phase * (w2 * (t[i] + e0 * (1.0 - t[i])));
                sum_y += phase * (w2 * (t[j] + e1 * (1.0 - t[j])));
            }
        const double area = g.a * g.b;
        const cd rx = amp * e[0] * area * sum_x, ry = amp * e[1] * area * sum_y;
        waves[r] = Wave{s0.imag(), s1.imag(), mk(rx.real(), rx.imag()), mk(ry.real(), ry.imag())};
    }
    DeviceArray<Wave> dw;
    dw.allocate(p.ctx, waves.size(), "plane-wave parameters");
    RT_HIP(hipMemcpyAsync(dw.ptr, waves.data(), sizeof(Wave) * waves.size(), hipMemcpyHostToDevice, s));
    for (int64_t r0 = 0; r0 < count; r0 += 65535) {
        const int64_t cnt = std::min<int64_t>(65535, count - r0);
        hipLaunchKernelGGL(k_plane_waves, dim3(blocks_for(g.n, 256, 4096), cnt), dim3(256), 0, s, dw.ptr + r0, g.n,
                           g.nx, g.rows_x, g.rows_y, p.exists.ptr, b + r0 * g.n);
        RT_LAUNCH_CHECK();
    }
    RT_HIP(hipStreamSynchronize(s));
}

// sigma[r] = (k Z0)^2 |b_r.' x_r|^2 / (4 pi intensity_r^4) for the device batches b (the plane waves) and x (their
// solutions), n x count each; the block sums are added in a fixed order on the host.
void radar_cross_sections_device(Problem &p, int64_t count, const c128 *b, const c128 *x, const double *intensities,
                                 double *sigma, hipStream_t s) {
    const int64_t n = p.g.n;
    const unsigned blocks = blocks_for(n, 256, 256);
    DeviceArray<c128> part;
    part.allocate(p.ctx, static_cast<size_t>(blocks) * count, "observable partial sums");
    for (int64_t r0 = 0; r0 < count; r0 += 65535) {
        const int64_t cnt = std::min<int64_t>(65535, count - r0);
        hipLaunchKernelGGL(k_tdot, dim3(blocks, cnt), dim3(256), 0, s, b + r0 * n, x + r0 * n, n,
                           part.ptr + r0 * blocks);
        RT_LAUNCH_CHECK();
    }
    std::vector<c128> h(static_cast<size_t>(blocks) * count);
    RT_HIP(hipMemcpyAsync(h.data(), part.ptr, sizeof(c128) * h.size(), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    const double kz = p.k * kZ0;
    for (int64_t r = 0; r < count; ++r) {
        double re = 0, im = 0;
        for (unsigned k = 0; k < blocks; ++k) {
            re += h[r * blocks + k].x;
            im += h[r * blocks + k].y;
        }
        const double e = intensities ? intensities[r] : 1.0;
        sigma[r] = kz * kz * (re * re + im * im) / (4 * 3.141592653589793 * e * e * e * e);
    }
}

// Z_in[r] = V_r / (edge lengths . x_r(port)) (real, imaginary) for the device batch x (n x count); port holds
// distinct 0-based rooftop indices, voltages [count] or null for 1 V.
void input_impedances_device(Problem &p, int64_t count, const int64_t *port, int64_t port_len, const double *voltages,
                             const c128 *x, double *zin, hipStream_t s) {
    const GridInfo &g = p.g;
    if (port_len <= 0 || port_len > (1 << 20))
        fail(RT_EINVAL, "rooftopMoM:excitation:invalidPort", "The port must hold 1 to 2^20 rooftop indices.");
    std::vector<int64_t> sorted(port, port + port_len);
    std::sort(sorted.begin(), sorted.end());
    if (sorted.front() < 0 || sorted.back() >= g.n || std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        fail(RT_EINVAL, "rooftopMoM:excitation:invalidPort", "The port must hold distinct rooftop indices up to %lld.",
             static_cast<long long>(g.n));
    DeviceArray<int64_t> dp;
    dp.allocate(p.ctx, port_len, "port indices");
    RT_HIP(hipMemcpyAsync(dp.ptr, port, sizeof(int64_t) * port_len, hipMemcpyHostToDevice, s));
    DeviceArray<c128> got;
    got.allocate(p.ctx, static_cast<size_t>(port_len) * count, "port currents");
    for (int64_t r0 = 0; r0 < count; r0 += 65535) {
        const int64_t cnt = std::min<int64_t>(65535, count - r0);
        hipLaunchKernelGGL(k_gather, dim3(1, cnt), dim3(256), 0, s, x + r0 * g.n, g.n, dp.ptr,
                           static_cast<int>(port_len), got.ptr + r0 * port_len);
        RT_LAUNCH_CHECK();
    }
    std::vector<c128> h(static_cast<size_t>(port_len) * count);
    RT_HIP(hipMemcpyAsync(h.data(), got.ptr, sizeof(c128) * h.size(), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    for (int64_t r = 0; r < count; ++r) {
        std::complex<double> sum = 0;
        for (int64_t q = 0; q < port_len; ++q) {
            const double len = port[q] < g.nx ? g.b : g.a;   // an x-directed rooftop crosses an edge of length b
            sum += len * std::complex<double>(h[r * port_len + q].x, h[r * port_len + q].y);
        }
        const std::complex<double> z = (voltages ? voltages[r] : 1.0) / sum;
        zin[2 * r] = z.real();
        zin[2 * r + 1] = z.imag();
    }
}

}  // namespace rt

// Green's Toeplitz generation on gfx950, FP64: the three generators of the Green's matrix from seven cell-pair
// moments per unique cell offset.
//
// Reference: src/+rooftopMoM/+assembler/+greensMatrix/@MomentAssembler/generate.m:10-58 and its helpers:
//   near pass      integrateCellPairs_.m:13-44 with +utils/integrateMoments.m:17-38 (remainder on touching offsets,
//                  +singularIntegrals/evaluateRemainder.m:13-14);
//   far pass       integrateLattice_.m:17-127 -- the kernel once per point of the shared lattice of
//                  (countX+1)(countY+1) blocks, four quadrant sums per block, an offset's moments are the quadrant
//                  sums of its own block and its +1 neighbours;
//   touching       +singularIntegrals/generatePowerMoments.m:17-155, integrateMonomial.m:18-80 (closed forms);
//   accumulation   +utils/accumulateMoments.m:27-133 -- here a gather: every generator entry reads at most three
//                  offsets' terms, so no atomics are needed.
// The moments of all passes are summed per offset before the (linear) accumulation.
#include <algorithm>
#include <cmath>
#include <vector>

#include "rooftop.hpp"

namespace rt {

static const double kZ0 = 376.730313668;   // +assembler/Constants.m:11

// ------------------------------------------------------------------------------------------------------------------
// Host: Gauss-Legendre rules (Newton on the Legendre recurrence in long double) and the reduced moment rule.
// ------------------------------------------------------------------------------------------------------------------
void gauss_legendre(int n, std::vector<double> &pts, std::vector<double> &wts) {
    pts.assign(n, 0.0);
    wts.assign(n, 0.0);
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<long double> x(n), w(n);
    // P_n(z) and P_{n-1}(z) by the three-term recurrence.
    auto legendre = [n](long double z, long double &pn, long double &pm) {
        pm = 1;
        pn = z;
        for (int k = 2; k <= n; ++k) {
            long double pk = ((2 * k - 1) * z * pn - (k - 1) * pm) / k;
            pm = pn;
            pn = pk;
        }
    };
    for (int i = 0; i < n; ++i) {
        long double z = std::cos(pi * (i + 0.75L) / (n + 0.5L));
        long double pn, pm;
        for (int it = 0; it < 100; ++it) {
            legendre(z, pn, pm);
            long double dz = pn / (n * (z * pn - pm) / (z * z - 1));
            z -= dz;
            if (std::fabs(dz) < 1e-22L) break;
        }
        legendre(z, pn, pm);
        long double dp = n * (z * pn - pm) / (z * z - 1);
        x[i] = z;
        w[i] = 2 / ((1 - z * z) * dp * dp);
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return x[a] < x[b]; });
    for (int i = 0; i < n; ++i) {   // map [-1, 1] -> [0, 1]
        pts[i] = static_cast<double>((x[idx[i]] + 1) / 2);
        wts[i] = static_cast<double>(w[idx[i]] / 2);
    }
}

struct Rule {
    std::vector<double> nodes_x, nodes_y;   // nu, eta
    std::vector<double> weights;            // [7][Px*Py], nu fastest
};

// +greensMatrix/+utils/generateReducedRule.m:21-66.
static void axis_rule(int order, int parts, std::vector<double> &nodes, std::vector<double> &w4) {
    std::vector<double> p, pw;
    gauss_legendre(order, p, pw);
    std::vector<double> right;
    for (int part = 0; part < parts; ++part)
        for (int i = 0; i < order; ++i) right.push_back((p[i] + part) / parts);
    nodes.clear();
    for (double r : right) nodes.push_back(r - 1);
    for (double r : right) nodes.push_back(r);
    const int P = static_cast<int>(nodes.size());
    w4.assign(static_cast<size_t>(P) * 4, 0.0);
    for (int i = 0; i < P; ++i) {
        const double nu = nodes[i];
        const double gw = pw[i % order] / parts;
        const bool left = nu < 0;
        const double dist = std::fabs(nu);
        const double w10 = left ? (1 + nu) * (1 + nu) / 2 : (1 - nu * nu) / 2;
        const double w01 = left ? (1 - nu * nu) / 2 : (1 - nu) * (1 - nu) / 2;
        w4[i * 4 + 0] = gw * (1 - dist);
        w4[i * 4 + 1] = gw * w10;
        w4[i * 4 + 2] = gw * w01;
        w4[i * 4 + 3] = gw * (1 - dist) * (1 - dist) * (2 + dist) / 6;
    }
}

static Rule reduced_rule(int order, int parts_x, int parts_y) {
    Rule r;
    std::vector<double> wx, wy;
    axis_rule(order, parts_x, r.nodes_x, wx);
    axis_rule(order, parts_y, r.nodes_y, wy);
    const int px = static_cast<int>(r.nodes_x.size()), py = static_cast<int>(r.nodes_y.size());
    r.weights.assign(static_cast<size_t>(7) * px * py, 0.0);
    for (int iy = 0; iy < py; ++iy)
        for (int ix = 0; ix < px; ++ix) {
            const size_t k = static_cast<size_t>(iy) * px + ix;
            for (int m = 0; m < 4; ++m) r.weights[m * static_cast<size_t>(px) * py + k] = wy[iy * 4] * wx[ix * 4 + m];
            for (int m = 0; m < 3; ++m)
                r.weights[(4 + m) * static_cast<size_t>(px) * py + k] = wy[iy * 4 + 1 + m] * wx[ix * 4];
        }
    return r;
}

// ------------------------------------------------------------------------------------------------------------------
// Device: kernel values.
// ------------------------------------------------------------------------------------------------------------------
// exp(-ikR)/R, and the remainder exp(-ikR)/R - (1/R - k^2 R/2 + k^4 R^3/24) with expm1(-i kR) = -2 sin^2(kR/2) - i sin(kR).
__device__ __forceinline__ c128 green(double r, double k) {
    double s, c;
    sincos(k * r, &s, &c);
    return mk(c / r, -s / r);
}

__device__ __forceinline__ c128 remainder_kernel(double r, double k) {
    double s, c;
    sincos(k * r, &s, &c);
    const double sh = sin(0.5 * k * r);
    const double k2 = k * k;
    return mk(-2.0 * sh * sh / r + k2 / 2 * r - k2 * k2 / 24 * r * r * r, -s / r);
}

// Near pass: one block per near offset; threads over the rule's nodes; seven weighted sums reduced in a fixed order.
__global__ void __launch_bounds__(256) near_moments(const double *__restrict__ nodes_x,
                                                    const double *__restrict__ nodes_y,
                                                    const double *__restrict__ weights, int px, int py, int near_x,
                                                    double a, double b, double k, c128 *__restrict__ out) {
    const int off = blockIdx.x;
    const int dx = off % near_x, dy = off / near_x;
    const bool touching = dx <= 1 && dy <= 1;
    const int P = px * py;
    double acc[14];
    for (int i = 0; i < 14; ++i) acc[i] = 0;
    for (int n = threadIdx.x; n < P; n += blockDim.x) {
        const int ix = n % px, iy = n / px;
        const double r = hypot(a * (nodes_x[ix] - dx), b * (nodes_y[iy] - dy));
        const c128 kv = touching ? remainder_kernel(r, k) : green(r, k);
        for (int m = 0; m < 7; ++m) {
            const double w = weights[static_cast<size_t>(m) * P + n];
            acc[2 * m] += w * kv.x;
            acc[2 * m + 1] += w * kv.y;
        }
    }
    __shared__ double red[14][256];
    for (int i = 0; i < 14; ++i) red[i][threadIdx.x] = acc[i];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int i = 0; i < 14; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 7) out[off * 7 + threadIdx.x] = mk(red[2 * threadIdx.x][0], red[2 * threadIdx.x + 1][0]);
}

// Far lattice: one thread per lattice block (m, n), 0 <= m <= countX, 0 <= n <= countY. The block's order^2 kernel
// values at x = a (p - m), y = b (q - n) (p, q the Gauss-Legendre points on [0, 1]) are contracted with the four
// quadrant weight sets; out[(q * 7 + i) * blocks + block] holds quadrant q's moment i.
__global__ void __launch_bounds__(256) lattice_sums(const double *__restrict__ points, const double *__restrict__ qw,
                                                    int order, int bx, int by, double a, double b, double k,
                                                    c128 *__restrict__ out) {
    extern __shared__ double sw[];   // [4][7][order*order]
    const int nw = 28 * order * order;
    for (int i = threadIdx.x; i < nw; i += blockDim.x) sw[i] = qw[i];
    __syncthreads();
    const int64_t blocks = static_cast<int64_t>(bx) * by;
    const int64_t blk = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (blk >= blocks) return;
    const int m = static_cast<int>(blk % bx), n = static_cast<int>(blk / bx);
    double acc[56];
    for (int i = 0; i < 56; ++i) acc[i] = 0;
    const int o2 = order * order;
    for (int iq = 0; iq < order; ++iq) {
        const double y = b * (points[iq] - n);
        for (int ip = 0; ip < order; ++ip) {
            const double x = a * (points[ip] - m);
            const c128 kv = green(hypot(x, y), k);
            const int pt = iq * order + ip;
#pragma unroll
            for (int qi = 0; qi < 28; ++qi) {
                const double w = sw[qi * o2 + pt];
                acc[2 * qi] += w * kv.x;
                acc[2 * qi + 1] += w * kv.y;
            }
        }
    }
    for (int qi = 0; qi < 28; ++qi) out[qi * blocks + blk] = mk(acc[2 * qi], acc[2 * qi + 1]);
}

// Terms of every offset: moments from the near pass (near offsets), the lattice quadrant sums (far offsets), plus
// the closed-form extracted moments on the touching offsets; then the seven combinations of accumulateMoments.m:43-49
// times the scales. terms[i * cx * cy + dx + cx * dy].
__global__ void offset_terms(const c128 *__restrict__ near, int near_x, int near_y, const c128 *__restrict__ lat,
                             int bx, int by, const c128 *__restrict__ extracted, int cx, int cy, c128 s_e, c128 s_m,
                             c128 *__restrict__ terms) {
    const int64_t count = static_cast<int64_t>(cx) * cy;
    const int64_t id = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (id >= count) return;
    const int dx = static_cast<int>(id % cx), dy = static_cast<int>(id / cx);
    c128 mo[7];
    if (dx < near_x && dy < near_y) {
        for (int i = 0; i < 7; ++i) mo[i] = near[(dx + near_x * dy) * 7 + i];
    } else {
        const int64_t blocks = static_cast<int64_t>(bx) * by;
        // Quadrant q = halfX + 2 halfY; a left half (0) takes the block of the offset plus 1.
        const int64_t b00 = (dx + 1) + static_cast<int64_t>(bx) * (dy + 1);
        const int64_t b10 = dx + static_cast<int64_t>(bx) * (dy + 1);
        const int64_t b01 = (dx + 1) + static_cast<int64_t>(bx) * dy;
        const int64_t b11 = dx + static_cast<int64_t>(bx) * dy;
        for (int i = 0; i < 7; ++i)
            mo[i] = lat[(0 * 7 + i) * blocks + b00] + lat[(1 * 7 + i) * blocks + b10] +
                    lat[(2 * 7 + i) * blocks + b01] + lat[(3 * 7 + i) * blocks + b11];
    }
    if (dx <= 1 && dy <= 1) {
        const int col = dx + 2 * dy;   // offsets (0,0), (1,0), (0,1), (1,1)
        for (int i = 0; i < 7; ++i) mo[i] = mo[i] + extracted[i * 4 + col];
    }
    c128 t[7];
    t[0] = s_e * mo[0];
    t[1] = s_m * (mo[0] - mo[1] - mo[2] + 2.0 * mo[3]);
    t[2] = s_m * (mo[1] - mo[3]);
    t[3] = s_m * (mo[2] - mo[3]);
    t[4] = s_m * (mo[0] - mo[4] - mo[5] + 2.0 * mo[6]);
    t[5] = s_m * (mo[4] - mo[6]);
    t[6] = s_m * (mo[5] - mo[6]);
    for (int i = 0; i < 7; ++i) terms[i * count + id] = t[i];
}

// Generators from the terms (accumulateMoments.m:62-132 as a gather).
__global__ void gather_generators(const c128 *__restrict__ terms, int cx, int cy, c128 *__restrict__ e,
                                  c128 *__restrict__ mxx, c128 *__restrict__ myy) {
    const int64_t count = static_cast<int64_t>(cx) * cy;
    const int64_t id = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (id >= count) return;
    const int dx = static_cast<int>(id % cx), dy = static_cast<int>(id / cx);
    auto T = [&](int i, int x, int y) { return terms[i * count + x + static_cast<int64_t>(cx) * y]; };
    e[id] = T(0, dx, dy);
    if (mxx && dx <= cx - 2) {
        c128 v = T(1, dx, dy) + (dx == 0 ? 2.0 : 1.0) * T(2, dx + 1, dy);
        if (dx >= 1) v = v + T(3, dx - 1, dy);
        mxx[dx + static_cast<int64_t>(cx - 1) * dy] = v;
    }
    if (myy && dy <= cy - 2) {
        c128 v = T(4, dx, dy) + (dy == 0 ? 2.0 : 1.0) * T(5, dx, dy + 1);
        if (dy >= 1) v = v + T(6, dx, dy - 1);
        myy[dx + static_cast<int64_t>(cx) * dy] = v;
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Touching closed forms (FP64).
// ------------------------------------------------------------------------------------------------------------------
__device__ double monomial(int power, int m, int n, double a, double b) {
    if (a == 0 || b == 0) return 0;
    const double R = sqrt(a * a + b * b);
    const double lb = asinh(b / a), la = asinh(a / b);
    const double a2 = a * a, b2 = b * b, a4 = a2 * a2, b4 = b2 * b2;
    const int key = (power + 1) * 100 + m * 10 + n;   // power -1 -> 0, 1 -> 200, 3 -> 400
    switch (key) {
        case 0: return a * lb + b * la;
        case 1: return (-a2 + a * R + b2 * la) / 2;
        case 10: return (-b2 + b * R + a2 * lb) / 2;
        case 11: return (-a2 * a - b2 * b + (a2 + b2) * R) / 3;
        case 20: return (a * b * R + 2 * a2 * a * lb - b2 * b * la) / 6;
        case 21: return (-2 * a4 + a * (2 * a2 + b2) * R - b4 * la) / 8;
        case 30: return (2 * b4 + b * (a2 - 2 * b2) * R + 3 * a4 * lb) / 12;
        case 31: return (-3 * a4 * a + 2 * b4 * b + (3 * a4 + a2 * b2 - 2 * b4) * R) / 15;
        case 200: return (2 * a * b * R + a2 * a * lb + b2 * b * la) / 6;
        case 201: return (-2 * a4 + a * (2 * a2 + 5 * b2) * R + 3 * b4 * la) / 24;
        case 210: return (-2 * b4 + b * (5 * a2 + 2 * b2) * R + 3 * a4 * lb) / 24;
        case 211: return (-a4 * a - b4 * b + (a4 + 2 * a2 * b2 + b4) * R) / 15;
        case 220: return (a * b * (6 * a2 + b2) * R + 4 * a4 * a * lb - b4 * b * la) / 40;
        case 221: return (-8 * a4 * a2 + a * (8 * a4 + 14 * a2 * b2 + 3 * b4) * R - 3 * b4 * b2 * la) / 144;
        case 230: return (4 * b4 * b2 + b * (21 * a4 + 2 * a2 * b2 - 4 * b4) * R + 15 * a4 * a2 * lb) / 180;
        case 231:
            return (-5 * a4 * a2 * a + 2 * b4 * b2 * b + (5 * a4 * a2 + 8 * a4 * b2 + a2 * b4 - 2 * b4 * b2) * R) / 105;
        case 400: return (7 * a * b * (a2 + b2) * R + 3 * a4 * a * lb + 3 * b4 * b * la) / 40;
        case 401: return (-8 * a4 * a2 + a * (8 * a4 + 26 * a2 * b2 + 33 * b4) * R + 15 * b4 * b2 * la) / 240;
        case 410: return (-8 * b4 * b2 + b * (33 * a4 + 26 * a2 * b2 + 8 * b4) * R + 15 * a4 * a2 * lb) / 240;
        case 411:
            return (-a4 * a2 * a - b4 * b2 * b + (a4 * a2 + 3 * a4 * b2 + 3 * a2 * b4 + b4 * b2) * R) / 35;
        case 420:
            return (a * b * (38 * a4 + 26 * a2 * b2 + 3 * b4) * R + 18 * a4 * a2 * a * lb - 3 * b4 * b2 * b * la) / 336;
        case 421:
            return (-48 * a4 * a4 + a * (48 * a4 * a2 + 136 * a4 * b2 + 118 * a2 * b4 + 15 * b4 * b2) * R -
                    15 * b4 * b4 * la) / 1920;
        case 430:
            return (16 * b4 * b4 + b * (215 * a4 * a2 + 134 * a4 * b2 + 8 * a2 * b4 - 16 * b4 * b2) * R +
                    105 * a4 * a4 * lb) / 2240;
        case 431:
            return (-7 * a4 * a4 * a + 2 * b4 * b4 * b +
                    (7 * a4 * a4 + 19 * a4 * a2 * b2 + 15 * a4 * b4 + a2 * b4 * b2 - 2 * b4 * b4) * R) / 315;
    }
    return 0;
}

// Correlation weights W00, W10, W01, W11, ascending coefficients on [-1, 0] and [0, 1] (generatePowerMoments.m:66-71).
__constant__ double kWeights[4][2][4] = {
    {{1, 1, 0, 0}, {1, -1, 0, 0}},
    {{0.5, 1, 0.5, 0}, {0.5, 0, -0.5, 0}},
    {{0.5, 0, -0.5, 0}, {0.5, -1, 0.5, 0}},
    {{1.0 / 3, 0.5, 0, -1.0 / 6}, {1.0 / 3, -0.5, 0, 1.0 / 6}}};
__constant__ int kWeightLen[4] = {2, 3, 3, 4};

// The weight as a polynomial in u = sigma * size * (nu - offset) >= 0 on one part, and its interval (0: [0, size],
// 1: [size, 2 size]) -- generatePowerMoments.m:121-155.
__device__ int compose_part(int k, int part, int offset, double size, double *coef, int count) {
    const double lo = (part ? 0.0 : -1.0) - offset, hi = (part ? 1.0 : 0.0) - offset;
    const double sigma = hi <= 0 ? -1.0 : 1.0;
    const int interval = static_cast<int>(fmin(sigma * lo, sigma * hi));
    double term[5] = {1, 0, 0, 0, 0};
    for (int i = 0; i < count; ++i) coef[i] = 0;
    int deg = 0;
    for (int j = 0; j < kWeightLen[k]; ++j) {
        for (int i = 0; i <= deg && i < count; ++i) coef[i] += kWeights[k][part][j] * term[i];
        // term *= (offset + sigma/size * u)
        double next[5] = {0, 0, 0, 0, 0};
        for (int i = 0; i <= deg; ++i) {
            next[i] += term[i] * offset;
            if (i + 1 < 5) next[i + 1] += term[i] * sigma / size;
        }
        for (int i = 0; i < 5; ++i) term[i] = next[i];
        ++deg;
    }
    return interval;
}

// Sum over the wave, the same in every lane (a butterfly of lane exchanges, in a fixed order).
__device__ __forceinline__ double wave_total(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// M00, M10, M01, M11 at the four touching offsets of an a x b cell (generatePowerMoments.m:57-153), computed by one
// wave into out[4][4] (LDS). The corner values of the mixed antiderivatives (72, two per lane), the rectangle integrals
// as their corner differences (64, one per lane), then the rectangles at least their longer side away from the origin
// by a 20 x 20 Gauss-Legendre rule whose 400 points are spread over the lanes (generatePowerMoments.m:94-118: there
// the closed form is a difference of large, nearly equal terms), and the 16 moments, one per lane.
__device__ void along_x_wave(double a, double b, int power, const double *gl_p, const double *gl_w, double *G,
                             double *rect, double out[4][4]) {
    const int lane = threadIdx.x & 63;
    const double ratio = b / a;
    // G[((m * 2 + n) * 3 + i) * 3 + j] = monomial(power, m, n, i, ratio j).
    for (int e = lane; e < 72; e += 64) {
        const int j = e % 3, i = (e / 3) % 3, mn = e / 9;
        G[e] = monomial(power, mn / 2, mn % 2, static_cast<double>(i), ratio * j);
    }
    __syncthreads();
    {   // rect[((m * 2 + n) * 2 + s) * 2 + t]
        const int t = lane & 1, s = (lane >> 1) & 1, mn = lane >> 2;
        const double *g = G + mn * 9;
        rect[lane] = g[(s + 1) * 3 + t + 1] - g[s * 3 + t + 1] - g[(s + 1) * 3 + t] + g[s * 3 + t];
    }
    __syncthreads();
    for (int kk = 0; kk < 2; ++kk)
        for (int l = 0; l < 2; ++l) {
            if (hypot(static_cast<double>(kk), ratio * l) < fmax(1.0, ratio)) continue;   // uniform
            double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int e = lane; e < 400; e += 64) {
                const int i = e / 20, j = e % 20;
                const double u = kk + gl_p[i], v = ratio * (l + gl_p[j]);
                const double r = sqrt(u * u + v * v);
                const double rp = power == -1 ? 1.0 / r : power == 1 ? r : r * r * r;
                const double w = gl_w[i] * gl_w[j] * rp;
                double um = 1.0;
                for (int m = 0; m < 4; ++m) {
                    acc[m * 2] += w * um;
                    acc[m * 2 + 1] += w * um * v;
                    um *= u;
                }
            }
            for (int mn = 0; mn < 8; ++mn) {
                const double total = wave_total(acc[mn]);
                if (lane == 0) rect[(mn * 2 + kk) * 2 + l] = ratio * total;
            }
        }
    __syncthreads();
    if (lane < 16) {
        const int offs[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
        const int k = lane / 4, l = lane % 4;
        double acc = 0;
        for (int s = 0; s < 2; ++s) {
            double ca[4];
            const int ia = compose_part(k, s, offs[l][0], 1.0, ca, 4);
            for (int t = 0; t < 2; ++t) {
                double cb[2];
                const int ib = compose_part(0, t, offs[l][1], ratio, cb, 2);
                for (int m = 0; m < 4; ++m)
                    for (int n = 0; n < 2; ++n) acc += ca[m] * rect[((m * 2 + n) * 2 + ia) * 2 + ib] * cb[n];
            }
        }
        out[k][l] = pow(a, power) * acc / ratio;
    }
}

// The extracted moments 7x4 (generateExtractedMoments.m:15-18): P(-1) - k^2/2 P(1) + k^4/24 P(3). Six waves
// evaluate the six power moment tables (three powers along x and along y) at once; thread 0 combines.
__global__ void __launch_bounds__(384) touching_moments(double a, double b, double k, const double *gl_p,
                                                        const double *gl_w, c128 *__restrict__ out) {
    __shared__ double tab[6][4][4];
    __shared__ double G[6][72], rect[6][64];
    const int powers[3] = {-1, 1, 3};
    const int w = threadIdx.x >> 6;
    if (w < 3) along_x_wave(a, b, powers[w], gl_p, gl_w, G[w], rect[w], tab[w]);
    else along_x_wave(b, a, powers[w - 3], gl_p, gl_w, G[w], rect[w], tab[w]);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double factors[3] = {1.0, -k * k / 2, k * k * k * k / 24};
    double acc[7][4];
    for (int i = 0; i < 7; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] = 0;
    const int perm[4] = {0, 2, 1, 3};
    for (int pi = 0; pi < 3; ++pi)
        for (int j = 0; j < 4; ++j) {
            for (int i = 0; i < 4; ++i) acc[i][j] += factors[pi] * tab[pi][i][j];
            for (int i = 0; i < 3; ++i) acc[4 + i][j] += factors[pi] * tab[3 + pi][1 + i][perm[j]];
        }
    for (int i = 0; i < 7; ++i)
        for (int j = 0; j < 4; ++j) out[i * 4 + j] = mk(acc[i][j], 0.0);
}

// ------------------------------------------------------------------------------------------------------------------
// Host orchestration.
// ------------------------------------------------------------------------------------------------------------------
template <typename T>
static void upload(Context *ctx, DeviceArray<T> &arr, const std::vector<T> &v, hipStream_t s, const char *what) {
    arr.allocate(ctx, v.size(), what);
    if (!v.empty()) RT_HIP(hipMemcpyAsync(arr.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

void Problem::assemble() {
    hipStream_t s = ctx->stream;
    const int cx = g.cx, cy = g.cy;
    const double a = g.a, b = g.b;
    const double sizes[2] = {a, b};
    const double big = std::max(a, b), small = std::min(a, b);
    // countNearOffsets_.m:5-10
    int nearc[2];
    const int counts[2] = {cx, cy};
    for (int d = 0; d < 2; ++d) {
        nearc[d] = 0;
        for (int o = 0; o < counts[d]; ++o)
            if (std::max(o - 1, 0) * sizes[d] < 3 * big) ++nearc[d];
    }
    // generate.m:30-31: the near rule split by the rounded aspect ratio (MATLAB round: half away from zero).
    const int parts_x = static_cast<int>(std::floor(a / small + 0.5));
    const int parts_y = static_cast<int>(std::floor(b / small + 0.5));
    Rule near_rule = reduced_rule(near_order, parts_x, parts_y);
    Rule far_rule = reduced_rule(far_order, 1, 1);

    const c128 s_e = mk(0.0, kZ0 / (4 * M_PI) * (a * b) * (a * b) / k);
    const c128 s_m = mk(0.0, kZ0 / (4 * M_PI) * (a * b) * (a * b) * k);

    // Far quadrant weights [q][7][order^2], x fastest (integrateLattice_.m:92-111).
    const int o = far_order;
    const int P2 = 2 * o;
    std::vector<double> qw(static_cast<size_t>(28) * o * o);
    for (int hy = 0; hy < 2; ++hy)
        for (int hx = 0; hx < 2; ++hx) {
            const int q = hx + 2 * hy;
            for (int m = 0; m < 7; ++m)
                for (int iq = 0; iq < o; ++iq)
                    for (int ip = 0; ip < o; ++ip) {
                        const int kx = hx * o + ip, ky = hy * o + iq;
                        qw[(static_cast<size_t>(q) * 7 + m) * o * o + iq * o + ip] =
                            far_rule.weights[static_cast<size_t>(m) * P2 * P2 + static_cast<size_t>(ky) * P2 + kx];
                    }
        }
    std::vector<double> points(far_rule.nodes_x.begin() + o, far_rule.nodes_x.end());
    std::vector<double> gl_p, gl_w;
    gauss_legendre(20, gl_p, gl_w);

    DeviceArray<double> d_nx, d_ny, d_nw, d_pts, d_qw, d_glp, d_glw;
    upload(ctx, d_nx, near_rule.nodes_x, s, "assembly rules");
    upload(ctx, d_ny, near_rule.nodes_y, s, "assembly rules");
    upload(ctx, d_nw, near_rule.weights, s, "assembly rules");
    upload(ctx, d_pts, points, s, "assembly rules");
    upload(ctx, d_qw, qw, s, "assembly rules");
    upload(ctx, d_glp, gl_p, s, "assembly rules");
    upload(ctx, d_glw, gl_w, s, "assembly rules");

    const int bx = cx + 1, by = cy + 1;
    const int64_t blocks = static_cast<int64_t>(bx) * by;
    const int64_t offsets = static_cast<int64_t>(cx) * cy;
    DeviceArray<c128> near, lat, extracted, terms;
    near.allocate(ctx, static_cast<size_t>(nearc[0]) * nearc[1] * 7, "assembly chunk");
    lat.allocate(ctx, static_cast<size_t>(blocks) * 28, "assembly chunk");
    extracted.allocate(ctx, 28, "assembly chunk");
    terms.allocate(ctx, static_cast<size_t>(offsets) * 7, "assembly chunk");

    const int npx = static_cast<int>(near_rule.nodes_x.size()), npy = static_cast<int>(near_rule.nodes_y.size());
    hipLaunchKernelGGL(near_moments, dim3(nearc[0] * nearc[1]), dim3(256), 0, s, d_nx.ptr, d_ny.ptr, d_nw.ptr, npx,
                       npy, nearc[0], a, b, k, near.ptr);
    RT_LAUNCH_CHECK();
    const size_t lds = static_cast<size_t>(28) * o * o * sizeof(double);
    if (lds > 64 * 1024)
        RT_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&lattice_sums),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL(lattice_sums, dim3(static_cast<unsigned>((blocks + 255) / 256)), dim3(256), lds, s, d_pts.ptr,
                       d_qw.ptr, o, bx, by, a, b, k, lat.ptr);
    RT_LAUNCH_CHECK();
    hipLaunchKernelGGL(touching_moments, dim3(1), dim3(384), 0, s, a, b, k, d_glp.ptr, d_glw.ptr, extracted.ptr);
    RT_LAUNCH_CHECK();
    hipLaunchKernelGGL(offset_terms, dim3(static_cast<unsigned>((offsets + 255) / 256)), dim3(256), 0, s, near.ptr,
                       nearc[0], nearc[1], lat.ptr, bx, by, extracted.ptr, cx, cy, s_e, s_m, terms.ptr);
    RT_LAUNCH_CHECK();
    if (g.has_x) mxx.allocate(ctx, static_cast<size_t>(cx - 1) * cy, "The generators");
    if (g.has_y) myy.allocate(ctx, static_cast<size_t>(cx) * (cy - 1), "The generators");
    e.allocate(ctx, static_cast<size_t>(cx) * cy, "The generators");
    hipLaunchKernelGGL(gather_generators, dim3(static_cast<unsigned>((offsets + 255) / 256)), dim3(256), 0, s,
                       terms.ptr, cx, cy, e.ptr, g.has_x ? mxx.ptr : nullptr, g.has_y ? myy.ptr : nullptr);
    RT_LAUNCH_CHECK();
    RT_HIP(hipStreamSynchronize(s));
    assembled = true;
}

void Problem::ensure_generators() {
    if (!assembled) assemble();
}

}  // namespace rt

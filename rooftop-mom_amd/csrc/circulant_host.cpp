// Host side of the circulant preconditioners that the device folds do not cover: the superoptimal kernel of
// Tyrtyshnikov (+preconditioner/+kernels/generateSuperoptimal.m with +utils/generatePartSpectra.m and
// +utils/generateOptimalOfProduct.m) and the kernels applied to generators that are not even (the XY coupling block of
// the blockwise builder, generateBlockwiseSpectra.m). These are one-off set-up computations on count_x x count_y
// grids (powers of two), in FP64, with a radix-2 FFT of their own.
#include "circulant_host.hpp"

#include <cmath>

#include "rooftop.hpp"

namespace rt {

namespace {

constexpr double kPi = 3.141592653589793;

// In-place radix-2 FFT of n (a power of two) values at a[0], a[stride], ...: forward sum x(k) exp(-2 pi i k j / n),
// inverse with the opposite sign and no scaling.
void fft1(cd *a, int n, int64_t stride, bool inverse) {
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i * stride], a[j * stride]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = (inverse ? 2 : -2) * kPi / len;
        for (int k = 0; k < len / 2; ++k) {
            const cd w(std::cos(ang * k), std::sin(ang * k));
            for (int i = 0; i < n; i += len) {
                cd &u = a[(i + k) * stride], &v = a[(i + k + len / 2) * stride];
                const cd t = w * v;
                v = u - t;
                u = u + t;
            }
        }
    }
}

// One-dimensional transforms of every column (axis 0, along x) or every row (axis 1, along y) of a[x + cx y].
void fft_axis(std::vector<cd> &a, int cx, int cy, int axis, bool inverse) {
    if (axis == 0) {
        for (int y = 0; y < cy; ++y) fft1(a.data() + static_cast<int64_t>(y) * cx, cx, 1, inverse);
    } else {
        for (int x = 0; x < cx; ++x) fft1(a.data() + x, cy, cx, inverse);
    }
    if (inverse) {
        const double s = 1.0 / (axis == 0 ? cx : cy);
        for (auto &v : a) v *= s;
    }
}

double window(int o, int P, int kind, double bandwidth) {
    const int a = o < 0 ? -o : o;
    switch (kind) {
    case RT_CIRCULANT_STRANG: return 2 * a < P ? 1.0 : (2 * a == P ? 0.5 : 0.0);
    case RT_CIRCULANT_TCHAN: return static_cast<double>(P - a) / P;
    case RT_CIRCULANT_RCHAN: return 1.0;
    default: return std::fmax(0.0, 1.0 - a / bandwidth);   // Huckle
    }
}

// The part generators of generatePartSpectra.m along x of the two-sided t: c(k) = (t(k) + t(k - cx)) / 2 and
// s(k) = (t(k) - t(k - cx)) / 2 for k = 0 .. cx - 1 (t(-cx) = 0), every y offset kept: [k + cx (q + cy - 1)].
void split_x(const TwoSided &t, std::vector<cd> &c, std::vector<cd> &s) {
    const int cx = t.cx, cy = t.cy, wy = 2 * cy - 1;
    c.assign(static_cast<size_t>(cx) * wy, 0.0);
    s.assign(static_cast<size_t>(cx) * wy, 0.0);
    for (int q = -(cy - 1); q < cy; ++q)
        for (int k = 0; k < cx; ++k) {
            const cd pos = t.at(k, q), neg = k == 0 ? cd(0.0) : t.at(k - cx, q);
            c[k + static_cast<size_t>(cx) * (q + cy - 1)] = (pos + neg) / 2.0;
            s[k + static_cast<size_t>(cx) * (q + cy - 1)] = (pos - neg) / 2.0;
        }
}

// The same along y of an x-split part h [k + cx (q + cy - 1)], into [k + cx l], l = 0 .. cy - 1.
void split_y(const std::vector<cd> &h, int cx, int cy, std::vector<cd> &c, std::vector<cd> &s) {
    c.assign(static_cast<size_t>(cx) * cy, 0.0);
    s.assign(static_cast<size_t>(cx) * cy, 0.0);
    for (int l = 0; l < cy; ++l)
        for (int k = 0; k < cx; ++k) {
            const cd pos = h[k + static_cast<size_t>(cx) * (l + cy - 1)];
            const cd neg = l == 0 ? cd(0.0) : h[k + static_cast<size_t>(cx) * (l - 1)];
            c[k + static_cast<size_t>(cx) * l] = (pos + neg) / 2.0;
            s[k + static_cast<size_t>(cx) * l] = (pos - neg) / 2.0;
        }
}

// The spectrum of K(d) = 1 / (count^2 sin^2(pi (2 d + 1) / (2 count))), d = 0 .. count - 1.
std::vector<cd> skew_kernel(int count) {
    std::vector<cd> k(count);
    for (int d = 0; d < count; ++d) {
        const double sn = std::sin(kPi * (2 * d + 1) / (2.0 * count));
        k[d] = 1.0 / (static_cast<double>(count) * count * sn * sn);
    }
    fft1(k.data(), count, 1, false);
    return k;
}

// The circular convolution along one axis with the kernel K (its spectrum kern).
void convolve(std::vector<cd> &v, int cx, int cy, int axis, const std::vector<cd> &kern) {
    fft_axis(v, cx, cy, axis, false);
    for (int y = 0; y < cy; ++y)
        for (int x = 0; x < cx; ++x) v[x + static_cast<size_t>(cx) * y] *= kern[axis == 0 ? x : y];
    fft_axis(v, cx, cy, axis, true);
}

}  // namespace

cd TwoSided::at(int p, int q) const {
    if (p <= -cx || p >= cx || q <= -cy || q >= cy) return 0.0;
    return v[(p + cx - 1) + static_cast<size_t>(2 * cx - 1) * (q + cy - 1)];
}

void fft2_host(std::vector<cd> &a, int cx, int cy, bool inverse) {
    fft_axis(a, cx, cy, 0, inverse);
    fft_axis(a, cx, cy, 1, inverse);
}

std::vector<cd> fold_two_sided(const TwoSided &t, int kind, double bwx, double bwy) {
    const int cx = t.cx, cy = t.cy;
    std::vector<cd> c(static_cast<size_t>(cx) * cy, 0.0);
    for (int q = -(cy - 1); q < cy; ++q) {
        const double wy = window(q, cy, kind, bwy);
        if (wy == 0) continue;
        const int ky = ((q % cy) + cy) % cy;
        for (int p = -(cx - 1); p < cx; ++p) {
            const double w = window(p, cx, kind, bwx) * wy;
            if (w == 0) continue;
            c[((p % cx) + cx) % cx + static_cast<size_t>(cx) * ky] += w * t.at(p, q);
        }
    }
    return c;
}

std::vector<cd> superoptimal_spectrum(const TwoSided &t) {
    const int cx = t.cx, cy = t.cy;
    const size_t n = static_cast<size_t>(cx) * cy;
    // The four part spectra, P[m][k]: m along x, k along y, 0 circulant, 1 skew-circulant (modulated).
    std::vector<cd> P[2][2];
    {
        std::vector<cd> hx[2];
        split_x(t, hx[0], hx[1]);
        for (int m = 0; m < 2; ++m) {
            split_y(hx[m], cx, cy, P[m][0], P[m][1]);
            for (int k = 0; k < 2; ++k) {
                std::vector<cd> &part = P[m][k];
                for (int l = 0; l < cy; ++l)
                    for (int x = 0; x < cx; ++x) {
                        cd f = 1.0;
                        if (m == 1) f *= std::polar(1.0, kPi * x / cx);
                        if (k == 1) f *= std::polar(1.0, kPi * l / cy);
                        part[x + static_cast<size_t>(cx) * l] *= f;
                    }
                fft2_host(part, cx, cy, false);
            }
        }
    }
    const std::vector<cd> kern[2] = {skew_kernel(cx), skew_kernel(cy)};
    // c(T T^H) (generateOptimalOfProduct.m): each pair of parts (A: ta, B: tb) adds A's eigenvalues times the
    // conjugate of B's, convolved first along a direction where one of the two is skew, the product convolved where
    // both are. The operands' convolutions are formed once per part; the products that need the same convolutions
    // are summed before them (the convolution is linear).
    std::vector<cd> conv[2][2][3];   // [mx][my][1 along x, 2 along y, 3 both] (index - 1)
    for (int mx = 0; mx < 2; ++mx)
        for (int my = 0; my < 2; ++my)
            for (int which = 1; which <= 3; ++which) {
                std::vector<cd> v = P[mx][my];
                if (which & 1) convolve(v, cx, cy, 0, kern[0]);
                if (which & 2) convolve(v, cx, cy, 1, kern[1]);
                conv[mx][my][which - 1] = std::move(v);
            }
    auto operand = [&](int mx, int my, int which) -> const std::vector<cd> & {
        return which ? conv[mx][my][which - 1] : P[mx][my];
    };
    std::vector<cd> acc[4];   // products to convolve: [1 along x | 2 along y]
    for (auto &a : acc) a.assign(n, 0.0);
    for (int ax = 0; ax < 2; ++ax)
        for (int ay = 0; ay < 2; ++ay)
            for (int bx = 0; bx < 2; ++bx)
                for (int by = 0; by < 2; ++by) {
                    const int wa = (ax == 1 && bx == 0 ? 1 : 0) | (ay == 1 && by == 0 ? 2 : 0);
                    const int wb = (ax == 0 && bx == 1 ? 1 : 0) | (ay == 0 && by == 1 ? 2 : 0);
                    const int wp = (ax == 1 && bx == 1 ? 1 : 0) | (ay == 1 && by == 1 ? 2 : 0);
                    const std::vector<cd> &va = operand(ax, ay, wa), &vb = operand(bx, by, wb);
                    std::vector<cd> &dst = acc[wp];
                    for (size_t i = 0; i < n; ++i) dst[i] += va[i] * std::conj(vb[i]);
                }
    convolve(acc[1], cx, cy, 0, kern[0]);
    convolve(acc[2], cx, cy, 1, kern[1]);
    convolve(acc[3], cx, cy, 0, kern[0]);
    convolve(acc[3], cx, cy, 1, kern[1]);
    // c(T^H) has the conjugate eigenvalues of c(T), T. Chan's fold of the two-sided generator.
    std::vector<cd> optimal = fold_two_sided(t, RT_CIRCULANT_TCHAN, 0, 0);
    fft2_host(optimal, cx, cy, false);
    std::vector<cd> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = (acc[0][i] + acc[1][i] + acc[2][i] + acc[3][i]) / std::conj(optimal[i]);
    return out;
}

std::vector<cd> kernel_spectrum(const TwoSided &t, int kind, double bwx, double bwy) {
    if (kind == RT_CIRCULANT_SUPEROPTIMAL) return superoptimal_spectrum(t);
    std::vector<cd> c = fold_two_sided(t, kind, bwx, bwy);
    fft2_host(c, t.cx, t.cy, false);
    return c;
}

}  // namespace rt

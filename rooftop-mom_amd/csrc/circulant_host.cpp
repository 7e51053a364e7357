// Host side of the circulant preconditioners that the device folds do not cover: the superoptimal kernel of
// Tyrtyshnikov (+preconditioner/+kernels/generateSuperoptimal.m with +utils/generatePartSpectra.m and
// +utils/generateOptimalOfProduct.m) and the kernels applied to generators that are not even (the XY coupling block of
// the blockwise builder, generateBlockwiseSpectra.m), and every kernel and builder on face grids that are not powers of
// two. These are one-off set-up computations on count_x x count_y grids, in FP64, with an FFT of their own (radix 2,
// Bluestein's chirp transform for other lengths).
#include "circulant_host.hpp"

#include <cmath>

#include "rooftop.hpp"

namespace rt {

namespace {

constexpr double kPi = 3.141592653589793;

// In-place radix-2 FFT of n (a power of two) values at a[0], a[stride], ...: forward sum x(k) exp(-2 pi i k j / n),
// inverse with the opposite sign and no scaling.
void fft_pow2(cd *a, int n, int64_t stride, bool inverse) {
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

// The same for any n: Bluestein's chirp transform, j k = (j^2 + k^2 - (k - j)^2) / 2, turns the DFT into a
// convolution with the chirp exp(+-i pi m^2 / n), done by power-of-two FFTs of length >= 2 n - 1.
void fft1(cd *a, int n, int64_t stride, bool inverse) {
    if ((n & (n - 1)) == 0) {
        fft_pow2(a, n, stride, inverse);
        return;
    }
    const double sgn = inverse ? 1.0 : -1.0;
    int L = 1;
    while (L < 2 * n - 1) L <<= 1;
    std::vector<cd> chirp(n), u(L, 0.0), v(L, 0.0);
    for (int m = 0; m < n; ++m) {
        // m^2 mod 2n keeps the angle small and exact.
        const long long m2 = (static_cast<long long>(m) * m) % (2LL * n);
        chirp[m] = std::polar(1.0, sgn * kPi * static_cast<double>(m2) / n);
    }
    for (int m = 0; m < n; ++m) u[m] = a[m * stride] * chirp[m];
    v[0] = std::conj(chirp[0]);
    for (int m = 1; m < n; ++m) v[m] = v[L - m] = std::conj(chirp[m]);
    fft_pow2(u.data(), L, 1, false);
    fft_pow2(v.data(), L, 1, false);
    for (int i = 0; i < L; ++i) u[i] *= v[i];
    fft_pow2(u.data(), L, 1, true);
    for (int k = 0; k < n; ++k) a[k * stride] = u[k] * chirp[k] / static_cast<double>(L);
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

cd host_at(const std::vector<cd> &g, int rows, int cols, int p, int q) {
    if (p < 0) p = -p;
    if (q < 0) q = -q;
    if (p >= rows || q >= cols) return 0.0;
    return g[static_cast<size_t>(q) * rows + p];
}

TwoSided even_extension(const std::vector<cd> &g, int rows, int cols, int cx, int cy) {
    return two_sided(cx, cy, [&](int p, int q) { return host_at(g, rows, cols, p, q); });
}

TwoSided spatial_block(int block, const HostGenerators &g) {
    const int cx = g.cx, cy = g.cy;
    const double a = g.a, b = g.b;
    auto E = [&](int p, int q) { return host_at(g.e, cx, cy, p, q); };
    return two_sided(cx, cy, [&](int p, int q) -> cd {
        if (block == 1) {
            if (std::abs(p) > cx - 2) return 0.0;
            return host_at(g.mxx, cx - 1, cy, p, q) - (2.0 * E(p, q) - E(p - 1, q) - E(p + 1, q)) / (a * a);
        }
        if (block == 2) {
            if (std::abs(q) > cy - 2) return 0.0;
            return host_at(g.myy, cx, cy - 1, p, q) - (2.0 * E(p, q) - E(p, q - 1) - E(p, q + 1)) / (b * b);
        }
        if (p > cx - 2 || q < -(cy - 2)) return 0.0;
        return -(E(p, q) - E(p, q - 1) - E(p + 1, q) + E(p + 1, q - 1)) / (a * b);
    });
}

std::array<std::vector<cd>, 4> circulant_blocks(const HostGenerators &g, int kernel, int builder, double bwx,
                                                double bwy) {
    const int cx = g.cx, cy = g.cy;
    const size_t n = static_cast<size_t>(cx) * cy;
    std::array<std::vector<cd>, 4> out;   // XX, XY, YX, YY
    auto spectrum = [&](const TwoSided &t) { return kernel_spectrum(t, kernel, bwx, bwy); };
    if (builder == RT_BUILDER_BLOCKWISE) {
        // generateBlockwiseSpectra.m: every spatial block's circulant; YX the transpose of XY (negative frequencies).
        if (g.has_x) out[0] = spectrum(spatial_block(1, g));
        if (g.has_y) out[3] = spectrum(spatial_block(2, g));
        if (g.has_x && g.has_y) {
            out[1] = spectrum(spatial_block(3, g));
            out[2].resize(n);
            for (int fy = 0; fy < cy; ++fy)
                for (int fx = 0; fx < cx; ++fx)
                    out[2][fx + static_cast<size_t>(cx) * fy] = out[1][(cx - fx) % cx + static_cast<size_t>(cx) * ((cy - fy) % cy)];
        }
        return out;
    }
    // Partwise (generatePartwiseSpectra.m) and hybrid (generateHybridSpectra.m): the electric circulant couples the
    // blocks as in generateBlockSpectra.m, Dx = (1 - exp(2 pi i fx / cx)) / a; the diagonal blocks from the magnetic
    // generators' circulants (partwise) or the spatial blocks' (hybrid).
    const std::vector<cd> E = spectrum(even_extension(g.e, cx, cy, cx, cy));
    std::vector<cd> A, B;
    if (builder == RT_BUILDER_HYBRID) {
        if (g.has_x) A = spectrum(spatial_block(1, g));
        if (g.has_y) B = spectrum(spatial_block(2, g));
    } else {
        if (g.has_x) A = spectrum(even_extension(g.mxx, cx - 1, cy, cx, cy));
        if (g.has_y) B = spectrum(even_extension(g.myy, cx, cy - 1, cx, cy));
    }
    if (g.has_x) out[0].resize(n);
    if (g.has_y) out[3].resize(n);
    if (g.has_x && g.has_y) {
        out[1].resize(n);
        out[2].resize(n);
    }
    for (int fy = 0; fy < cy; ++fy)
        for (int fx = 0; fx < cx; ++fx) {
            const size_t i = fx + static_cast<size_t>(cx) * fy;
            const cd dx = (1.0 - std::polar(1.0, 2 * kPi * fx / cx)) / g.a;
            const cd dy = (1.0 - std::polar(1.0, 2 * kPi * fy / cy)) / g.b;
            const bool given = builder == RT_BUILDER_HYBRID;
            if (g.has_x) out[0][i] = given ? A[i] : A[i] - std::norm(dx) * E[i];
            if (g.has_y) out[3][i] = given ? B[i] : B[i] - std::norm(dy) * E[i];
            if (g.has_x && g.has_y) {
                out[1][i] = -dx * std::conj(dy) * E[i];
                out[2][i] = -std::conj(dx) * dy * E[i];
            }
        }
    return out;
}

bool invert_blocks(std::array<std::vector<cd>, 4> &blk, bool has_x, bool has_y) {
    const size_t n = has_x ? blk[0].size() : blk[3].size();
    bool ok = true;
    auto finite = [](cd v) { return std::isfinite(v.real()) && std::isfinite(v.imag()); };
    for (size_t f = 0; f < n; ++f) {
        if (has_x && has_y) {
            const cd xx = blk[0][f], xy = blk[1][f], yx = blk[2][f], yy = blk[3][f];
            const cd det = xx * yy - xy * yx;
            blk[0][f] = yy / det;
            blk[1][f] = -xy / det;
            blk[2][f] = -yx / det;
            blk[3][f] = xx / det;
            ok = ok && det != cd(0.0) && finite(blk[0][f]) && finite(blk[1][f]) && finite(blk[2][f]) && finite(blk[3][f]);
        } else {
            cd &v = has_x ? blk[0][f] : blk[3][f];
            const cd det = v;
            v = 1.0 / det;
            ok = ok && det != cd(0.0) && finite(v);
        }
    }
    return ok;
}

}  // namespace rt

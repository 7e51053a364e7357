// Host-side circulant approximations on count_x x count_y grids, FP64 (circulant_host.cpp).
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <vector>

namespace rt {

using cd = std::complex<double>;

// A two-sided generator: the offsets -(cx-1) .. cx-1 by -(cy-1) .. cy-1, column-major, offset 0 at (cx-1, cy-1).
struct TwoSided {
    int cx = 0, cy = 0;
    std::vector<cd> v;
    cd at(int p, int q) const;   // zero outside the offsets held
};

template <class F>
TwoSided two_sided(int cx, int cy, F f) {
    TwoSided t;
    t.cx = cx;
    t.cy = cy;
    t.v.resize(static_cast<size_t>(2 * cx - 1) * (2 * cy - 1));
    for (int q = -(cy - 1); q < cy; ++q)
        for (int p = -(cx - 1); p < cx; ++p) t.v[(p + cx - 1) + static_cast<size_t>(2 * cx - 1) * (q + cy - 1)] = f(p, q);
    return t;
}

// The generators on the host (after the modifier): mXX (cx-1) x cy, mYY cx x (cy-1), e cx x cy, column-major.
struct HostGenerators {
    std::vector<cd> mxx, myy, e;
    int cx = 0, cy = 0;
    double a = 0, b = 0;
    bool has_x = false, has_y = false;
};

// The value of an even generator (rows x cols) at the offset (p, q); zero beyond it.
cd host_at(const std::vector<cd> &g, int rows, int cols, int p, int q);
// The even two-sided extension of a one-sided generator padded to cx x cy (extendEvenly.m, padWithZeros.m).
TwoSided even_extension(const std::vector<cd> &g, int rows, int cols, int cx, int cy);
// The two-sided spatial blocks (generateSpatialBlocks.m:19-61): 1 XX, 2 YY, 3 XY.
TwoSided spatial_block(int block, const HostGenerators &g);
// The 2x2 block spectra (XX, XY, YX, YY; [fx + cx fy], empty for a missing block) of the circulant approximation of
// an RT_CIRCULANT_* kernel with an RT_BUILDER_* builder on the cx x cy period.
std::array<std::vector<cd>, 4> circulant_blocks(const HostGenerators &g, int kernel, int builder, double bwx,
                                                double bwy);
// In place, frequency by frequency, the inverse blocks (@Spectra/invert.m:12-20); false when singular (a zero
// determinant or a non-finite inverse).
bool invert_blocks(std::array<std::vector<cd>, 4> &blocks, bool has_x, bool has_y);

// In-place 2D FFT of a[x + cx y] (forward: exp(-2 pi i ...); inverse scaled by 1 / (cx cy)).
void fft2_host(std::vector<cd> &a, int cx, int cy, bool inverse);
// The fold of a two-sided generator onto the period by the window of an RT_CIRCULANT_* kernel (foldGenerator.m:27-61),
// [x + cx y].
std::vector<cd> fold_two_sided(const TwoSided &t, int kind, double bwx, double bwy);
// The eigenvalues (fft2 of the generator) of the superoptimal circulant c(T T^H) / c(T^H) (generateSuperoptimal.m).
std::vector<cd> superoptimal_spectrum(const TwoSided &t);
// The eigenvalues of an RT_CIRCULANT_* kernel's circulant of t, [fx + cx fy].
std::vector<cd> kernel_spectrum(const TwoSided &t, int kind, double bwx, double bwy);

}  // namespace rt

// Host-side circulant approximations on count_x x count_y grids (powers of two), FP64 (circulant_host.cpp).
#pragma once

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

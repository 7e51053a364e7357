// Host side of the FFT plans: radix schedule and the output order of the forward transform.
#include "fft.hpp"

namespace rt {

FftPlan make_plan(int M) {
    FftPlan p;
    if (M < 1 || (M & (M - 1)))
        fail(RT_EINVAL, "rooftop:fft", "transform length %d is not a power of two", M);
    p.M = M;
    while ((1 << p.log2M) < M) ++p.log2M;
    int rem = p.log2M;
    int len = p.log2M;
    // Radix 4 stages, with one radix 2 stage first when log2 M is odd.
    if (rem & 1) {
        p.radix[p.nstages] = 2;
        p.log2radix[p.nstages] = 1;
        p.log2len[p.nstages] = len;
        ++p.nstages;
        len -= 1;
        rem -= 1;
    }
    while (rem > 0) {
        if (p.nstages >= kMaxStages) fail(RT_EINVAL, "rooftop:fft", "transform length %d is too long", M);
        p.radix[p.nstages] = 4;
        p.log2radix[p.nstages] = 2;
        p.log2len[p.nstages] = len;
        ++p.nstages;
        len -= 2;
        rem -= 2;
    }
    return p;
}

// Position j * (M / r0) + q of the forward output holds frequency r0 * k_sub(q) + j, where k_sub is the order of the
// transform of length M / r0 with the remaining stages (decimation in frequency).
std::vector<int> output_frequencies(const FftPlan &p) {
    std::vector<int> freq(1, 0);
    for (int s = p.nstages - 1; s >= 0; --s) {
        const int r = p.radix[s];
        const int sub = static_cast<int>(freq.size());
        std::vector<int> next(static_cast<size_t>(sub) * r);
        for (int j = 0; j < r; ++j)
            for (int q = 0; q < sub; ++q) next[static_cast<size_t>(j) * sub + q] = r * freq[q] + j;
        freq.swap(next);
    }
    return freq;
}

}  // namespace rt

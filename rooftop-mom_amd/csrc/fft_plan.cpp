// Host side of the team FFT plans: round schedule, output order and twiddle tables.
#include <cmath>

#include "fft.hpp"

namespace rt {

FftPlan make_plan(int M, int maxbits) {
    FftPlan p;
    if (M < 1 || (M & (M - 1)))
        fail(RT_EINVAL, "rooftop:fft", "transform length %d is not a power of two", M);
    if (maxbits < 1 || maxbits > 4) fail(RT_EINVAL, "rooftop:fft", "radices up to 16 are supported");
    p.M = M;
    p.bits = maxbits;
    while ((1 << p.log2M) < M) ++p.log2M;
    const int m = p.log2M;
    if (m == 0) {
        p.K = 1;   // one round of radix 1
        p.twcount = 1;
        return p;
    }
    p.K = (m + maxbits - 1) / maxbits;
    if (p.K > kMaxRounds) fail(RT_EINVAL, "rooftop:fft", "transform length %d is too long", M);
    const int base = m / p.K, extra = m % p.K;
    int ll = m, off = 0;
    for (int s = 0; s < p.K; ++s) {
        p.lr[s] = base + (s < extra ? 1 : 0);
        p.ll[s] = ll;
        p.twoff[s] = off;
        off += 1 << ll;
        ll -= p.lr[s];
        if (p.lr[s] > p.log2E) p.log2E = p.lr[s];
    }
    p.twcount = off;
    return p;
}

// Position pos = sum_s q_s S_s (S_s = L_s / r_s, digits q_s < r_s) holds frequency q_0 + r_0 (q_1 + r_1 (q_2 + ...)):
// round s splits a block of length L into r sub-blocks of length S, sub-block q taking the frequencies = q mod r.
std::vector<int> output_frequencies(const FftPlan &p) {
    std::vector<int> freq(p.M);
    for (int pos = 0; pos < p.M; ++pos) {
        int k = 0, mul = 1;
        for (int s = 0; s < p.K; ++s) {
            const int ls = p.ll[s] - p.lr[s];
            const int q = (pos >> ls) & ((1 << p.lr[s]) - 1);
            k += q * mul;
            mul <<= p.lr[s];
        }
        freq[pos] = k;
    }
    return freq;
}

std::vector<c128> plan_twiddles(const FftPlan &p) {
    std::vector<c128> tw(p.twcount, c128{1.0, 0.0});
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int s = 0; s < p.K && p.log2M > 0; ++s) {
        const int L = 1 << p.ll[s], r = 1 << p.lr[s], S = L / r;
        for (int n = 0; n < S; ++n)
            for (int q = 0; q < r; ++q) {
                const long double ang = -2.0L * pi * static_cast<long double>(q * n) / static_cast<long double>(L);
                tw[p.twoff[s] + n * r + q] = {static_cast<double>(std::cos(ang)), static_cast<double>(std::sin(ang))};
            }
    }
    return tw;
}

}  // namespace rt

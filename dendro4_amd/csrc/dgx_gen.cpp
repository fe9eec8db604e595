// dgx_gen.cpp -- deterministic synthetic point clouds for the benchmark
// configurations.  Counter-based SplitMix64: the i-th 64-bit draw of stream s
// under seed k is mix64(k * G1 + s * G2 + (i + 1) * G), so every coordinate is
// a pure function of (seed, stream, counter) and the arrays are reproducible
// on any host with the same libm.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/dendro_mi355x.h"

namespace {

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(seed * 0xD1B54A32D192ED03ull + stream * 0xABC98388FB8FAC03ull + (i + 1) * 0x9E3779B97F4A7C15ull);
}
inline double unif(uint64_t seed, uint64_t stream, uint64_t i) {  // [0,1)
    return (double)(draw(seed, stream, i) >> 11) * (1.0 / 9007199254740992.0);
}
inline double normal(uint64_t seed, uint64_t stream, uint64_t i) {  // Box-Muller, cosine branch
    double u1 = 1.0 - unif(seed, stream, 2 * i);                     // (0,1]
    double u2 = unif(seed, stream, 2 * i + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
inline double clip01(double v) {
    const double top = std::nextafter(1.0, 0.0);
    return v < 0.0 ? 0.0 : (v > top ? top : v);
}

// Runs body(lo, hi) over [0, n) on up to 16 host threads (every value is a
// pure function of its index, so the split does not change the result).
template <typename F>
void parallel_range(size_t n, F body) {
    const size_t grain = size_t(1) << 16;
    unsigned t = std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    if (n < 2 * grain || t < 2) {
        body(0, n);
        return;
    }
    t = (unsigned)std::min<size_t>(t, n / grain);
    std::vector<std::thread> th;
    for (unsigned k = 0; k < t; ++k) th.emplace_back(body, n * k / t, n * (k + 1) / t);
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" int dgx_gen_points(int dist, uint64_t seed, size_t n, int n_gauss, double lo, double hi, double sig_lo,
                              double sig_hi, int dirichlet, double *out) {
    if (!out && n) return -1;
    if (dist == 0) {
        parallel_range(3 * n, [&](size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; ++i) out[i] = unif(seed, 0, i);
        });
        return 0;
    }
    if (dist == 1) {
        parallel_range(3 * n, [&](size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; ++i) out[i] = clip01(std::exp(-1.0 + 0.5 * normal(seed, 0, i)));
        });
        return 0;
    }
    if (dist != 2 || n_gauss <= 0) return -1;
    const int K = n_gauss;
    std::vector<double> cen(3 * K), sig(K), cdf(K);
    for (int k = 0; k < 3 * K; ++k) cen[k] = lo + (hi - lo) * unif(seed, 1, k);
    for (int k = 0; k < K; ++k) {
        if (sig_hi > 0.0) {
            double a = std::log(sig_lo), b = std::log(sig_hi);
            sig[k] = std::exp(a + (b - a) * unif(seed, 2, k));
        } else {
            sig[k] = sig_lo;
        }
    }
    double tot = 0.0;
    for (int k = 0; k < K; ++k) {
        double w = dirichlet ? -std::log(1.0 - unif(seed, 3, k)) : 1.0;  // Dirichlet(1) = normalised Exp(1)
        tot += w;
        cdf[k] = tot;
    }
    for (int k = 0; k < K; ++k) cdf[k] /= tot;
    parallel_range(n, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            double u = unif(seed, 4, i);
            int k = 0;
            while (k < K - 1 && u >= cdf[k]) ++k;
            for (int a = 0; a < 3; ++a)
                out[3 * i + a] = clip01(cen[3 * k + a] + sig[k] * normal(seed, 5, 3 * i + a));
        }
    });
    return 0;
}

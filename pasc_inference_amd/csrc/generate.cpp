// Deterministic synthetic inputs for the FEM-H1 K-means path.
//
// The reference reads its signals from PETSc binary files produced by separate generators
// (examples/test_signal1D/test_signal1D_generate.cpp, Matlab/signal1D_comparison/generate.m).
// These generators define the workloads of this library exactly, so that the CPU oracle and the
// device path see bit-identical inputs:
//   uniform  : SplitMix64 -> (z >> 11) * 2^-53 in [0, 1)
//   normal   : Box-Muller on (1 - u1, u2), cosine branch only
//   Markov   : initial state floor(u*K); at every t > 0, with probability p the state moves to
//              one of the other K-1 states chosen uniformly; x_t = level[state] + sigma*normal
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/pasc_hip.h"
#include "pasc_internal.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double normal() {
        const double u1 = 1.0 - uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925 * u2);
    }
    int below(int n) {  // uniform integer in [0, n)
        int v = (int)(uniform() * (double)n);
        return v >= n ? n - 1 : v;
    }
};

}  // namespace

extern "C" int pasc_generate_markov(uint64_t seed, int64_t T, int32_t K, const double *levels,
                                    double p_switch, double sigma, double *x, int32_t *labels) {
    if (T < 0 || K < 1 || !levels || !x) return pasc::fail(PASC_ERR_ARG, "generate_markov: bad args");
    SplitMix64 rng(seed);
    int state = rng.below(K);
    for (int64_t t = 0; t < T; t++) {
        if (t > 0 && K > 1) {
            if (rng.uniform() < p_switch) state = (state + 1 + rng.below(K - 1)) % K;
        }
        x[t] = levels[state] + sigma * rng.normal();
        if (labels) labels[t] = state;
    }
    return PASC_OK;
}

extern "C" int pasc_generate_uniform(uint64_t seed, int64_t n, double *out) {
    if (n < 0 || (!out && n > 0)) return pasc::fail(PASC_ERR_ARG, "generate_uniform: bad args");
    SplitMix64 rng(seed);
    for (int64_t i = 0; i < n; i++) out[i] = rng.uniform();
    return PASC_OK;
}

extern "C" int pasc_generate_mask(uint64_t seed, int64_t n, double fraction, uint8_t *mask) {
    if (n < 0 || !mask) return pasc::fail(PASC_ERR_ARG, "generate_mask: bad args");
    SplitMix64 rng(seed ^ 0x6A09E667F3BCC909ull);
    for (int64_t i = 0; i < n; i++) mask[i] = rng.uniform() < fraction ? 1 : 0;
    return PASC_OK;
}

// Image: cluster 0 is the background; each of n_rect rectangles [c0,c1] x [r0,r1] (corners
// drawn uniformly, sorted) is painted with a label drawn uniformly from 1..K-1.  Colours are
// uniform in [0,1]^3 per cluster; pixels are row-major, channels innermost.
extern "C" int pasc_generate_rect_image(uint64_t seed, int32_t width, int32_t height, int32_t K,
                                        int32_t n_rect, double sigma, double *x, double *colours,
                                        int32_t *labels) {
    if (width < 1 || height < 1 || K < 1 || !x || !colours || !labels)
        return pasc::fail(PASC_ERR_ARG, "generate_rect_image: bad args");
    SplitMix64 rng(seed);
    for (int k = 0; k < K; k++)
        for (int c = 0; c < 3; c++) colours[k * 3 + c] = rng.uniform();
    const int64_t R = (int64_t)width * height;
    for (int64_t i = 0; i < R; i++) labels[i] = 0;
    for (int q = 0; q < n_rect; q++) {
        int c0 = rng.below(width), c1 = rng.below(width);
        int r0 = rng.below(height), r1 = rng.below(height);
        if (c0 > c1) { int s = c0; c0 = c1; c1 = s; }
        if (r0 > r1) { int s = r0; r0 = r1; r1 = s; }
        const int lab = K > 1 ? 1 + rng.below(K - 1) : 0;
        for (int r = r0; r <= r1; r++)
            for (int c = c0; c <= c1; c++) labels[(int64_t)r * width + c] = lab;
    }
    for (int64_t i = 0; i < R; i++)
        for (int c = 0; c < 3; c++) x[i * 3 + c] = colours[labels[i] * 3 + c] + sigma * rng.normal();
    return PASC_OK;
}

extern "C" int pasc_generate_batch_params(uint64_t seed, int32_t n_series, double lg_tmin,
                                          double lg_tmax, int32_t kmin, int32_t kmax,
                                          double level_max, double pmin, double pmax,
                                          double smin, double smax, int64_t *T, int32_t *K,
                                          double *levels, double *p_switch, double *sigma,
                                          uint64_t *series_seed) {
    if (n_series < 0 || kmin < 1 || kmax < kmin || kmax > PASC_MAX_K || !T || !K || !levels ||
        !p_switch || !sigma || !series_seed)
        return pasc::fail(PASC_ERR_ARG, "generate_batch_params: bad args");
    SplitMix64 rng(seed);
    for (int i = 0; i < n_series; i++) {
        const double lg = lg_tmin + (lg_tmax - lg_tmin) * rng.uniform();
        T[i] = (int64_t)std::llround(std::pow(10.0, lg));
        K[i] = kmin + rng.below(kmax - kmin + 1);
        for (int k = 0; k < PASC_MAX_K; k++) levels[(int64_t)i * PASC_MAX_K + k] = 0.0;
        for (int k = 0; k < K[i]; k++) levels[(int64_t)i * PASC_MAX_K + k] = level_max * rng.uniform();
        p_switch[i] = pmin + (pmax - pmin) * rng.uniform();
        sigma[i] = smin + (smax - smin) * rng.uniform();
        series_seed[i] = rng.next();
    }
    return PASC_OK;
}

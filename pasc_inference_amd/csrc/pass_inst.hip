// Instantiations of the fused 1-D SPG pass for one storage type and block size.  This file is
// compiled once per (PASS_S, PASS_NT) pair so that the instantiations build in parallel.
#include "spg_kernels.h"

#ifndef PASS_S
#error "PASS_S (float or double) must be defined"
#endif
#ifndef PASS_NT
#error "PASS_NT (threads per block) must be defined"
#endif

#define PASC_CAT3_(a, b, c) a##b##c
#define PASC_CAT3(a, b, c) PASC_CAT3_(a, b, c)
#define PASS_FN PASC_CAT3(launch_pass_1d_, PASS_S, PASC_CAT3(_, PASS_NT, _))

namespace pasc {

template <int K, int CH>
static int launch_kc(const PassArgs &a, int grid, bool pf, hipStream_t s) {
    if (pf)
        hipLaunchKernelGGL((spg_pass_1d<PASS_S, PASS_S, K, CH, PASS_NT, true>), dim3(grid),
                           dim3(PASS_NT), 0, s, a);
    else
        hipLaunchKernelGGL((spg_pass_1d<PASS_S, PASS_S, K, CH, PASS_NT, false>), dim3(grid),
                           dim3(PASS_NT), 0, s, a);
    return 0;
}

template <int K>
static int launch_k(const PassArgs &a, int grid, int chunk, bool pf, hipStream_t s) {
    constexpr bool f32 = sizeof(PASS_S) == 4;
    constexpr int CH_BIG = f32 ? 8 : 4;
    constexpr int CH_SMALL = f32 ? 4 : 2;
    if (chunk == CH_BIG) return launch_kc<K, CH_BIG>(a, grid, pf, s);
    if (chunk == CH_SMALL) return launch_kc<K, CH_SMALL>(a, grid, pf, s);
    return fail(PASC_ERR_ARG, "no fused pass for chunk %d", chunk);
}

int PASS_FN(int K, const PassArgs &a, int grid, int chunk, int pf, hipStream_t s) {
    switch (K) {
        case 1: return launch_k<1>(a, grid, chunk, pf != 0, s);
        case 2: return launch_k<2>(a, grid, chunk, pf != 0, s);
        case 3: return launch_k<3>(a, grid, chunk, pf != 0, s);
        case 4: return launch_k<4>(a, grid, chunk, pf != 0, s);
        case 5: return launch_k<5>(a, grid, chunk, pf != 0, s);
        case 6: return launch_k<6>(a, grid, chunk, pf != 0, s);
        case 7: return launch_k<7>(a, grid, chunk, pf != 0, s);
        case 8: return launch_k<8>(a, grid, chunk, pf != 0, s);
    }
    return fail(PASC_ERR_ARG, "unsupported K %d", K);
}

}  // namespace pasc

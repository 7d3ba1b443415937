// Instantiations of the fused image pass for one storage type (PASS_S).
#include "img_kernels.h"

#ifndef PASS_S
#error "PASS_S (float or double) must be defined"
#endif

#define PASC_CAT2_(a, b) a##b
#define PASC_CAT2(a, b) PASC_CAT2_(a, b)
#define IMG_FN PASC_CAT2(launch_pass_img_, PASS_S)

namespace pasc {

// pd: raw image rows in flight ahead of their use besides the next one (1..3)
template <int K>
static int launch_k(const ImgArgs &a, int grid, int pd, hipStream_t s) {
    switch (pd) {
        case 1: hipLaunchKernelGGL((spg_pass_img<PASS_S, PASS_S, K, 1>), dim3(grid), dim3(WAVE), 0, s, a); return 0;
        case 2: hipLaunchKernelGGL((spg_pass_img<PASS_S, PASS_S, K, 2>), dim3(grid), dim3(WAVE), 0, s, a); return 0;
        case 3: hipLaunchKernelGGL((spg_pass_img<PASS_S, PASS_S, K, 3>), dim3(grid), dim3(WAVE), 0, s, a); return 0;
    }
    return fail(PASC_ERR_ARG, "image pass: prefetch depth %d (1..3)", pd);
}

int IMG_FN(int K, const ImgArgs &a, int grid, int pd, hipStream_t s) {
    switch (K) {
        case 1: return launch_k<1>(a, grid, pd, s);
        case 2: return launch_k<2>(a, grid, pd, s);
        case 3: return launch_k<3>(a, grid, pd, s);
        case 4: return launch_k<4>(a, grid, pd, s);
        case 5: return launch_k<5>(a, grid, pd, s);
        case 6: return launch_k<6>(a, grid, pd, s);
        case 7: return launch_k<7>(a, grid, pd, s);
        case 8: return launch_k<8>(a, grid, pd, s);
    }
    return fail(PASC_ERR_ARG, "unsupported K %d", K);
}

}  // namespace pasc

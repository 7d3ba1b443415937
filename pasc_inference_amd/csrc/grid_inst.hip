// Instantiations of the grid-graph SPG kernels for one storage type (PASS_S).
#include "spg_kernels.h"

#ifndef PASS_S
#error "PASS_S (float or double) must be defined"
#endif

#define PASC_CAT2_(a, b) a##b
#define PASC_CAT2(a, b) PASC_CAT2_(a, b)
#define GRID_FN PASC_CAT2(launch_pass_grid_, PASS_S)

namespace pasc {

template <int K, bool IMG>
static void launch_kg(const GridArgs &a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((spg_pass_grid_dir<PASS_S, PASS_S, K, 256, IMG>), dim3(grid), dim3(256), 0, s, a);
    hipLaunchKernelGGL((spg_pass_grid_ad<PASS_S, PASS_S, K, 256, IMG>), dim3(grid), dim3(256), 0, s, a);
}

// images (T == 1 on an implicit 4-neighbour grid) take the 32-bit neighbour walk
template <int K>
static void launch_k(const GridArgs &a, int grid, hipStream_t s) {
    const bool img = a.geo.T == 1 && a.geo.width > 0 && !a.geo.ptr && a.geo.R < (int64_t(1) << 31);
    if (img)
        launch_kg<K, true>(a, grid, s);
    else
        launch_kg<K, false>(a, grid, s);
}

int GRID_FN(int K, const GridArgs &a, int grid, hipStream_t s) {
    switch (K) {
        case 1: launch_k<1>(a, grid, s); return 0;
        case 2: launch_k<2>(a, grid, s); return 0;
        case 3: launch_k<3>(a, grid, s); return 0;
        case 4: launch_k<4>(a, grid, s); return 0;
        case 5: launch_k<5>(a, grid, s); return 0;
        case 6: launch_k<6>(a, grid, s); return 0;
        case 7: launch_k<7>(a, grid, s); return 0;
        case 8: launch_k<8>(a, grid, s); return 0;
    }
    return fail(PASC_ERR_ARG, "unsupported K %d", K);
}

}  // namespace pasc

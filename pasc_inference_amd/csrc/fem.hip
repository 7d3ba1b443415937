// FEM coarsening of the time axis on the device: reduction of fine-grid rows (gamma or the
// k-means residual (x_t - theta_k)^2) to a coarse grid and prolongation back.
//
// Fem (piecewise constant, fem.cpp:28-182): diff = T1/T2, coarse cell t2 holds the fine steps
// round(t2*diff) .. round((t2+1)*diff)-1; reduce sums over the cell, prolongate copies.
// FemHat (hat functions, femhat.cpp:26-217): diff = (T1-1)/(T2-1); reduce is the hat-weighted
// sum as the reference writes it (left loop from floor(left), right loop below
// floor(diff*(T2-1)) -- femhat.h:154-180 with one rank), prolongate the linear interpolation
// (a right neighbour index of T2, whose weight is 0, reads the last coarse value).
// Rows are k-major with strides ld1 (fine) / ld2 (coarse), as the engine stores gamma.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "device_common.h"
#include "pasc_internal.h"

namespace pasc {

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(PASC_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                              \
    } while (0)

struct FemArgs {
    FemGeom g;
    int32_t K;
    const double *in;       // fine rows (reduce of values) or coarse rows (prolongate, FP64)
    const float *in32;      // coarse rows in FP32 (prolongate of an FP32 iterate) or null
    double *out;
    // residual mode of the reduction: fine values (x_t - theta_k)^2 summed over channels
    const double *x;        // data rows [xdim][ld1] or null
    int32_t xdim;
    const uint8_t *heldout; // fine mask: held-out steps carry no residual
    double theta[PASC_MAX_K * 3];
    double scale;           // multiplies every reduced value
};

__device__ __forceinline__ void fine_values(const FemArgs &A, int64_t t1, double (&v)[PASC_MAX_K]) {
    if (A.x) {
        const bool obs = !(A.heldout && A.heldout[t1]);
        for (int k = 0; k < A.K; k++) {
            double q = 0.0;
            if (obs)
                for (int c = 0; c < A.xdim; c++) {
                    const double dx = A.x[c * A.g.ld1 + t1] - A.theta[k * A.xdim + c];
                    q += dx * dx;
                }
            v[k] = q;
        }
    } else {
        for (int k = 0; k < A.K; k++) v[k] = A.in[k * A.g.ld1 + t1];
    }
}

__global__ __launch_bounds__(256) void fem_reduce_kernel(FemArgs A) {
    const int64_t t2 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t2 >= A.g.T2) return;
    double s[PASC_MAX_K], v[PASC_MAX_K];
    for (int k = 0; k < A.K; k++) s[k] = 0.0;
    const double diff = A.g.diff;
    if (A.g.type == PASC_FEM_CONSTANT) {
        const int64_t a = (int64_t)round((double)t2 * diff);
        const int64_t b = (int64_t)round((double)(t2 + 1) * diff);
        for (int64_t i = a; i < b && i < A.g.T1; i++) {
            fine_values(A, i, v);
            for (int k = 0; k < A.K; k++) s[k] += v[k];
        }
    } else {
        const double center = (double)t2 * diff;
        const double left = (double)(t2 - 1) * diff;
        const double right = (double)(t2 + 1) * diff;
        int64_t t1 = (int64_t)floor(left);
        for (; (double)t1 <= center; t1++) {
            if (t1 < 0 || t1 >= A.g.T1) continue;
            const double phi = ((double)t1 - left) / (center - left);
            fine_values(A, t1, v);
            for (int k = 0; k < A.K; k++) s[k] += phi * v[k];
        }
        for (; (double)t1 < right; t1++) {
            if (t1 >= A.g.right_idx || t1 < 0) continue;
            const double phi = ((double)t1 - right) / (center - right);
            fine_values(A, t1, v);
            for (int k = 0; k < A.K; k++) s[k] += phi * v[k];
        }
    }
    for (int k = 0; k < A.K; k++) A.out[k * A.g.ld2 + t2] = A.scale * s[k];
}

__device__ __forceinline__ double coarse_at(const FemArgs &A, int k, int64_t t2) {
    return A.in32 ? (double)A.in32[k * A.g.ld2 + t2] : A.in[k * A.g.ld2 + t2];
}

__global__ __launch_bounds__(256) void fem_prolongate_kernel(FemArgs A) {
    const int64_t t1 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t1 >= A.g.T1) return;
    const double diff = A.g.diff;
    const int64_t T2 = A.g.T2;
    if (A.g.type == PASC_FEM_CONSTANT) {
        // the cell holding t1: cells are contiguous and cover [0, T1)
        int64_t t2 = (int64_t)floor((double)t1 / diff);
        if (t2 > T2 - 1) t2 = T2 - 1;
        if (t2 < 0) t2 = 0;
        while (t2 > 0 && (int64_t)round((double)t2 * diff) > t1) t2--;
        while (t2 < T2 - 1 && (int64_t)round((double)(t2 + 1) * diff) <= t1) t2++;
        for (int k = 0; k < A.K; k++) A.out[k * A.g.ld1 + t1] = coarse_at(A, k, t2);
    } else {
        const int64_t tl = (int64_t)floor((double)t1 / diff);
        const int64_t tr = tl + 1;
        const double t1l = (double)tl * diff, t1r = (double)tr * diff;
        const int64_t il = tl < T2 ? tl : T2 - 1;
        const int64_t ir = tr < T2 ? tr : T2 - 1;
        const double wr = ((double)t1 - t1l) / (t1r - t1l);
        const double wl = ((double)t1 - t1r) / (t1l - t1r);
        for (int k = 0; k < A.K; k++)
            A.out[k * A.g.ld1 + t1] = wr * coarse_at(A, k, ir) + wl * coarse_at(A, k, il);
    }
}

FemGeom fem_geom(int type, int64_t T1, int64_t T2, int64_t ld1, int64_t ld2) {
    FemGeom g;
    g.type = type;
    g.T1 = T1;
    g.T2 = T2;
    g.ld1 = ld1;
    g.ld2 = ld2;
    if (type == PASC_FEM_CONSTANT) {
        g.diff = (double)T1 / (double)T2;                         // fem.cpp:16
        g.right_idx = T1;
    } else {
        g.diff = (double)(T1 - 1) / (double)(T2 - 1);             // femhat.cpp:15
        g.right_idx = (int64_t)std::floor(g.diff * (double)(T2 - 1));  // femhat.h:179
    }
    return g;
}

int fem_check_geom(int type, int64_t T1, int64_t T2) {
    if (type != PASC_FEM_CONSTANT && type != PASC_FEM_HAT)
        return fail(PASC_ERR_ARG, "fem: unknown type %d", type);
    if (T1 < 1 || T2 < 1 || T2 > T1) return fail(PASC_ERR_ARG, "fem: need 1 <= T2 <= T1");
    if (type == PASC_FEM_HAT && T2 < 2) return fail(PASC_ERR_ARG, "fem: hat functions need T2 >= 2");
    return PASC_OK;
}

int fem_reduce_rows(void *stream, const FemGeom &g, int K, const double *in, double *out,
                    double scale) {
    FemArgs A;
    std::memset(&A, 0, sizeof(A));
    A.g = g;
    A.K = K;
    A.in = in;
    A.out = out;
    A.scale = scale;
    const unsigned blocks = (unsigned)((g.T2 + 255) / 256);
    hipLaunchKernelGGL(fem_reduce_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, A);
    HIP_TRY(hipGetLastError());
    return PASC_OK;
}

int fem_residual_rows(void *stream, const FemGeom &g, int K, const double *x, int xdim,
                      const double *theta, const uint8_t *heldout, double *out, double scale) {
    FemArgs A;
    std::memset(&A, 0, sizeof(A));
    A.g = g;
    A.K = K;
    A.x = x;
    A.xdim = xdim;
    A.heldout = heldout;
    for (int i = 0; i < K * xdim; i++) A.theta[i] = theta[i];
    A.out = out;
    A.scale = scale;
    const unsigned blocks = (unsigned)((g.T2 + 255) / 256);
    hipLaunchKernelGGL(fem_reduce_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, A);
    HIP_TRY(hipGetLastError());
    return PASC_OK;
}

int fem_prolongate_rows(void *stream, const FemGeom &g, int K, const void *in, bool in_f32,
                        double *out) {
    FemArgs A;
    std::memset(&A, 0, sizeof(A));
    A.g = g;
    A.K = K;
    if (in_f32) A.in32 = (const float *)in;
    else A.in = (const double *)in;
    A.out = out;
    const unsigned blocks = (unsigned)((g.T1 + 255) / 256);
    hipLaunchKernelGGL(fem_prolongate_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, A);
    HIP_TRY(hipGetLastError());
    return PASC_OK;
}

}  // namespace pasc

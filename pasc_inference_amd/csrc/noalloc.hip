// The theta step without device allocations (pasc_theta_update_noalloc).
//
// pasc_theta_update writes one row of partial sums per 256-step block into the engine's scratch
// buffer, which grows by hipMalloc when a larger problem needs more, and sums the rows on the
// host.  Callers that cannot allocate (code under HIP graph capture, tight serving loops) use
// this variant instead: the partial sums go to `scratch`, a file-scope device array of
// NOALLOC_SCRATCH_DOUBLES doubles (4 MiB on every device the library runs on, held until the
// process ends), and a one-workgroup kernel sums them in block order -- the order of the host
// loop of the allocating path, so the two give bit-identical theta.  Only nv = K (1 + xdim)
// totals come back to the host, into a fixed array on the stack.
//
// Every call on a device shares that device's array: calls that overlap on one device race.
// The array is addressed through the pointer hipGetSymbolAddress returned for the device
// (noalloc_init, run by pasc_ctx_create with the device current); the array's name is a
// host-side stand-in that no host code may pass to a kernel.
#include <hip/hip_runtime.h>

#include <mutex>

#include "util_kernels.h"

namespace pasc {

constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)1 << 19;  // 4 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on `device` (the caller made it current).  Once per device.
int noalloc_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return fail(PASC_ERR_ARG, "noalloc: device %d beyond %d", device, NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return PASC_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(PASC_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                    hipGetErrorString(e));
    g_scratch_ptr[device] = (double *)ptr;
    return PASC_OK;
}

// Thread v sums the partials of value v in block order (as the host loop of rows_theta).
__global__ __launch_bounds__(64) void theta_reduce_ordered(const double *part, int blocks, int nv,
                                                           double *out) {
    const int v = threadIdx.x;
    if (v >= nv) return;
    double a = 0.0;
    for (int b = 0; b < blocks; b++) a += part[(int64_t)b * nv + v];
    out[v] = a;
}

// Largest number of steps the static scratch serves for K clusters and xdim channels.
int64_t noalloc_theta_max_T(int K, int xdim) {
    const int64_t nv = (int64_t)K * (1 + xdim);
    const int64_t blocks = (int64_t)NOALLOC_SCRATCH_DOUBLES / nv - 1;
    return blocks * 256;
}

// The theta step of rows_theta with the partials in the static scratch: gamma rows g (type f32
// or FP64) described by d, data rows x of stride d.ld; theta_out receives theta[k*xdim + c].
int rows_theta_noalloc(Engine &e, const void *g, bool f32, const InstDesc &d, const double *x,
                       double *theta_out) {
    const int device = e.ctx->device;
    if (device < 0 || device >= NOALLOC_MAX_DEVICES || !g_scratch_ptr[device])
        return fail(PASC_ERR_STATE, "theta_update_noalloc: the static scratch of device %d was "
                                    "not set up (pasc_ctx_create does that)", device);
    const int nv = d.K * (1 + d.xdim);
    const int64_t blocks = (d.T + 255) / 256;
    if (d.T > noalloc_theta_max_T(d.K, d.xdim))
        return fail(PASC_ERR_ARG, "theta_update_noalloc: T*R = %lld exceeds %lld, the most the "
                                  "static scratch (%zu doubles) serves at K = %d, xdim = %d",
                    (long long)d.T, (long long)noalloc_theta_max_T(d.K, d.xdim),
                    NOALLOC_SCRATCH_DOUBLES, d.K, d.xdim);
    double *part = g_scratch_ptr[device];
    double *tot_dev = part + (size_t)blocks * nv;
    hipStream_t s = (hipStream_t)e.ctx->stream;
    if (blocks > 0) {
        if (f32)
            hipLaunchKernelGGL((theta_partials<float>), dim3((unsigned)blocks), dim3(256), 0, s,
                               (const float *)g, x, d, part);
        else
            hipLaunchKernelGGL((theta_partials<double>), dim3((unsigned)blocks), dim3(256), 0, s,
                               (const double *)g, x, d, part);
    }
    hipLaunchKernelGGL(theta_reduce_ordered, dim3(1), dim3(64), 0, s, (const double *)part,
                       (int)blocks, nv, tot_dev);
    hipError_t er = hipGetLastError();
    if (er != hipSuccess) return fail(PASC_ERR_HIP, "theta_update_noalloc launch: %s", hipGetErrorString(er));
    double tot[PASC_MAX_K * 4];
    er = hipMemcpyAsync(tot, tot_dev, sizeof(double) * nv, hipMemcpyDeviceToHost, s);
    if (er == hipSuccess) er = hipStreamSynchronize(s);
    if (er != hipSuccess) return fail(PASC_ERR_HIP, "theta_update_noalloc: %s", hipGetErrorString(er));
    // graphh1fem.cpp:374-379: theta_k = <x, gamma_k> / sum(gamma_k), 0 if the sum is 0
    for (int k = 0; k < d.K; k++)
        for (int c = 0; c < d.xdim; c++) {
            const double gs = tot[k];
            theta_out[k * d.xdim + c] = gs != 0 ? tot[d.K + k * d.xdim + c] / gs : 0.0;
        }
    return PASC_OK;
}

}  // namespace pasc

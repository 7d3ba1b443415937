// Per-iteration scalar kernels of the SPG engine (included by engine.hip only).
#pragma once

#include "device_common.h"
#include "step_scalar.h"

namespace pasc {

// Step kernel: workgroups stride over the instances; each reduces one active instance's partial
// slots in a fixed order (deterministic) and runs step_scalar.
// nactive: decremented when an instance stops (single problems, which launch no scan); null
// when the active-tile scan recounts the active instances.
__global__ __launch_bounds__(256) void spg_step(const InstDesc *desc, InstState *state,
                                                 const double *part, int n_inst, StepParams P,
                                                 int32_t *dirty, int32_t *nactive) {
    // one workgroup per instance: a single instance of T = 1e7 owns ~1e4 partial slots
    __shared__ double lds[NPART * 4];
    __shared__ InstState lds_state;
    for (int w = blockIdx.x; w < n_inst; w += gridDim.x) {
        if (!uni(state[w].active)) continue;
        const InstDesc &D = desc[w];
        const double *pw = part + (int64_t)uni(D.tile0) * NPART;
        step_instance<256>(state + w, uni(D.ntiles), D.alpha * D.s_kkt, P, dirty, &lds_state, lds,
                           [&](int i, int q) { return pw[(int64_t)i * NPART + q]; },
                           [](int, int, double v) { return v; });
        if (nactive && threadIdx.x == 0 && !lds_state.active) atomicSub(nactive, 1);
        __syncthreads();  // lds_state is reused by the next instance
    }
}

// ---------------------------------------------------------------------------------------------
// Active-tile scan: one workgroup rebuilds, for every K group, the exclusive scan of the tile
// counts of its active instances and the tile -> position hint table, and counts the active
// instances for the host poll.
// ---------------------------------------------------------------------------------------------
struct ScanGroup {
    const int32_t *ids;
    int32_t *off;
    int32_t *hint;
    int32_t n;
};

constexpr int MAX_GROUPS = PASC_MAX_K;

struct ScanArgs {
    ScanGroup g[MAX_GROUPS];
    int32_t n_groups;
    const InstDesc *desc;
    const InstState *state;
    int32_t *nactive;
    int32_t *dirty;  // set by the step kernel when an instance stopped; cleared here
};

__global__ __launch_bounds__(1024) void spg_scan(ScanArgs A) {
    __shared__ int32_t sh[1024];
    __shared__ int32_t carry;
    if (*A.dirty == 0) return;  // nothing stopped since the last rebuild
    int active_total = 0;
    for (int gi = 0; gi < A.n_groups; gi++) {
        const ScanGroup G = A.g[gi];
        if (threadIdx.x == 0) carry = 0;
        __syncthreads();
        for (int base = 0; base < G.n; base += 1024) {
            const int a = base + threadIdx.x;
            int cnt = 0;
            if (a < G.n) {
                const int inst = G.ids[a];
                cnt = A.state[inst].active ? A.desc[inst].ntiles : 0;
                active_total += A.state[inst].active ? 1 : 0;
            }
            sh[threadIdx.x] = cnt;
            __syncthreads();
            for (int o = 1; o < 1024; o <<= 1) {
                const int v = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
                __syncthreads();
                sh[threadIdx.x] += v;
                __syncthreads();
            }
            const int excl = carry + sh[threadIdx.x] - cnt;
            if (a < G.n) G.off[a] = excl;
            __syncthreads();
            if (threadIdx.x == 1023) carry += sh[1023];
            __syncthreads();
        }
        if (threadIdx.x == 0) G.off[G.n] = carry;
        __syncthreads();
        // hints: bucket b (tiles [64b, 64b+64)) starts at the position whose range holds 64b;
        // one thread per bucket, binary search over the scan
        const int total = G.off[G.n];
        for (int b = threadIdx.x; (b << 6) < total; b += 1024) {
            const int j = b << 6;
            int lo = 0, hi = G.n;  // invariant: off[lo] <= j < off[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (G.off[mid] <= j) lo = mid; else hi = mid;
            }
            G.hint[b] = lo;
        }
        __syncthreads();
    }
    __shared__ int32_t red[1024];
    red[threadIdx.x] = active_total;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *A.nactive = red[0];
        *A.dirty = 0;
    }
}

}  // namespace pasc

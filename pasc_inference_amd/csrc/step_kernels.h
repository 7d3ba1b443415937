// Per-iteration scalar kernels of the SPG engine (included by engine.hip only).
#pragma once

#include "device_common.h"

namespace pasc {

// ---------------------------------------------------------------------------------------------
// Step: the scalar part of the reference iteration for one instance -- spg_fs
// (spg_fs.cpp:17-63), the generalised Armijo step (spgqpsolver.cpp:182-195), the BB step (:216)
// and the stopping tests (:260-274) plus the KKT test -- run by thread 0 of the workgroup that
// reduced the instance's partial slots.
// ---------------------------------------------------------------------------------------------
__device__ inline void step_scalar(const InstDesc &D, InstState &s, const double (&v)[NPART],
                                   const StepParams &P, int32_t *dirty) {
    const double dd = v[0], dAd = v[1], gd = v[2];
    // a negative residual marks a pass that did not evaluate it: keep the last value
    const bool have_kkt = v[4] >= 0.0 || v[4] != v[4];
    const double kkt = have_kkt ? v[4] : s.kkt;
    // The FP32 phase takes f from the exact quadratic identity (the generic SPG's
    // get_fx(fx_old, beta, gd, dAd), spgqpsolver.h:835-842): its increments are accurate, while
    // the FP32 sum of 1/2 <g - b, gamma> is noisy at the level the line search compares.
    const double f = (P.fx_identity && !s.first)
                         ? s.f + s.beta_apply * (s.gd + 0.5 * s.beta_apply * s.dAd)
                         : v[3];
    s.parity ^= 1;  // the pass wrote gamma_i into the other buffer
    s.kkt = kkt;
    const bool kkt_fresh = have_kkt;
    bool stop = false;
    int reason = PASC_REASON_MAXIT;
    if (!(f == f) || !(dd == dd) || !(dAd == dAd) || !(gd == gd) || !(kkt == kkt)) {
        s.status = PASC_ERR_NUMERIC;
        stop = true;
        reason = PASC_REASON_NUMERIC;
    } else if (s.first) {
        for (int i = 0; i < P.m; i++) s.fs[i] = f;
        s.fs_last = 0;
        s.it = 0;
        if ((P.stop_mask & PASC_STOP_KKT) && kkt_fresh && kkt <= P.kkt_tol) { stop = true; reason = PASC_REASON_KKT; }
    } else {
        s.fs_last = (s.fs_last + 1 >= P.m) ? 0 : s.fs_last + 1;
        s.fs[s.fs_last] = f;
        s.it += 1;
        if ((P.stop_mask & PASC_STOP_DIFFF) && fabs(f - s.f) < P.eps) { stop = true; reason = PASC_REASON_DIFFF; }
        else if ((P.stop_mask & PASC_STOP_NORMGP) && s.dd < P.eps) { stop = true; reason = PASC_REASON_NORMGP; }
        else if ((P.stop_mask & PASC_STOP_NORMGP_NORMB) && s.dd < P.eps * s.normb) { stop = true; reason = PASC_REASON_NORMGP_NORMB; }
        else if ((P.stop_mask & PASC_STOP_ANORMGP) && s.dAd < P.eps) { stop = true; reason = PASC_REASON_ANORMGP; }
        else if ((P.stop_mask & PASC_STOP_ANORMGP_NORMB) && s.dAd < P.eps * s.normb) { stop = true; reason = PASC_REASON_ANORMGP_NORMB; }
        else if ((P.stop_mask & PASC_STOP_KKT) && kkt_fresh && kkt <= P.kkt_tol) { stop = true; reason = PASC_REASON_KKT; }
    }
    if (!stop && s.it >= P.maxit) { stop = true; reason = PASC_REASON_MAXIT; }
    if (!stop && dd == 0.0) { stop = true; reason = PASC_REASON_STATIONARY; }
    if (!stop && P.fp32_phase) {
        // FP32 phase ends at the switch threshold or when the KKT residual stagnates
        if (kkt_fresh && kkt < 0.9 * s.best_kkt) { s.best_kkt = kkt; s.stall = 0; }
        else s.stall += 1;
        // With gamma rounded to FP32 after every update, SPG stalls once the residual nears
        // ~1e-5 * epsilon (per-sample epsilon = alpha * s_kkt): the stencil amplifies the
        // rounding of gamma by up to 8 epsilon and the BB steps degrade well above that floor.
        // The FP32 phase therefore hands over no later than at that level.
        const double floor32 = 1e-5 * D.alpha * D.s_kkt;
        const double sw = P.fp32_switch > floor32 ? P.fp32_switch : floor32;
        if ((kkt_fresh && kkt <= sw) || s.stall >= P.stall_window) {
            stop = true;
            reason = PASC_REASON_MAXIT;
        }
    }
    s.f_prev = s.f;
    s.f = f;
    s.dd = dd;
    s.dAd = dAd;
    s.gd = gd;
    if (stop) {
        s.active = 0;
        s.reason = reason;
        s.first = 0;
        *dirty = 1;  // the active-tile enumeration must be rebuilt
        return;
    }
    double fmax = s.fs[0];
    for (int i = 0; i < P.m; i++) fmax = s.fs[i] > fmax ? s.fs[i] : fmax;
    double beta, alpha_next;
    if (dAd > 0.0) {
        const double xi = (fmax - f) / dAd;
        const double beta_bar = -gd / dAd;
        double beta_hat = P.gamma * beta_bar + sqrt(P.gamma * P.gamma * beta_bar * beta_bar + 2.0 * xi);
        if (beta_hat < P.sigma1) beta_hat = P.sigma1;
        beta = beta_hat < P.sigma2 ? beta_hat : P.sigma2;
        alpha_next = dd / dAd;
    } else {
        // d lies in the null space of A (objective linear along d): the reference's IEEE
        // arithmetic gives beta = sigma2 there; the BB step would be infinite, so restart it.
        beta = P.sigma2;
        alpha_next = P.alphainit;
    }
    s.dd_prev = dd;
    s.dAd_prev = dAd;
    s.beta_apply = beta;
    s.alpha_apply = s.alpha_dir;
    s.alpha_dir = alpha_next;
    s.first = 0;
}


// Step kernel: workgroups stride over the instances; each reduces one active instance's partial
// slots in a fixed order (deterministic) and runs step_scalar.
__global__ __launch_bounds__(256) void spg_step(const InstDesc *desc, InstState *state,
                                                 const double *part, int n_inst, StepParams P,
                                                 int32_t *dirty) {
    // one workgroup per instance: a single instance of T = 1e7 owns ~1e4 partial slots
    __shared__ double lds[NPART * 4];
    for (int w = blockIdx.x; w < n_inst; w += gridDim.x) {
    InstState &s = state[w];
    if (!s.active) continue;
    const InstDesc &D = desc[w];
    double v[NPART] = {0.0, 0.0, 0.0, 0.0, -1.0};  // residual -1: not evaluated
    for (int i = threadIdx.x; i < D.ntiles; i += blockDim.x) {
        const double *p = part + (int64_t)(D.tile0 + i) * NPART;
#pragma unroll
        for (int q = 0; q < NPART - 1; q++) v[q] += p[q];
        const double m = p[NPART - 1];
        v[NPART - 1] = (m > v[NPART - 1] || m != m) ? m : v[NPART - 1];
    }
    block_reduce_part<256>(v, lds);
    if (threadIdx.x == 0) step_scalar(D, s, v, P, dirty);
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

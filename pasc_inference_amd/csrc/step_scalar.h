// The scalar SPG logic of one instance per iteration (shared by the step kernel and the fused
// 1-D pass, whose last workgroup per instance runs it).
#pragma once

#include "device_common.h"

namespace pasc {

// ---------------------------------------------------------------------------------------------
// Step: the scalar part of the reference iteration for one instance -- spg_fs
// (spg_fs.cpp:17-63), the generalised Armijo step (spgqpsolver.cpp:182-195), the BB step (:216)
// and the stopping tests (:260-274) plus the KKT test -- run by thread 0 of the workgroup that
// reduced the instance's partial slots.
// ---------------------------------------------------------------------------------------------
// sample_eps: the per-sample epsilon alpha * s_kkt of the instance (FP32 hand-over floor).
__device__ inline void step_scalar(double sample_eps, InstState &s, const double (&v)[NPART],
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
    if (!stop && s.it + s.it_base >= P.maxit) { stop = true; reason = PASC_REASON_MAXIT; }
    if (!stop && dd == 0.0) { stop = true; reason = PASC_REASON_STATIONARY; }
    if (!stop && P.fp32_phase) {
        // FP32 phase ends at the switch threshold or when the KKT residual stagnates
        if (kkt_fresh && kkt < 0.9 * s.best_kkt) { s.best_kkt = kkt; s.stall = 0; }
        else s.stall += 1;
        // With gamma rounded to FP32 after every update, SPG stalls once the residual nears
        // ~1e-5 * epsilon (per-sample epsilon = alpha * s_kkt): the stencil amplifies the
        // rounding of gamma by up to 8 epsilon and the BB steps degrade well above that floor.
        // The FP32 phase therefore hands over no later than at that level.
        const double floor32 = 1e-5 * sample_eps;
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

// One instance's step by a workgroup of NT threads.  The state is staged in LDS word by word by
// many lanes at once, so that step_scalar's chain of dependent reads and writes costs LDS
// latency rather than a device-memory round trip each; the partial slots are read meanwhile.
// load_slot(i, q) loads component q of the instance's partial slot i (i < ntiles); settle(i,
// q, v) turns that load into the value (the fused step waits there for a value not yet
// published and marks the slot empty again; the step kernel returns v).
static_assert(sizeof(InstState) % 8 == 0, "InstState is staged in 8-byte words");
constexpr int STATE_WORDS = (int)(sizeof(InstState) / 8);

template <int NT, typename LoadSlot, typename Settle>
__device__ __forceinline__ void step_instance(InstState *state, int ntiles, double sample_eps,
                                              const StepParams &P, int32_t *dirty,
                                              InstState *lds_state, double *lds_red,
                                              LoadSlot load_slot, Settle settle) {
    constexpr int WPT = (STATE_WORDS + NT - 1) / NT;  // state words per thread
    uint64_t *gw = reinterpret_cast<uint64_t *>(state);
    uint64_t *lw = reinterpret_cast<uint64_t *>(lds_state);
    uint64_t word[WPT];
#pragma unroll
    for (int q = 0; q < WPT; q++) {
        const int w = threadIdx.x + q * NT;
        word[q] = w < STATE_WORDS ? gw[w] : 0;
    }
    double v[NPART] = {0.0, 0.0, 0.0, 0.0, -1.0};  // residual -1: not evaluated
    // B slots per thread are loaded before any is used, so that their latencies overlap
    constexpr int B = 4;
    for (int i0 = threadIdx.x; i0 < ntiles; i0 += B * NT) {
        double w[B][NPART];
#pragma unroll
        for (int b = 0; b < B; b++) {
            // past the end: the neutral element (each slot is read exactly once)
            const int i = i0 + b * NT;
#pragma unroll
            for (int q = 0; q < NPART; q++)
                w[b][q] = i < ntiles ? load_slot(i, q) : (q == NPART - 1 ? -1.0 : 0.0);
        }
#pragma unroll
        for (int b = 0; b < B; b++) {
            const int i = i0 + b * NT;
            if (i < ntiles) {
#pragma unroll
                for (int q = 0; q < NPART; q++) w[b][q] = settle(i, q, w[b][q]);
            }
        }
#pragma unroll
        for (int b = 0; b < B; b++) {
#pragma unroll
            for (int q = 0; q < NPART - 1; q++) v[q] += w[b][q];
            const double m = w[b][NPART - 1];
            v[NPART - 1] = (m > v[NPART - 1] || m != m) ? m : v[NPART - 1];
        }
    }
#pragma unroll
    for (int q = 0; q < WPT; q++) {
        const int w = threadIdx.x + q * NT;
        if (w < STATE_WORDS) lw[w] = word[q];
    }
    block_reduce_part<NT>(v, lds_red);  // its barriers also publish the staged state
    if (threadIdx.x == 0) step_scalar(sample_eps, *lds_state, v, P, dirty);
    __syncthreads();
    for (int w = threadIdx.x; w < STATE_WORDS; w += NT) gw[w] = lw[w];
}


}  // namespace pasc

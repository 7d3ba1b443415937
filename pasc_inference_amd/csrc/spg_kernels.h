// SPG-QP kernels for gfx950: the fused single-pass 1-D iteration, the grid-graph iteration,
// the device-side step-size/stopping logic and the active-tile scan.
//
// One reference SPG iteration (spgqpsolver.cpp:140-297) is ~14 PETSc launches, 3 MPI
// reductions and ~38 vector passes.  Here the iterate gamma is the only state: the gradient
// g = alpha*A0*gamma - b is rebuilt from gamma and the data x wherever it is needed (it equals
// the reference's incrementally updated g in exact arithmetic), so one pass per iteration reads
// gamma_{i-1} and x once and writes gamma_i once:
//
//   gamma_i = gamma_{i-1} + beta_{i-1} * (P(gamma_{i-1} - alpha_{i-1} g_{i-1}) - gamma_{i-1})
//   d_i     = P(gamma_i - alpha_i g_i) - gamma_i,   Ad_i = alpha*A0*d_i
//   partials: dd, dAd, gd of d_i;  f_i = 1/2 <g_i - b, gamma_i>;  KKT residual of gamma_i
//
// The time stencil has radius 1 and is applied three times (g_{i-1}, g_i, Ad_i), so each
// thread owns CH consecutive time steps and recomputes a 3-step halo on each side in registers
// (neighbouring threads load the same cache lines; HBM sees each element once).
#pragma once

#include "device_common.h"

namespace pasc {

struct PassArgs {
    void *gamma0;
    void *gamma1;
    const void *x;
    const InstDesc *desc;
    const InstState *state;
    double *part;
    GroupView grp;
};

// 2*(gamma_t - gamma_{t-1})[t>0] + 2*(gamma_t - gamma_{t+1})[t<T-1]: row of A0
// (blockgraphsparse.cpp:60-89) in difference form, which keeps FP32 rounding relative to the
// differences rather than to 4*gamma_t.
template <typename C, int W>
__device__ __forceinline__ C lap_time(const C (&v)[W], int j, int64_t t, int64_t T) {
    C a = C(0);
    if (t > 0) a += v[j] - v[j - 1];
    if (t < T - 1) a += v[j] - v[j + 1];
    return C(2) * a;
}

template <typename S, typename C, int W>
__device__ __forceinline__ void load_window(const S *__restrict__ row, int64_t tw0, int64_t T,
                                            int64_t ld, C (&v)[W]) {
    using VT = V16<S>;
    constexpr int VW = VT::n;
    static_assert(W % VW == 0, "window must be whole vectors");
#pragma unroll
    for (int q = 0; q < W / VW; q++) {
        const int64_t t = tw0 + q * VW;
        if (t >= 0 && t < ld) {
            const typename VT::type vec = *reinterpret_cast<const typename VT::type *>(row + t);
#pragma unroll
            for (int i = 0; i < VW; i++) v[q * VW + i] = (t + i < T) ? C(VT::get(vec, i)) : C(0);
        } else {
#pragma unroll
            for (int i = 0; i < VW; i++) v[q * VW + i] = C(0);
        }
    }
}

template <typename S, typename C, int K, int CH, int NT>
__device__ __forceinline__ void pass_tile_1d(const PassArgs &A, int inst, int chunk,
                                             double *lds) {
    constexpr int W = CH + 8;  // window starts 4 steps before the first owned step
    const InstDesc &D = A.desc[inst];
    const InstState &st = A.state[inst];
    const int64_t T = D.T;
    const int64_t ld = D.ld;
    const int64_t t0 = (int64_t)chunk * (NT * CH) + (int64_t)threadIdx.x * CH;
    const int par = st.parity;
    const S *src = reinterpret_cast<const S *>(par ? A.gamma1 : A.gamma0) + D.goff;
    S *dst = reinterpret_cast<S *>(par ? A.gamma0 : A.gamma1) + D.goff;
    const S *xr = reinterpret_cast<const S *>(A.x) + D.xoff;

    const C alpha = C(D.alpha);
    const C bscale = C(D.bscale);
    const C s_kkt = C(D.s_kkt);
    const C a_apply = C(st.alpha_apply);
    const C b_apply = C(st.beta_apply);
    const C a_dir = C(st.alpha_dir);
    const bool first = st.first != 0;

    double acc[NPART] = {0.0, 0.0, 0.0, 0.0, 0.0};

    if (t0 < T) {
        C X[W];
        load_window<S, C, W>(xr, t0 - 4, T, ld, X);
        C th[K];
#pragma unroll
        for (int k = 0; k < K; k++) th[k] = C(D.theta[k]);
        // data term q = bscale*(x - theta_k)^2 = -b (zero when held out or outside [0,T))
        bool obs[W];
#pragma unroll
        for (int j = 0; j < W; j++) {
            const int64_t t = t0 - 4 + j;
            bool o = (t >= 0 && t < T);
            if (o && D.heldout) o = D.heldout[t] == 0;
            obs[j] = o;
        }
        C G[K][W];
#pragma unroll
        for (int k = 0; k < K; k++) load_window<S, C, W>(src + k * ld, t0 - 4, T, ld, G[k]);

        // level 1: gamma_i on window positions [2, W-2)
        C G1[K][W];
#pragma unroll
        for (int j = 2; j < W - 2; j++) {
            const int64_t t = t0 - 4 + j;
            C y[K];
            if (first) {
#pragma unroll
                for (int k = 0; k < K; k++) y[k] = G[k][j];
                project_simplex<K, C>(y);
#pragma unroll
                for (int k = 0; k < K; k++) G1[k][j] = y[k];
            } else {
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const C dx = X[j] - th[k];
                    const C q = obs[j] ? bscale * dx * dx : C(0);
                    const C g = alpha * lap_time<C, W>(G[k], j, t, T) + q;
                    y[k] = G[k][j] - a_apply * g;
                }
                project_simplex<K, C>(y);
#pragma unroll
                for (int k = 0; k < K; k++) G1[k][j] = G[k][j] + b_apply * (y[k] - G[k][j]);
            }
        }
        // level 2: g_i and d_i on [3, W-3)
        C Dd[K][W];
        C Gi[K][CH];
#pragma unroll
        for (int j = 3; j < W - 3; j++) {
            const int64_t t = t0 - 4 + j;
            C y[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const C dx = X[j] - th[k];
                const C q = obs[j] ? bscale * dx * dx : C(0);
                const C g = alpha * lap_time<C, W>(G1[k], j, t, T) + q;
                if (j >= 4 && j < W - 4) Gi[k][j - 4] = g;
                y[k] = G1[k][j] - a_dir * g;
            }
            project_simplex<K, C>(y);
#pragma unroll
            for (int k = 0; k < K; k++) Dd[k][j] = y[k] - G1[k][j];
        }
        // level 3: owned steps [4, W-4)
#pragma unroll
        for (int j = 4; j < W - 4; j++) {
            const int64_t t = t0 - 4 + j;
            if (t < T) {
                C z[K];
                // <g, d> with the per-step common mode g_0t removed: every block of d sums to 0,
                // so the value is unchanged, but the rounding of sum_k d_kt no longer multiplies
                // the large data term (which otherwise stalls SPG near 1e-9 KKT).
                const C gc = Gi[0][j - 4];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const C d = Dd[k][j];
                    const C Ad = alpha * lap_time<C, W>(Dd[k], j, t, T);
                    const C g = Gi[k][j - 4];
                    const C dx = X[j] - th[k];
                    const C q = obs[j] ? bscale * dx * dx : C(0);
                    acc[0] += (double)d * (double)d;
                    acc[1] += (double)d * (double)Ad;
                    acc[2] += ((double)g - (double)gc) * (double)d;
                    acc[3] += 0.5 * ((double)g + (double)q) * (double)G1[k][j];
                    z[k] = G1[k][j] - s_kkt * g;
                }
                project_simplex<K, C>(z);
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double e = fabs((double)G1[k][j] - (double)z[k]);
                    acc[4] = (e > acc[4] || e != e) ? e : acc[4];
                }
            }
        }
        // store gamma_i for the owned steps
        using VT = V16<S>;
        constexpr int VW = VT::n;
#pragma unroll
        for (int k = 0; k < K; k++) {
#pragma unroll
            for (int q = 0; q < CH / VW; q++) {
                const int64_t t = t0 + q * VW;
                if (t + VW <= T) {
                    typename VT::type vec;
#pragma unroll
                    for (int i = 0; i < VW; i++) VT::set(vec, i, S(G1[k][4 + q * VW + i]));
                    *reinterpret_cast<typename VT::type *>(dst + k * ld + t) = vec;
                } else {
#pragma unroll
                    for (int i = 0; i < VW; i++)
                        if (t + i < T) dst[k * ld + t + i] = S(G1[k][4 + q * VW + i]);
                }
            }
        }
    }
    block_reduce_part<NT>(acc, lds);
    if (threadIdx.x == 0) {
        double *o = A.part + (int64_t)(D.tile0 + chunk) * NPART;
#pragma unroll
        for (int i = 0; i < NPART; i++) o[i] = acc[i];
    }
}

template <typename S, typename C, int K, int CH, int NT>
__global__ __launch_bounds__(NT) void spg_pass_1d(PassArgs A) {
    __shared__ double lds[NPART * (NT / WAVE)];
    const int total = A.grp.off[A.grp.n];
    for (int j = blockIdx.x; j < total; j += gridDim.x) {
        int inst, chunk;
        find_tile(A.grp, j, inst, chunk);
        pass_tile_1d<S, C, K, CH, NT>(A, inst, chunk, lds);
    }
}

// ---------------------------------------------------------------------------------------------
// Grid graph (T = 1, R = width*height, 4 neighbours; blockgraphsparse.cpp:51-66,93-104 with
// T == 1: diagonal deg(r), -1 per neighbour).  Unfused: the pass rebuilds gamma_i pointwise
// from gamma_{i-1} + beta*d_{i-1} (also at the 4 neighbours), forms g_i and d_i, and stores
// gamma_i and d_i; a second kernel forms A*d_i.  Data is stored SoA, rows c = 0..xdim-1.
// ---------------------------------------------------------------------------------------------
struct GridArgs {
    void *gamma0;
    void *gamma1;
    void *dir0;  // d_{i-1} / d_i ping-pong, indexed like gamma by the instance parity
    void *dir1;
    const void *x;
    const InstDesc *desc;
    const InstState *state;
    double *part;
    GroupView grp;
    int32_t width;
};

template <typename S, typename C, int K>
__device__ __forceinline__ void grid_gamma_at(const GridArgs &A, const InstDesc &D,
                                              const InstState &st, const S *src, int64_t r,
                                              C (&v)[K]) {
    if (st.first) {
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = C(src[k * D.ld + r]);
        project_simplex<K, C>(v);
    } else {
        const S *dr = reinterpret_cast<const S *>(st.parity ? A.dir1 : A.dir0) + D.goff;
        const C b = C(st.beta_apply);
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = C(src[k * D.ld + r]) + b * C(dr[k * D.ld + r]);
    }
}

template <typename C, int K>
__device__ __forceinline__ void grid_dataterm(const InstDesc &D, const void *xv, int64_t r,
                                              bool is_f32, C (&q)[K]) {
    const bool obs = !(D.heldout && D.heldout[r]);
#pragma unroll
    for (int k = 0; k < K; k++) q[k] = C(0);
    if (!obs) return;
    for (int c = 0; c < D.xdim; c++) {
        const C xc = is_f32 ? C(reinterpret_cast<const float *>(xv)[D.xoff + c * D.ld + r])
                            : C(reinterpret_cast<const double *>(xv)[D.xoff + c * D.ld + r]);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const C dx = xc - C(D.theta[k * D.xdim + c]);
            q[k] += dx * dx;
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) q[k] *= C(D.bscale);
}

template <typename S, typename C, int K, int NT>
__global__ __launch_bounds__(NT) void spg_pass_grid_dir(GridArgs A) {
    __shared__ double lds[NPART * (NT / WAVE)];
    const int total = A.grp.off[A.grp.n];
    const bool is_f32 = sizeof(S) == 4;
    for (int j = blockIdx.x; j < total; j += gridDim.x) {
        int inst, chunk;
        find_tile(A.grp, j, inst, chunk);
        const InstDesc &D = A.desc[inst];
        const InstState &st = A.state[inst];
        const int64_t R = D.T;
        const int64_t W = A.width;
        const int64_t r = (int64_t)chunk * NT + threadIdx.x;
        const int par = st.parity;
        const S *src = reinterpret_cast<const S *>(par ? A.gamma1 : A.gamma0) + D.goff;
        S *dst = reinterpret_cast<S *>(par ? A.gamma0 : A.gamma1) + D.goff;
        S *dr = reinterpret_cast<S *>(par ? A.dir0 : A.dir1) + D.goff;
        double acc[NPART] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (r < R) {
            const int64_t row = r / W, col = r - row * W;
            C g0[K];
            grid_gamma_at<S, C, K>(A, D, st, src, r, g0);
            C lap[K];
#pragma unroll
            for (int k = 0; k < K; k++) lap[k] = C(0);
            const int64_t nb[4] = {row > 0 ? r - W : -1, col > 0 ? r - 1 : -1,
                                   col < W - 1 ? r + 1 : -1, r + W < R ? r + W : -1};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                if (nb[e] < 0) continue;
                C gn[K];
                grid_gamma_at<S, C, K>(A, D, st, src, nb[e], gn);
#pragma unroll
                for (int k = 0; k < K; k++) lap[k] += g0[k] - gn[k];
            }
            C q[K];
            grid_dataterm<C, K>(D, A.x, r, is_f32, q);
            C y[K], z[K], g[K];
            const C alpha = C(D.alpha), a_dir = C(st.alpha_dir), s = C(D.s_kkt);
#pragma unroll
            for (int k = 0; k < K; k++) {
                g[k] = alpha * lap[k] + q[k];
                y[k] = g0[k] - a_dir * g[k];
                z[k] = g0[k] - s * g[k];
            }
            project_simplex<K, C>(y);
            project_simplex<K, C>(z);
#pragma unroll
            for (int k = 0; k < K; k++) {
                const C d = y[k] - g0[k];
                dst[k * D.ld + r] = S(g0[k]);
                dr[k * D.ld + r] = S(d);
                acc[2] += ((double)g[k] - (double)g[0]) * (double)d;  // common mode removed
                acc[3] += 0.5 * ((double)g[k] + (double)q[k]) * (double)g0[k];
                const double e = fabs((double)g0[k] - (double)z[k]);
                acc[4] = (e > acc[4] || e != e) ? e : acc[4];
            }
        }
        block_reduce_part<NT>(acc, lds);
        if (threadIdx.x == 0) {
            double *o = A.part + (int64_t)(D.tile0 + chunk) * NPART;
            o[2] = acc[2];
            o[3] = acc[3];
            o[4] = acc[4];
        }
    }
}

template <typename S, typename C, int K, int NT>
__global__ __launch_bounds__(NT) void spg_pass_grid_ad(GridArgs A) {
    __shared__ double lds[NPART * (NT / WAVE)];
    const int total = A.grp.off[A.grp.n];
    for (int j = blockIdx.x; j < total; j += gridDim.x) {
        int inst, chunk;
        find_tile(A.grp, j, inst, chunk);
        const InstDesc &D = A.desc[inst];
        const int64_t R = D.T;
        const int64_t W = A.width;
        const int64_t r = (int64_t)chunk * NT + threadIdx.x;
        // d_i was written by spg_pass_grid_dir into the buffer opposite to the parity
        const S *dr = reinterpret_cast<const S *>(A.state[inst].parity ? A.dir0 : A.dir1) + D.goff;
        double acc[NPART] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (r < R) {
            const int64_t row = r / W, col = r - row * W;
            const int64_t nb[4] = {row > 0 ? r - W : -1, col > 0 ? r - 1 : -1,
                                   col < W - 1 ? r + 1 : -1, r + W < R ? r + W : -1};
            const C alpha = C(D.alpha);
#pragma unroll
            for (int k = 0; k < K; k++) {
                const C d = C(dr[k * D.ld + r]);
                C lap = C(0);
#pragma unroll
                for (int e = 0; e < 4; e++)
                    if (nb[e] >= 0) lap += d - C(dr[k * D.ld + nb[e]]);
                const C Ad = alpha * lap;
                acc[0] += (double)d * (double)d;
                acc[1] += (double)d * (double)Ad;
            }
        }
        block_reduce_part<NT>(acc, lds);
        if (threadIdx.x == 0) {
            double *o = A.part + (int64_t)(D.tile0 + chunk) * NPART;
            o[0] = acc[0];
            o[1] = acc[1];
        }
    }
}

}  // namespace pasc

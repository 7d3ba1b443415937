// Auxiliary kernels: stand-alone projection and matvec (the kernel-level C-ABI entry points),
// explicit FP64 objective/KKT evaluation, the theta update, labels, abserr and conversions.
#pragma once

#include "device_common.h"

namespace pasc {

// ---- stand-alone projection, K <= 8: one thread per block of K values ----------------------
// layout 0: t-major (block t = x[t*K .. t*K+K-1]);  layout 1: k-major (x[k*ld + t])
template <typename S, typename C, int K>
__global__ __launch_bounds__(256) void project_thread(S *x, int64_t nblocks, int layout,
                                                      int64_t ld) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblocks) return;
    C y[K];
#pragma unroll
    for (int k = 0; k < K; k++) y[k] = C(layout == 0 ? x[t * K + k] : x[k * ld + t]);
    project_simplex<K, C>(y);
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (layout == 0) x[t * K + k] = S(y[k]);
        else x[k * ld + t] = S(y[k]);
    }
}

// ---- stand-alone projection, large K: one wavefront per block ------------------------------
// Lanes hold the entries k = lane + 64*i.  The threshold is bracketed in [max-1, max) and
// bisected (wave-wide sums), then finished exactly by the support fixed point.
template <typename S>
__global__ __launch_bounds__(256) void project_wave(S *x, int64_t nblocks, int K, int layout,
                                                    int64_t ld) {
    const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    if (b >= nblocks) return;
    auto at = [&](int k) -> int64_t { return layout == 0 ? b * K + k : (int64_t)k * ld + b; };
    double s = 0.0, mx = -1e308;
    int neg = 0;
    for (int k = lane; k < K; k += WAVE) {
        const double v = (double)x[at(k)];
        s += v;
        neg |= v < 0.0;
        mx = v > mx ? v : mx;
    }
    s = wave_sum(s);
    neg = wave_sum_i(neg);
    mx = wave_max(mx);
    if (neg == 0 && s == 1.0) return;
    double lo = mx - 1.0, hi = mx;
    for (int it = 0; it < 24; it++) {
        const double mid = 0.5 * (lo + hi);
        double phi = 0.0;
        for (int k = lane; k < K; k += WAVE) {
            const double v = (double)x[at(k)] - mid;
            phi += v > 0.0 ? v : 0.0;
        }
        phi = wave_sum(phi) - 1.0;
        if (phi >= 0.0) lo = mid; else hi = mid;
    }
    double tau = lo;
    int cnt_prev = -1;
    for (int it = 0; it < K; it++) {
        double sum = 0.0;
        int cnt = 0;
        for (int k = lane; k < K; k += WAVE) {
            const double v = (double)x[at(k)];
            if (v > tau) { sum += v; cnt++; }
        }
        sum = wave_sum(sum);
        cnt = wave_sum_i(cnt);
        if (cnt == cnt_prev || cnt == 0) break;
        cnt_prev = cnt;
        tau = (sum - 1.0) / (double)cnt;
    }
    for (int k = lane; k < K; k += WAVE) {
        const double v = (double)x[at(k)] - tau;
        x[at(k)] = S(v > 0.0 ? v : 0.0);
    }
}

// ---- stand-alone matvec y = alpha*A0*x (k-major rows of stride ld) --------------------------
template <typename S>
__global__ __launch_bounds__(256) void matvec_time(const S *x, S *y, int64_t T, int K,
                                                   int64_t ld, double alpha) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (t >= T || k >= K) return;
    const S *r = x + k * ld;
    double a = 0.0;
    if (t > 0) a += 2.0 * ((double)r[t] - (double)r[t - 1]);
    if (t < T - 1) a += 2.0 * ((double)r[t] - (double)r[t + 1]);
    y[k * ld + t] = S(alpha * a);
}

template <typename S>
__global__ __launch_bounds__(256) void matvec_graph(const S *x, S *y, GraphGeo G, int K,
                                                    int64_t ld, double alpha) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (n >= G.T * G.R || k >= K) return;
    const S *v = x + k * ld;
    const double c = (double)v[n];
    double a = 0.0;
    for_neighbours(G, n, [&](int64_t n2, double w) { a += w * (c - (double)v[n2]); });
    y[k * ld + n] = S(alpha * a);
}

// ---- explicit FP64 evaluation of f = 1/2 g'Ag - b'g, the KKT residual and ||b||^2 -----------
// Partials per block: [0] f, [1] ||b||^2, [2] kkt (max).
template <typename S, int K>
__global__ __launch_bounds__(256) void eval_kernel(const S *gam, const double *x,
                                                   InstDesc D, GraphGeo G, double *part) {
    __shared__ double lds[3 * 4];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t T = D.T;
    double f = 0.0, nb2 = 0.0, kkt = 0.0;
    if (t < T) {
        double g0[K], lap[K], q[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            g0[k] = (double)gam[k * D.ld + t];
            lap[k] = 0.0;
            q[k] = 0.0;
        }
        for_neighbours(G, t, [&](int64_t n2, double w) {
#pragma unroll
            for (int k = 0; k < K; k++) lap[k] += w * (g0[k] - (double)gam[k * D.ld + n2]);
        });
        if (D.b64) {
#pragma unroll
            for (int k = 0; k < K; k++) q[k] = -D.b64[k * D.ld + t];
        } else if (!(D.heldout && D.heldout[t])) {
            for (int c = 0; c < D.xdim; c++) {
                const double xc = x[D.xoff + c * D.ld + t];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double dx = xc - D.theta[k * D.xdim + c];
                    q[k] += dx * dx;
                }
            }
#pragma unroll
            for (int k = 0; k < K; k++) q[k] *= D.bscale;
        }
        double z[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const double g = D.alpha * lap[k] + q[k];
            f += 0.5 * D.alpha * lap[k] * g0[k] + q[k] * g0[k];
            nb2 += q[k] * q[k];
            z[k] = g0[k] - D.s_kkt * g;
        }
        project_simplex<K, double>(z);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const double e = fabs(g0[k] - z[k]);
            kkt = (e > kkt || e != e) ? e : kkt;
        }
    }
    f = wave_sum(f);
    nb2 = wave_sum(nb2);
    kkt = wave_max(kkt);
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    if (lane == 0) {
        lds[w * 3 + 0] = f;
        lds[w * 3 + 1] = nb2;
        lds[w * 3 + 2] = kkt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q2 = 1; q2 < 4; q2++) {
            f += lds[q2 * 3 + 0];
            nb2 += lds[q2 * 3 + 1];
            const double m = lds[q2 * 3 + 2];
            kkt = (m > kkt || m != m) ? m : kkt;
        }
        part[blockIdx.x * 3 + 0] = f;
        part[blockIdx.x * 3 + 1] = nb2;
        part[blockIdx.x * 3 + 2] = kkt;
    }
}

// ---- theta update partials: per block sum_t gamma_k and sum_t x_c gamma_k (FP64) -------------
// part layout per block: [k] sums for k < K, then [K + k*xdim + c].
template <typename S>
__global__ __launch_bounds__(256) void theta_partials(const S *gam, const double *x, InstDesc D,
                                                      double *part) {
    const int K = D.K, xd = D.xdim;
    const int nv = K * (1 + xd);
    __shared__ double lds[PASC_MAX_K * 4 * 4];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    for (int v = 0; v < nv; v++) {
        double a = 0.0;
        if (t < D.T && !(D.heldout && D.heldout[t])) {
            if (v < K) {
                a = (double)gam[v * D.ld + t];
            } else {
                const int k = (v - K) / xd, c = (v - K) % xd;
                a = x[D.xoff + c * D.ld + t] * (double)gam[k * D.ld + t];
            }
        }
        a = wave_sum(a);
        if (lane == 0) lds[v * 4 + w] = a;
    }
    __syncthreads();
    if (threadIdx.x < nv) {
        const int v = threadIdx.x;
        part[(int64_t)blockIdx.x * nv + v] = lds[v * 4 + 0] + lds[v * 4 + 1] + lds[v * 4 + 2] + lds[v * 4 + 3];
    }
}

// ---- labels (argmax over k, lowest k on ties) and |sum_k theta_k gamma_k - sol| partials -----
template <typename S>
__global__ __launch_bounds__(256) void labels_kernel(const S *gam, InstDesc D, int32_t *lab) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.T) return;
    int best = 0;
    double bv = (double)gam[t];
    for (int k = 1; k < D.K; k++) {
        const double v = (double)gam[k * D.ld + t];
        if (v > bv) { bv = v; best = k; }
    }
    lab[t] = best;
}

template <typename S>
__global__ __launch_bounds__(256) void abserr_kernel(const S *gam, InstDesc D, const double *sol,
                                                     double *part) {
    __shared__ double lds[4];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0;
    if (t < D.T) {
        double rec = -sol[t];
        for (int k = 0; k < D.K; k++) rec += D.theta[k] * (double)gam[k * D.ld + t];
        a = fabs(rec);
    }
    a = wave_sum(a);
    if ((threadIdx.x & (WAVE - 1)) == 0) lds[threadIdx.x / WAVE] = a;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}

// ---- element-wise conversion (FP32 state -> FP64 polish state) ------------------------------
template <typename A, typename B>
__global__ __launch_bounds__(256) void convert_kernel(const A *src, B *dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = B(src[i]);
}

// ---- fill of 64-bit words (the PART_EMPTY marks of the fused step's partial sums) ----------
static __global__ __launch_bounds__(256) void fill_u64(uint64_t *p, int64_t n, uint64_t v) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        p[i] = v;
}

// ---- last gamma row ----------------------------------------------------------------------------
// The fused 1-D pass and the graph kernels store K - 1 rows of each iterate (the last is 1 - the
// sum of the others, clamped at 0).  At the end of a phase this writes the full last row of every instance's
// current iterate, with the pass's own arithmetic, so that the row equals the value the pass
// used.  Instances that ran no pass in the phase (first still set) keep their rows.  grid.y
// strides over the instances, grid.x over the steps (the nodes, for a graph).
template <typename S>
__global__ __launch_bounds__(256) void complete_last_row(const InstDesc *desc,
                                                         const InstState *state, int n_inst,
                                                         S *g0, S *g1) {
    for (int i = blockIdx.y; i < n_inst; i += gridDim.y) {
        const InstDesc &D = desc[i];
        const int K = D.K;
        if (K < 2 || state[i].first) continue;
        S *g = (state[i].parity ? g1 : g0) + D.goff;
        for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < D.T;
             t += (int64_t)gridDim.x * 256) {
            S r = S(1);
            for (int k = 0; k < K - 1; k++) r -= g[k * D.ld + t];
            g[(K - 1) * D.ld + t] = r > S(0) ? r : S(0);
        }
    }
}

}  // namespace pasc

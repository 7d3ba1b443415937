// Adjoint of the converged gamma step (any graph of the engine, xdim <= 3).
//
// At the minimiser gamma* of  min 1/2 g'Ag - b'g  over the product of simplices, let F be the
// free set {(k,t): gamma_kt > 0} and P the orthogonal projector onto
// {v : v = 0 off F, sum_{k in F_t} v_kt = 0 for every t}.  The KKT conditions P(A gamma - b) = 0
// differentiate to  (P A P) dgamma = -P(dA gamma - db).  For the held-out loss
//   J = sum_{t in M} (x_t - sum_k theta_k gamma_kt)^2
// one CG solve of (P A P) lambda = P dJ/dgamma gives every parameter derivative:
//   dJ/dlog(eps) = -lambda' A gamma                      (A = eps/T A0, so dA/dlog eps = A)
//   dJ/dtheta_k  = dJ/dtheta_k|direct + sum_{t with data} lambda_kt * 2 (x_t - theta_k) / T
// All vectors are FP64, k-major with row stride ld, one thread per time step.  The CG scalars
// (p'APp, r'r, the step and the update factor) are reduced and formed on the device, so an
// iteration needs no host round trip; the host reads the stop flag every CG_POLL iterations.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "device_common.h"

namespace pasc {

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(PASC_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

constexpr int ANT = 256;
constexpr double FREE_EPS = 1e-12;  // gamma entries at or below this are on the bound

struct AdjArgs {
    const double *gam;      // gamma* (k-major, stride ld)
    const double *x;        // data rows [xdim][ld]
    const uint8_t *mask;    // held-out nodes (1 = in the loss)
    const uint8_t *qmask;   // nodes without a data term in the QP (problem's held-out mask)
    uint8_t *freeset;       // [K][ld] 1 = free
    int64_t T, ld;          // T: nodes
    int K, xdim;
    double alpha, bscale;
    double theta[PASC_MAX_K * 3];
    GraphGeo G;
};

constexpr int NTH = PASC_MAX_K * 3;  // theta entries

// Scalars of the CG iteration on the device.  adj_cg_step and adj_cg_rr form them from the
// block partials of the kernel before; the vector kernels read step and b from here and do
// nothing once done is set.
struct CgState {
    double rr, rr0, step, b, tol2;
    int32_t it, maxit, done;
};
constexpr int CG_POLL = 16;  // CG iterations launched between host reads of the stop flag
constexpr int RNT = 1024;    // threads of the one-workgroup partial reductions

// (A0 v)_n of one row
__device__ inline double adj_lap(const AdjArgs &a, const double *row, int64_t n) {
    const double c = row[n];
    double l = 0.0;
    for_neighbours(a.G, n, [&](int64_t n2, double w) { l += w * (c - row[n2]); });
    return l;
}

// (A0 v)_n of one row of v = rv + b pv, formed at n and at its neighbours
__device__ inline double adj_lap_upd(const AdjArgs &a, const double *rv, const double *pv,
                                     double b, int64_t n) {
    const double c = rv[n] + b * pv[n];
    double l = 0.0;
    for_neighbours(a.G, n, [&](int64_t n2, double w) { l += w * (c - (rv[n2] + b * pv[n2])); });
    return l;
}

template <int NV>
__device__ inline void block_sum(double (&v)[NV], double *lds, double *out) {
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) lds[w * NV + i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < NV; i++) {
            double s = 0.0;
            for (int q = 0; q < ANT / WAVE; q++) s += lds[q * NV + i];
            out[(int64_t)blockIdx.x * NV + i] = s;
        }
    }
}

// P applied to the K values of one time step (in place); free flags per k.
__device__ inline void project_face(double *v, const bool *fr, int K) {
    int nf = 0;
    double s = 0.0;
    for (int k = 0; k < K; k++)
        if (fr[k]) { nf++; s += v[k]; }
    const double m = nf > 0 ? s / nf : 0.0;
    for (int k = 0; k < K; k++) v[k] = (fr[k] && nf > 1) ? v[k] - m : 0.0;
}

// r = P dJ/dgamma; partials: [0] J, [1] r'r, [2 + k*xdim + c] direct dJ/dtheta_kc
__global__ __launch_bounds__(ANT) void adj_init(AdjArgs a, double *r, double *part) {
    __shared__ double lds[(2 + NTH) * (ANT / WAVE)];
    const int64_t t = (int64_t)blockIdx.x * ANT + threadIdx.x;
    double v[2 + NTH];
    for (int i = 0; i < 2 + NTH; i++) v[i] = 0.0;
    if (t < a.T) {
        double g[PASC_MAX_K], dj[PASC_MAX_K];
        bool fr[PASC_MAX_K];
        for (int k = 0; k < a.K; k++) {
            g[k] = a.gam[k * a.ld + t];
            fr[k] = g[k] > FREE_EPS;
            a.freeset[k * a.ld + t] = fr[k] ? 1 : 0;
            dj[k] = 0.0;
        }
        const bool held = a.mask[t] != 0;
        for (int c = 0; c < a.xdim; c++) {
            double rec = 0.0;
            for (int k = 0; k < a.K; k++) rec += a.theta[k * a.xdim + c] * g[k];
            const double res = held ? a.x[c * a.ld + t] - rec : 0.0;
            v[0] += res * res;
            for (int k = 0; k < a.K; k++) {
                dj[k] += -2.0 * res * a.theta[k * a.xdim + c];
                v[2 + k * a.xdim + c] = -2.0 * res * g[k];
            }
        }
        project_face(dj, fr, a.K);
        for (int k = 0; k < a.K; k++) {
            r[k * a.ld + t] = dj[k];
            v[1] += dj[k] * dj[k];
        }
    }
    block_sum<2 + NTH>(v, lds, part);
}

// Sum of part[0 .. blocks) in a fixed order (thread i adds blocks i, i + RNT, ..., then the
// workgroup's tree); the result is valid in thread 0.
__device__ inline double reduce_partials(const double *part, int blocks, double *lds) {
    double v = 0.0;
    for (int b = threadIdx.x; b < blocks; b += RNT) v += part[b];
    v = wave_sum(v);
    if ((threadIdx.x & (WAVE - 1)) == 0) lds[threadIdx.x / WAVE] = v;
    __syncthreads();
    double sum = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < RNT / WAVE; i++) sum += lds[i];
    return sum;
}

// after adj_matvec: step = r'r / p'APp, or the stop when p'APp <= 0 (p in the null space of
// P A P, where the system is consistent)
__global__ __launch_bounds__(RNT) void adj_cg_step(const double *part, int blocks, CgState *cs) {
    __shared__ double lds[RNT / WAVE];
    if (cs->done) return;
    const double pq = reduce_partials(part, blocks, lds);
    if (threadIdx.x == 0) {
        if (!(pq > 0.0)) cs->done = 1;
        else cs->step = cs->rr / pq;
    }
}

// after adj_axpy: b = r'r (new) / r'r (old), the iteration counted, and the stop once r'r is
// below tol^2 r'r (initial) or the iterations reach maxit
__global__ __launch_bounds__(RNT) void adj_cg_rr(const double *part, int blocks, CgState *cs) {
    __shared__ double lds[RNT / WAVE];
    if (cs->done) return;
    const double rr_new = reduce_partials(part, blocks, lds);
    if (threadIdx.x == 0) {
        cs->b = rr_new / cs->rr;
        cs->rr = rr_new;
        cs->it += 1;
        if (!(cs->it < cs->maxit && rr_new > cs->tol2 * cs->rr0 && rr_new > 0.0)) cs->done = 1;
    }
}

// The search direction p = r + b pold of this step (b = 0 on the first, where pold = r), written
// to pnew, and q = P A P p (p in range(P)); partial p'q.  pnew and pold are two buffers, since
// the neighbours' pold are read while other workgroups write.
__global__ __launch_bounds__(ANT) void adj_matvec(AdjArgs a, const double *r, const double *pold,
                                                  double *pnew, double *q, double *part,
                                                  const CgState *cs) {
    __shared__ double lds[ANT / WAVE];
    if (cs->done) return;
    const double b = cs->b;
    const int64_t t = (int64_t)blockIdx.x * ANT + threadIdx.x;
    double v[1] = {0.0};
    if (t < a.T) {
        double w[PASC_MAX_K], pn[PASC_MAX_K];
        bool fr[PASC_MAX_K];
        for (int k = 0; k < a.K; k++) {
            const double *rv = r + k * a.ld, *pv = pold + k * a.ld;
            pn[k] = rv[t] + b * pv[t];
            w[k] = a.alpha * adj_lap_upd(a, rv, pv, b, t);
            fr[k] = a.freeset[k * a.ld + t] != 0;
        }
        project_face(w, fr, a.K);
        for (int k = 0; k < a.K; k++) {
            pnew[k * a.ld + t] = pn[k];
            q[k * a.ld + t] = w[k];
            v[0] += pn[k] * w[k];
        }
    }
    block_sum<1>(v, lds, part);
}

// lam += s p; r -= s q with s the step; partial r'r
__global__ __launch_bounds__(ANT) void adj_axpy(AdjArgs a, const CgState *cs, const double *p,
                                                const double *q, double *lam, double *r,
                                                double *part) {
    __shared__ double lds[ANT / WAVE];
    if (cs->done) return;
    const double s = cs->step;
    const int64_t t = (int64_t)blockIdx.x * ANT + threadIdx.x;
    double v[1] = {0.0};
    if (t < a.T) {
        for (int k = 0; k < a.K; k++) {
            const int64_t i = k * a.ld + t;
            lam[i] += s * p[i];
            r[i] -= s * q[i];
            v[0] += r[i] * r[i];
        }
    }
    block_sum<1>(v, lds, part);
}

// partials: [0] lambda' A gamma, [1 + k*xdim + c] sum_{n with data} lambda_kn 2 (x_nc - theta_kc) bscale
__global__ __launch_bounds__(ANT) void adj_final(AdjArgs a, const double *lam, double *part) {
    __shared__ double lds[(1 + NTH) * (ANT / WAVE)];
    const int64_t t = (int64_t)blockIdx.x * ANT + threadIdx.x;
    double v[1 + NTH];
    for (int i = 0; i < 1 + NTH; i++) v[i] = 0.0;
    if (t < a.T) {
        const bool nodata = a.qmask && a.qmask[t] != 0;  // b_kn does not depend on theta there
        for (int k = 0; k < a.K; k++) {
            const double lk = lam[k * a.ld + t];
            v[0] += lk * a.alpha * adj_lap(a, a.gam + k * a.ld, t);
            if (!nodata)
                for (int c = 0; c < a.xdim; c++)
                    v[1 + k * a.xdim + c] =
                        lk * 2.0 * (a.x[c * a.ld + t] - a.theta[k * a.xdim + c]) * a.bscale;
        }
    }
    block_sum<1 + NTH>(v, lds, part);
}

static int sum_partials(hipStream_t s, const double *d_part, int blocks, int nv,
                        std::vector<double> &out) {
    std::vector<double> h((size_t)blocks * nv);
    HIP_TRY(hipMemcpyAsync(h.data(), d_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    out.assign(nv, 0.0);
    for (int b = 0; b < blocks; b++)
        for (int i = 0; i < nv; i++) out[i] += h[(size_t)b * nv + i];
    return PASC_OK;
}

template <typename A, typename B>
__global__ __launch_bounds__(256) void adj_convert(const A *src, B *dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = B(src[i]);
}

int engine_adjoint(Engine &e, int inst, const uint8_t *mask_host, double cg_tol, int cg_maxit,
                   double *loss, double *dlogeps, double *dtheta, int *cg_it) {
    const InstDesc &D = e.h_desc[inst];
    if (D.b64) return fail(PASC_ERR_ARG, "adjoint: needs the data term");
    hipStream_t s = (hipStream_t)e.ctx->stream;
    const int64_t T = D.T, ld = D.ld, n = (int64_t)D.K * ld;
    const int blocks = (int)((T + ANT - 1) / ANT);
    const int par = e.h_parity[inst];
    double *buf = nullptr;
    uint8_t *fr = nullptr, *mk = nullptr;
    double *part = nullptr;
    CgState *d_cs = nullptr;
    // gamma (FP64 copy when the state is FP32), lambda, r, q and two buffers of p
    HIP_TRY(hipMalloc(&buf, sizeof(double) * 6 * (size_t)n));
    HIP_TRY(hipMalloc(&fr, (size_t)n));
    HIP_TRY(hipMalloc(&mk, (size_t)T));
    HIP_TRY(hipMalloc(&part, sizeof(double) * (size_t)blocks * (2 + NTH)));
    HIP_TRY(hipMalloc(&d_cs, sizeof(CgState)));
    double *gam = buf, *lam = buf + n, *r = buf + 2 * n, *q = buf + 3 * n;
    double *pb[2] = {buf + 4 * n, buf + 5 * n};
    int rc = PASC_OK;
    auto cleanup = [&]() {
        (void)hipFree(buf);
        (void)hipFree(fr);
        (void)hipFree(mk);
        (void)hipFree(part);
        (void)hipFree(d_cs);
    };
    const unsigned cb = (unsigned)((n + 255) / 256);
    if (e.cur64) {
        hipLaunchKernelGGL((adj_convert<double, double>), dim3(cb), dim3(256), 0, s,
                           (const double *)e.d_gamma64[par] + D.goff, gam, n);
    } else if (e.precision == PASC_FP32_BULK) {
        hipLaunchKernelGGL((adj_convert<float, double>), dim3(cb), dim3(256), 0, s,
                           (const float *)e.d_gamma[par] + D.goff, gam, n);
    } else {
        hipLaunchKernelGGL((adj_convert<double, double>), dim3(cb), dim3(256), 0, s,
                           (const double *)e.d_gamma[par] + D.goff, gam, n);
    }
    HIP_TRY(hipMemcpyAsync(mk, mask_host, (size_t)T, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(buf + n, 0, sizeof(double) * 5 * (size_t)n, s));
    AdjArgs a;
    a.gam = gam;
    a.x = (const double *)e.d_x64 + D.xoff;
    a.mask = mk;
    a.qmask = D.heldout;
    a.freeset = fr;
    a.T = T;
    a.ld = ld;
    a.K = D.K;
    a.xdim = D.xdim;
    a.alpha = D.alpha;
    a.bscale = D.bscale;
    for (int i = 0; i < NTH; i++) a.theta[i] = i < D.K * D.xdim ? D.theta[i] : 0.0;
    a.G = e.stencil == STENCIL_TIME ? GraphGeo{T, 1, 0, nullptr, nullptr} : e.geo;

    std::vector<double> v;
    hipLaunchKernelGGL(adj_init, dim3(blocks), dim3(ANT), 0, s, a, r, part);
    HIP_TRY(hipGetLastError());
    if ((rc = sum_partials(s, part, blocks, 2 + NTH, v))) { cleanup(); return rc; }
    *loss = v[0];
    std::vector<double> direct(v.begin() + 2, v.begin() + 2 + D.K * D.xdim);
    HIP_TRY(hipMemcpyAsync(pb[0], r, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    // conjugate gradients on P A P (symmetric positive semi-definite on range(P)), the scalars
    // on the device; the host reads the stop flag every CG_POLL iterations
    CgState h0{};  // the initial state; stays alive until the copies below have completed
    h0.rr = h0.rr0 = v[1];
    h0.tol2 = cg_tol * cg_tol;
    h0.maxit = cg_maxit;
    h0.done = (cg_maxit > 0 && h0.rr > h0.tol2 * h0.rr0 && h0.rr > 0.0) ? 0 : 1;
    HIP_TRY(hipMemcpyAsync(d_cs, &h0, sizeof(CgState), hipMemcpyHostToDevice, s));
    CgState h = h0;
    for (int launched = 0; !h.done && launched <= cg_maxit + CG_POLL; launched += CG_POLL) {
        for (int i = 0; i < CG_POLL; i++) {
            // step launched + i reads p from pb[(launched + i) & 1] and writes the other
            const double *pold = pb[i & 1];
            double *pnew = pb[(i + 1) & 1];
            hipLaunchKernelGGL(adj_matvec, dim3(blocks), dim3(ANT), 0, s, a, r, pold, pnew, q,
                               part, d_cs);
            hipLaunchKernelGGL(adj_cg_step, dim3(1), dim3(RNT), 0, s, part, blocks, d_cs);
            hipLaunchKernelGGL(adj_axpy, dim3(blocks), dim3(ANT), 0, s, a, d_cs, pnew, q, lam, r,
                               part);
            hipLaunchKernelGGL(adj_cg_rr, dim3(1), dim3(RNT), 0, s, part, blocks, d_cs);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&h, d_cs, sizeof(CgState), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    const int it = h.it;
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(adj_final, dim3(blocks), dim3(ANT), 0, s, a, lam, part);
    if ((rc = sum_partials(s, part, blocks, 1 + NTH, v))) { cleanup(); return rc; }
    *dlogeps = -v[0];
    for (int i = 0; i < D.K * D.xdim; i++) dtheta[i] = direct[i] + v[1 + i];
    if (cg_it) *cg_it = it;
    cleanup();
    if (!std::isfinite(*dlogeps)) return fail(PASC_ERR_NUMERIC, "adjoint: non-finite gradient");
    return PASC_OK;
}

}  // namespace pasc

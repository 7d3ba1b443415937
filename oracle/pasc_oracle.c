/*
 * pasc_oracle.c -- FP64 CPU restatement of the PASC_inference FEM-H1 K-means gamma/theta path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path, and the timed CPU baseline).  The
 * structure deliberately stays op-by-op and unfused like the PETSc specialisation it restates
 * (src/petscvector/solver/spgqpsolver.cpp:55-367): every vector operation is its own pass.
 * Parity pins for each routine are listed in pasc_oracle.h.
 */
#include "pasc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define OMP_FOR _Pragma("omp parallel for schedule(static) num_threads(nthreads)")

int or_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static int fix_threads(int n) { return n < 1 ? 1 : n; }

/* Sums over many threads are formed as one partial per thread (static schedule) added in
 * thread order, so that a result does not depend on the order in which threads finish (an
 * OpenMP reduction clause leaves that order open, which made SPG trajectories vary run to
 * run). */
#define OR_MAX_THREADS 256
#define DET_SUM_BEGIN(nthreads)                                                              \
    double part_[OR_MAX_THREADS] = {0.0};                                                    \
    if ((nthreads) > OR_MAX_THREADS) (nthreads) = OR_MAX_THREADS;                           \
    _Pragma("omp parallel num_threads(nthreads)") {                                          \
        double s_ = 0.0;
#define DET_SUM_END(result)                                                                  \
        part_[omp_get_thread_num()] = s_;                                                    \
    }                                                                                        \
    (result) = 0.0;                                                                          \
    for (int i_ = 0; i_ < OR_MAX_THREADS; i_++) (result) += part_[i_];

/* ------------------------------------------------------------------------------------------ */
/* simplex projection                                                                         */
/* ------------------------------------------------------------------------------------------ */

/* simplex_local.h:153-174: bubble sort, ascending, with the "last swap" shortening. */
void or_sort_bubble(double *x, int n) {
    int m = n;
    while (m > 0) {
        int mnew = 0;
        for (int i = 1; i < m; i++) {
            if (x[i] < x[i - 1]) {
                double s = x[i];
                x[i] = x[i - 1];
                x[i - 1] = s;
                mnew = i;
            }
        }
        m = mnew;
    }
}

/* simplex_local.h:176-249 (project_sub): skip exactly-feasible blocks, otherwise sort and use
 * the Chen & Ye (2011) threshold t_hat, then clamp x - t_hat at zero. */
static void project_block(double *xb, int K, double *y) {
    int inside = 1;
    double sum = 0.0;
    for (int k = 0; k < K; k++) {
        if (xb[k] < 0.0) inside = 0;
        sum += xb[k];
    }
    if (sum != 1) inside = 0;
    if (inside) return;

    for (int k = 0; k < K; k++) y[k] = xb[k];
    or_sort_bubble(y, K);

    double t_hat = 0.0;
    int i = K - 1;
    while (i >= 1) {
        double sum_y = 0.0;
        for (int j = i; j < K; j++) sum_y += y[j];
        double ti = (sum_y - 1.0) / (double)(K - i);
        if (ti >= y[i - 1]) {
            t_hat = ti;
            i = -1;
        } else {
            i = i - 1;
        }
    }
    if (i == 0) t_hat = (sum - 1.0) / (double)K;

    for (int k = 0; k < K; k++) {
        double ti = xb[k] - t_hat;
        xb[k] = ti > 0.0 ? ti : 0.0;
    }
}

void or_project_simplex(double *x, int64_t nblocks, int K, int nthreads) {
    nthreads = fix_threads(nthreads);
#pragma omp parallel num_threads(nthreads)
    {
        double *y = (double *)malloc(sizeof(double) * (size_t)K);
#pragma omp for schedule(static)
        for (int64_t t = 0; t < nblocks; t++) project_block(x + t * K, K, y);
        free(y);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* H1 graph-time matvec                                                                       */
/* ------------------------------------------------------------------------------------------ */

/* blockgraphsparse.cpp:57-73 */
int or_h1_wsum(int64_t t, int64_t T, int deg) {
    if (t == 0 || t == T - 1) {
        if (T > 1) return 2 * deg + 2;
        return deg;
    }
    if (T > 1) return 3 * deg + 4;
    return deg + 1; /* unreachable: t is interior only when T > 2 */
}

/* Row (t, r, k) of A0 (blockgraphsparse.cpp:75-104):
 *   diagonal Wsum; time neighbours t-1, t+1 of the same node carry -2;
 *   every spatial neighbour s carries -1 at t, t-1 and t+1. */
void or_h1_matvec(const double *x, double *y, int64_t T, int64_t R, int K, double alpha,
                  const int64_t *nbr_ptr, const int32_t *nbr_idx, int nthreads) {
    nthreads = fix_threads(nthreads);
    const int64_t RK = R * K;
    OMP_FOR
    for (int64_t t = 0; t < T; t++) {
        for (int64_t r = 0; r < R; r++) {
            const int64_t beg = nbr_ptr ? nbr_ptr[r] : 0;
            const int64_t end = nbr_ptr ? nbr_ptr[r + 1] : 0;
            const int deg = (int)(end - beg);
            const double wsum = (double)or_h1_wsum(t, T, deg);
            for (int k = 0; k < K; k++) {
                const int64_t i = t * RK + r * K + k;
                double acc = wsum * x[i];
                if (T > 1) {
                    if (t > 0) acc += -2.0 * x[i - RK];
                    if (t < T - 1) acc += -2.0 * x[i + RK];
                }
                for (int64_t e = beg; e < end; e++) {
                    const int64_t i2 = t * RK + (int64_t)nbr_idx[e] * K + k;
                    acc += -1.0 * x[i2];
                    if (t > 0) acc += -1.0 * x[i2 - RK];
                    if (t < T - 1) acc += -1.0 * x[i2 + RK];
                }
                y[i] = alpha * acc; /* MatMult then VecScale(alpha), spgqpsolver.cpp:122-123 */
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* model hooks                                                                                */
/* ------------------------------------------------------------------------------------------ */

void or_assemble_b(const double *x, const double *theta, double *b, int64_t T, int64_t R, int K,
                   int xdim, int scalef, const uint8_t *heldout, int nthreads) {
    nthreads = fix_threads(nthreads);
    /* graphh1fem.cpp:280-286: coeff = -1/T_reduced when scalef */
    const double coeff = scalef ? -1.0 / (double)T : -1.0;
    OMP_FOR
    for (int64_t t = 0; t < T; t++) {
        for (int64_t r = 0; r < R; r++) {
            const int64_t node = t * R + r;
            const double *xr = x + node * xdim;
            for (int k = 0; k < K; k++) {
                double v = 0.0;
                if (!(heldout && heldout[node])) {
                    for (int c = 0; c < xdim; c++) {
                        const double df = xr[c] - theta[k * xdim + c];
                        v += df * df; /* graphh1fem.cpp:261 for xdim == 1 */
                    }
                }
                b[node * K + k] = coeff * v;
            }
        }
    }
}

void or_theta_update(const double *x, const double *gamma, double *theta, int64_t T, int64_t R,
                     int K, int xdim, const uint8_t *heldout, int nthreads) {
    nthreads = fix_threads(nthreads);
    const int64_t nodes = T * R;
    for (int k = 0; k < K; k++) {
        double gsum;
        /* VecSum(gamma_k), graphh1fem.cpp:364 */
        {
            DET_SUM_BEGIN(nthreads)
#pragma omp for schedule(static)
            for (int64_t i = 0; i < nodes; i++) {
                if (heldout && heldout[i]) continue;
                s_ += gamma[i * K + k];
            }
            DET_SUM_END(gsum)
        }
        for (int c = 0; c < xdim; c++) {
            double gx;
            /* VecDot(data, gamma_k), graphh1fem.cpp:361 */
            {
                DET_SUM_BEGIN(nthreads)
#pragma omp for schedule(static)
                for (int64_t i = 0; i < nodes; i++) {
                    if (heldout && heldout[i]) continue;
                    s_ += x[i * xdim + c] * gamma[i * K + k];
                }
                DET_SUM_END(gx)
            }
            /* graphh1fem.cpp:375-379 */
            theta[k * xdim + c] = (gsum != 0) ? gx / gsum : 0.0;
        }
    }
}

double or_kkt_residual(const double *gamma, const double *g, int64_t nblocks, int K, double s,
                       int nthreads) {
    nthreads = fix_threads(nthreads);
    double r = 0.0;
#pragma omp parallel num_threads(nthreads)
    {
        double *z = (double *)malloc(sizeof(double) * (size_t)K);
        double *y = (double *)malloc(sizeof(double) * (size_t)K);
        double rl = 0.0;
#pragma omp for schedule(static)
        for (int64_t t = 0; t < nblocks; t++) {
            const double *gb = gamma + t * K;
            const double *gg = g + t * K;
            for (int k = 0; k < K; k++) z[k] = gb[k] - s * gg[k];
            project_block(z, K, y);
            for (int k = 0; k < K; k++) {
                const double e = fabs(gb[k] - z[k]);
                if (e > rl || e != e) rl = e;
            }
        }
#pragma omp critical
        {
            if (rl > r || rl != rl) r = rl;
        }
        free(z);
        free(y);
    }
    return r;
}

/* ------------------------------------------------------------------------------------------ */
/* vector helpers (one pass each, like the PETSc Vec calls they stand for)                    */
/* ------------------------------------------------------------------------------------------ */

static void v_copy(double *dst, const double *src, int64_t n, int nthreads) {
    OMP_FOR
    for (int64_t i = 0; i < n; i++) dst[i] = src[i];
}

static void v_axpy(double *y, double a, const double *x, int64_t n, int nthreads) {
    OMP_FOR
    for (int64_t i = 0; i < n; i++) y[i] += a * x[i];
}

static double v_dot(const double *a, const double *b, int64_t n, int nthreads) {
    double s;
    DET_SUM_BEGIN(nthreads)
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; i++) s_ += a[i] * b[i];
    DET_SUM_END(s)
    return s;
}

/* sum_t sum_k (g_kt - g_0t) d_kt: the value of dot(g, d) when every block of d sums to 0 */
static double dot_centered(const double *d, const double *g, int64_t nblocks, int K,
                           int nthreads) {
    double s;
    DET_SUM_BEGIN(nthreads)
#pragma omp for schedule(static)
    for (int64_t t = 0; t < nblocks; t++) {
        const double c = g[t * K];
        for (int k = 1; k < K; k++) s_ += (g[t * K + k] - c) * d[t * K + k];
    }
    DET_SUM_END(s)
    return s;
}

/* gd of the projected-gradient direction in a form whose sign survives rounding (the
 * gd_certified option): d = P(y) - x with y = x - a g, so g = (x - y)/a and
 * <g, d> = -(|d|^2 + sum_{k: z_k = 0} (tau - y_k) x_k) / a, tau the threshold of P(y). */
static double dot_certified(const double *x, const double *g, const double *d, int64_t nblocks,
                            int K, double a, int nthreads) {
    double s;
    DET_SUM_BEGIN(nthreads)
    double *y = (double *)malloc(sizeof(double) * (size_t)K * 3);
    double *z = y + K, *w = y + 2 * K;
#pragma omp for schedule(static)
    for (int64_t t = 0; t < nblocks; t++) {
        int m = 0;
        for (int k = 0; k < K; k++) {
            y[k] = x[t * K + k] - a * g[t * K + k];
            z[k] = y[k];
        }
        project_block(z, K, w);
        for (int k = 1; k < K; k++)
            if (z[k] > z[m]) m = k;
        const double tau = y[m] - z[m];
        double o = 0.0;
        for (int k = 0; k < K; k++) {
            const double dk = d[t * K + k];
            o += dk * dk;
            if (y[k] <= tau) o += (tau - y[k]) * x[t * K + k];
        }
        s_ += o;
    }
    free(y);
    DET_SUM_END(s)
    return -s / a;
}

/* d'A0 d as a sum over the edges of A0 (the dad_edges option): each node adds half of
 * w (d_i - d_j)^2 over its neighbours j, with the weights of or_h1_matvec's rows. */
static double dad_by_edges(const double *d, int64_t T, int64_t R, int K, const int64_t *nbr_ptr,
                           const int32_t *nbr_idx, int nthreads) {
    const int64_t RK = R * K;
    double s;
    DET_SUM_BEGIN(nthreads)
#pragma omp for schedule(static)
    for (int64_t t = 0; t < T; t++) {
        for (int64_t r = 0; r < R; r++) {
            const int64_t beg = nbr_ptr ? nbr_ptr[r] : 0;
            const int64_t end = nbr_ptr ? nbr_ptr[r + 1] : 0;
            for (int k = 0; k < K; k++) {
                const int64_t i = t * RK + r * K + k;
                double acc = 0.0, e;
                if (t > 0) { e = d[i] - d[i - RK]; acc += 2.0 * e * e; }
                if (t < T - 1) { e = d[i] - d[i + RK]; acc += 2.0 * e * e; }
                for (int64_t q = beg; q < end; q++) {
                    const int64_t i2 = t * RK + (int64_t)nbr_idx[q] * K + k;
                    e = d[i] - d[i2]; acc += e * e;
                    if (t > 0) { e = d[i] - d[i2 - RK]; acc += e * e; }
                    if (t < T - 1) { e = d[i] - d[i2 + RK]; acc += e * e; }
                }
                s_ += 0.5 * acc;
            }
        }
    }
    DET_SUM_END(s)
    return s;
}

/* spg_fs.cpp:17-63 */
typedef struct {
    int m;
    int last_idx;
    double *list;
} spg_fs;

static void fs_init(spg_fs *fs, double fx) {
    for (int i = 0; i < fs->m; i++) fs->list[i] = fx;
    fs->last_idx = 0;
}

static double fs_get_max(const spg_fs *fs) {
    double mv = fs->list[0];
    for (int i = 0; i < fs->m; i++)
        if (fs->list[i] > mv) mv = fs->list[i];
    return mv;
}

static void fs_update(spg_fs *fs, double fx) {
    fs->last_idx++;
    if (fs->last_idx >= fs->m) fs->last_idx = 0;
    fs->list[fs->last_idx] = fx;
}

void or_spg_default_params(or_spg_params *p) {
    memset(p, 0, sizeof(*p));
    p->maxit = 1000;
    p->eps = 1e-9;
    p->m = 20;
    p->gamma = 0.9;
    p->sigma1 = 0.0;
    p->sigma2 = 1.0;
    p->alphainit = 2.0;
    p->stop_difff = 1;
    p->kkt_tol = 1e-6;
    p->kkt_scale = 1.0;
    p->nthreads = 1;
}

/* get_fx, spgqpsolver.cpp:316-345: temp = g - b; fx = 0.5 * dot(x, temp) */
static double spg_get_fx(const double *x, const double *g, const double *b, double *temp,
                         int64_t n, int nthreads) {
    v_copy(temp, g, n, nthreads);
    v_axpy(temp, -1.0, b, n, nthreads);
    return 0.5 * v_dot(x, temp, n, nthreads);
}

int or_spg_solve(double *x, const double *b, int64_t T, int64_t R, int K, double alpha,
                 const int64_t *nbr_ptr, const int32_t *nbr_idx, const or_spg_params *p,
                 or_spg_result *res) {
    const int nthreads = fix_threads(p->nthreads);
    const int64_t n = T * R * K;
    const int64_t nblocks = T * R;
    double *g = (double *)malloc(sizeof(double) * (size_t)n);
    double *d = (double *)malloc(sizeof(double) * (size_t)n);
    double *Ad = (double *)malloc(sizeof(double) * (size_t)n);
    double *temp = (double *)malloc(sizeof(double) * (size_t)n);
    spg_fs fs;
    fs.m = p->m;
    fs.list = (double *)malloc(sizeof(double) * (size_t)(p->m > 0 ? p->m : 1));
    if (!g || !d || !Ad || !temp || !fs.list) {
        free(g); free(d); free(Ad); free(temp); free(fs.list);
        return -1;
    }

    int it = 0, hessmult = 0, reason = 0;
    double fx, fx_old, fx_max, xi, beta_bar, beta_hat, beta;
    double dd = DBL_MAX, gd = 0.0, dAd = 0.0, alpha_bb, kkt = DBL_MAX;

    const double normb = sqrt(v_dot(b, b, n, nthreads)); /* VecNorm(b) :104 */
    alpha_bb = p->alphainit;

    or_project_simplex(x, nblocks, K, nthreads); /* :116 */

    or_h1_matvec(x, g, T, R, K, alpha, nbr_ptr, nbr_idx, nthreads); /* :122-123 */
    hessmult += 1;
    v_axpy(g, -1.0, b, n, nthreads); /* :128 */

    fx = spg_get_fx(x, g, b, temp, n, nthreads); /* :133-136 */
    fx_old = DBL_MAX;
    fs_init(&fs, fx);
    (void)fx_old;

    while (it < p->maxit) {
        it += 1;
        /* d = P(x - alpha_bb*g) - x  (:146-159) */
        v_copy(d, x, n, nthreads);
        v_axpy(d, -alpha_bb, g, n, nthreads);
        or_project_simplex(d, nblocks, K, nthreads);
        v_axpy(d, -1.0, x, n, nthreads);

        /* Ad = alpha*A0*d (:164-167) */
        or_h1_matvec(d, Ad, T, R, K, alpha, nbr_ptr, nbr_idx, nthreads);
        hessmult += 1;

        /* compute_dots :348-367 */
        dd = v_dot(d, d, n, nthreads);
        dAd = p->dad_edges ? alpha * dad_by_edges(d, T, R, K, nbr_ptr, nbr_idx, nthreads)
                           : v_dot(d, Ad, n, nthreads);
        if (p->gd_certified)
            gd = dot_certified(x, g, d, nblocks, K, alpha_bb, nthreads);
        else
            gd = p->gd_centered ? dot_centered(d, g, nblocks, K, nthreads) : v_dot(d, g, n, nthreads);

        fx_max = fs_get_max(&fs); /* :176 */

        /* step-size from the generalised Armijo condition (:182-195) */
        if (p->safeguard && !(dAd > 0.0)) {
            beta = p->sigma2;
        } else {
            xi = (fx_max - fx) / dAd;
            beta_bar = -gd / dAd;
            beta_hat = p->gamma * beta_bar + sqrt(p->gamma * p->gamma * beta_bar * beta_bar + 2 * xi);
            if (beta_hat < p->sigma1) beta_hat = p->sigma1;
            if (beta_hat < p->sigma2)
                beta = beta_hat;
            else
                beta = p->sigma2;
        }

        /* x += beta*d; g += beta*Ad (:200-201) */
        v_axpy(x, beta, d, n, nthreads);
        if (p->g_recompute) {
            or_h1_matvec(x, g, T, R, K, alpha, nbr_ptr, nbr_idx, nthreads);
            v_axpy(g, -1.0, b, n, nthreads);
        } else {
            v_axpy(g, beta, Ad, n, nthreads);
        }

        /* :207-211 */
        fx_old = fx;
        if (p->fx_identity)
            fx = fx_old + beta * (gd + 0.5 * beta * dAd); /* spgqpsolver.h:835-842 */
        else
            fx = spg_get_fx(x, g, b, temp, n, nthreads);
        fs_update(&fs, fx);

        /* :216 */
        alpha_bb = (p->safeguard && !(dAd > 0.0)) ? p->alphainit : dd / dAd;

        /* stopping criteria :260-274 */
        if (p->stop_difff && fabs(fx - fx_old) < p->eps) { reason = 1; break; }
        if (p->stop_normgp && dd < p->eps) { reason = 2; break; }
        if (p->stop_normgp_normb && dd < p->eps * normb) { reason = 3; break; }
        if (p->stop_Anormgp && dAd < p->eps) { reason = 4; break; }
        if (p->stop_Anormgp_normb && dAd < p->eps * normb) { reason = 5; break; }
        if (p->stop_kkt) {
            kkt = or_kkt_residual(x, g, nblocks, K, p->kkt_scale, nthreads);
            if (kkt <= p->kkt_tol) { reason = 6; break; }
        }
    }

    if (reason != 6) kkt = or_kkt_residual(x, g, nblocks, K, p->kkt_scale, nthreads);

    if (res) {
        res->it = it;
        res->hessmult = hessmult;
        res->fx = fx;
        res->dd = dd;
        res->dAd = dAd;
        res->kkt = kkt;
        res->normb = normb;
        res->stop_reason = reason;
    }
    free(g); free(d); free(Ad); free(temp); free(fs.list);
    return 0;
}

double or_objective(const double *x, const double *b, int64_t T, int64_t R, int K, double alpha,
                    const int64_t *nbr_ptr, const int32_t *nbr_idx, int nthreads) {
    nthreads = fix_threads(nthreads);
    const int64_t n = T * R * K;
    double *ax = (double *)malloc(sizeof(double) * (size_t)n);
    if (!ax) return NAN;
    or_h1_matvec(x, ax, T, R, K, alpha, nbr_ptr, nbr_idx, nthreads);
    double q = v_dot(x, ax, n, nthreads);
    double l = v_dot(x, b, n, nthreads);
    free(ax);
    return 0.5 * q - l;
}

/* ------------------------------------------------------------------------------------------ */
/* outer alternation                                                                          */
/* ------------------------------------------------------------------------------------------ */

/* One run of the alternation (tssolver.h:542-650): the theta step (unless theta is fixed, or
 * skipped on the first iteration), the gamma step, L = fx of the gamma solve (get_L), and the
 * stop at the first |L - L_old| < eps.  The reference breaks before its it++, so it_last is the
 * index of the breaking iteration, or maxit when the loop runs out. */
typedef struct {
    int it;        /* it_last */
    int spg;       /* SPG iterations of every gamma solve of the run (QPSolver::it_sum) */
    double L, deltaL;
    int increased;
} ts_run_out;

static int ts_run_fem(const double *x, double *gamma, double *theta, int64_t T, int K, int xdim,
                      double epssqr, const uint8_t *heldout, const or_ts_params *tp,
                      const or_spg_params *sp, int skip_first, ts_run_out *o);

int64_t or_fem_reduced_T(int64_t T1, double fem_reduce) {
    /* fem.cpp:198 / femhat.cpp:224: T_reduced = ceil(T * fem_reduce) when reduced */
    if (fem_reduce >= 1.0) return T1;
    return (int64_t)ceil((double)T1 * fem_reduce);
}

void or_fem_reduce(const double *in, double *out, int64_t T1, int64_t T2, int K, int type) {
    if (type == 1) {
        const double diff = (double)T1 / (double)T2;            /* fem.cpp:16 */
        for (int k = 0; k < K; k++)
            for (int64_t t2 = 0; t2 < T2; t2++) {
                double s = 0.0;
                for (int64_t i = (int64_t)round((double)t2 * diff);
                     i < (int64_t)round((double)(t2 + 1) * diff); i++)
                    s += in[i * K + k];
                out[t2 * K + k] = s;
            }
        return;
    }
    const double diff = (double)(T1 - 1) / (double)(T2 - 1);    /* femhat.cpp:15 */
    const int64_t right_idx = (int64_t)floor(diff * (double)(T2 - 1));  /* femhat.h:179 */
    for (int k = 0; k < K; k++)
        for (int64_t t2 = 0; t2 < T2; t2++) {
            const double center = (double)t2 * diff;
            const double left = (double)(t2 - 1) * diff;
            const double right = (double)(t2 + 1) * diff;
            double s = 0.0;
            int64_t t1 = (int64_t)floor(left);
            for (; (double)t1 <= center; t1++)
                if (t1 >= 0) s += ((double)t1 - left) / (center - left) * in[t1 * K + k];
            for (; (double)t1 < right; t1++)
                if (t1 < right_idx) s += ((double)t1 - right) / (center - right) * in[t1 * K + k];
            out[t2 * K + k] = s;
        }
}

void or_fem_prolongate(const double *in, double *out, int64_t T1, int64_t T2, int K, int type) {
    if (type == 1) {
        const double diff = (double)T1 / (double)T2;
        for (int k = 0; k < K; k++)
            for (int64_t t2 = 0; t2 < T2; t2++)
                for (int64_t i = (int64_t)round((double)t2 * diff);
                     i < (int64_t)round((double)(t2 + 1) * diff); i++)
                    out[i * K + k] = in[t2 * K + k];
        return;
    }
    const double diff = (double)(T1 - 1) / (double)(T2 - 1);
    for (int k = 0; k < K; k++)
        for (int64_t t1 = 0; t1 < T1; t1++) {
            const int64_t tl = (int64_t)floor((double)t1 / diff);
            const int64_t tr = tl + 1;
            const double t1l = (double)tl * diff, t1r = (double)tr * diff;
            const double vr = in[(tr < T2 ? tr : T2 - 1) * K + k];
            const double vl = in[(tl < T2 ? tl : T2 - 1) * K + k];
            out[t1 * K + k] = ((double)t1 - t1l) / (t1r - t1l) * vr +
                              ((double)t1 - t1r) / (t1l - t1r) * vl;
        }
}

static int ts_run_plain(const double *x, double *gamma, double *theta, int64_t T, int64_t R,
                        int K, int xdim, double epssqr, const int64_t *nbr_ptr,
                        const int32_t *nbr_idx, const uint8_t *heldout, const or_ts_params *tp,
                        const or_spg_params *sp, int skip_first, ts_run_out *o) {
    const int nthreads = fix_threads(sp->nthreads);
    const int64_t n = T * R * K;
    double *b = (double *)malloc(sizeof(double) * (size_t)n);
    if (!b) return -1;
    /* graphh1fem.cpp:146-151 with no FEM reduction */
    const double alpha = tp->scalef ? epssqr / (double)T : epssqr;
    double L = DBL_MAX, L_old, deltaL = DBL_MAX;
    int it, spg_sum = 0, increased = 0;
    for (it = 0; it < tp->maxit; it++) {
        /* theta step (tssolver.h:548-560) */
        if (!tp->theta_fixed && !(it == 0 && skip_first))
            or_theta_update(x, gamma, theta, T, R, K, xdim, heldout, nthreads);
        /* gamma step (tssolver.h:584-596) */
        or_assemble_b(x, theta, b, T, R, K, xdim, tp->scalef, heldout, nthreads);
        or_spg_result r;
        if (or_spg_solve(gamma, b, T, R, K, alpha, nbr_ptr, nbr_idx, sp, &r) != 0) {
            free(b);
            return -1;
        }
        spg_sum += r.it;
        /* tssolver.h:628-650 */
        L_old = L;
        L = r.fx;
        deltaL = fabs(L - L_old);
        if (L - L_old > 0) increased++;
        if (deltaL < tp->eps) break;
    }
    o->it = it;
    o->spg = spg_sum;
    o->L = L;
    o->deltaL = deltaL;
    o->increased = increased;
    free(b);
    return 0;
}

void or_uniform(uint64_t seed, int64_t n, double *out) {
    uint64_t s = seed;
    for (int64_t i = 0; i < n; i++) {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        out[i] = (double)(z >> 11) * 0x1.0p-53;
    }
}

/* TSSolver::solve (tssolver.h:491-715): `annealing` runs of the alternation.  Each run ends with
 * AIC = 2 log L + K (graphh1fem.h:443-447); with more than one run, the gamma and theta of the
 * run with the smallest AIC are kept (tssolver.h:669-692, the temp vectors), and every run after
 * the first starts from the random gamma of set_random (tssolver.h:687-689).  skip_first_theta
 * applies to the first run only (the random restarts begin with a theta step, as in the
 * reference). */
int or_ts_solve(const double *x, double *gamma, double *theta, int64_t T, int64_t R, int K,
                int xdim, double epssqr, const int64_t *nbr_ptr, const int32_t *nbr_idx,
                const uint8_t *heldout, const or_ts_params *tp, const or_spg_params *sp,
                or_ts_result *res) {
    const int fem = tp->fem_reduce > 0.0 && tp->fem_reduce < 1.0 && tp->fem_type > 0;
    if (fem && R != 1) return -1;
    const int runs = tp->annealing > 1 ? tp->annealing : 1;
    const int64_t n = T * R * K;
    const int nth = K * xdim;
    double *gbest = NULL, *tbest = NULL;
    if (runs > 1) {
        gbest = (double *)malloc(sizeof(double) * (size_t)n);
        tbest = (double *)malloc(sizeof(double) * (size_t)nth);
        if (!gbest || !tbest) { free(gbest); free(tbest); return -1; }
    }
    double best_aic = DBL_MAX, L_keep = DBL_MAX, dL_keep = DBL_MAX;
    int it_last = 0, it_sum = 0, spg_sum = 0, increased = 0, best_run = 0;
    for (int a = 0; a < runs; a++) {
        ts_run_out o;
        const int skip = a == 0 ? tp->skip_first_theta : 0;
        const int rc = fem ? ts_run_fem(x, gamma, theta, T, K, xdim, epssqr, heldout, tp, sp, skip, &o)
                           : ts_run_plain(x, gamma, theta, T, R, K, xdim, epssqr, nbr_ptr, nbr_idx,
                                          heldout, tp, sp, skip, &o);
        if (rc) { free(gbest); free(tbest); return -1; }
        it_last = o.it;
        it_sum += o.it;
        spg_sum += o.spg;
        increased += o.increased;
        const double aic = 2.0 * log(o.L) + (double)K;
        if (runs <= 1 || aic < best_aic) {
            if (runs > 1) {
                memcpy(gbest, gamma, sizeof(double) * (size_t)n);
                memcpy(tbest, theta, sizeof(double) * (size_t)nth);
            }
            best_aic = aic;
            L_keep = o.L;
            dL_keep = o.deltaL;
            best_run = a;
        }
        if (a < runs - 1) or_uniform(tp->restart_seed, n, gamma);
    }
    if (runs > 1) {
        memcpy(gamma, gbest, sizeof(double) * (size_t)n);
        memcpy(theta, tbest, sizeof(double) * (size_t)nth);
    }
    free(gbest);
    free(tbest);
    if (res) {
        res->it = it_last;
        res->spg_it_sum = spg_sum;
        res->L = L_keep;
        res->deltaL = dL_keep;
        res->increased = increased;
        res->aic = best_aic;
        res->it_sum = it_sum;
        res->best_run = best_run;
    }
    return 0;
}

double or_abserr_reconstructed(const double *gamma, const double *theta,
                               const double *solution, int64_t T, int64_t R, int K) {
    double s = 0.0;
    for (int64_t i = 0; i < T * R; i++) {
        double rec = -solution[i];
        for (int k = 0; k < K; k++) rec += theta[k] * gamma[i * K + k];
        s += fabs(rec);
    }
    return s;
}

double or_host_copy_bandwidth(int64_t nbytes, int reps, int nthreads) {
    nthreads = fix_threads(nthreads);
    const int64_t n = nbytes / 8;
    double *a = (double *)malloc(sizeof(double) * (size_t)n);
    double *c = (double *)malloc(sizeof(double) * (size_t)n);
    if (!a || !c) { free(a); free(c); return 0.0; }
    OMP_FOR
    for (int64_t i = 0; i < n; i++) { a[i] = (double)i; c[i] = 0.0; }
    double best = 1e300;
    for (int r = 0; r < reps; r++) {
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        OMP_FOR
        for (int64_t i = 0; i < n; i++) c[i] = a[i];
        clock_gettime(CLOCK_MONOTONIC, &t1);
        double dt = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
        if (dt < best) best = dt;
    }
    volatile double sink = c[n / 2];
    (void)sink;
    free(a);
    free(c);
    return 2.0 * (double)n * 8.0 / best;
}

/* TSSolver alternation with FEM coarsening (graphh1fem.cpp:115-214,236-301): the residual
 * (x_t - theta_k)^2 (graphh1fem.cpp:261) and the fine gamma are reduced to the coarse grid
 * (:274-277), b = -reduced residual / T2 with scalef (:281-286), A = epssqr/T * A0 on the coarse
 * grid with scalef, epssqr*T2/T without (:146-151), the coarse solution is prolongated back
 * (:296-298); theta is updated from the fine gamma.  Held-out steps carry no residual. */
static int ts_run_fem(const double *x, double *gamma, double *theta, int64_t T, int K, int xdim,
                      double epssqr, const uint8_t *heldout, const or_ts_params *tp,
                      const or_spg_params *sp, int skip_first, ts_run_out *o) {
    const int nthreads = fix_threads(sp->nthreads);
    const int64_t T2 = or_fem_reduced_T(T, tp->fem_reduce);
    if (tp->fem_type == 2 && T2 < 2) return -1;
    double *resid = (double *)malloc(sizeof(double) * (size_t)(T * K));
    double *b2 = (double *)malloc(sizeof(double) * (size_t)(T2 * K));
    double *g2 = (double *)malloc(sizeof(double) * (size_t)(T2 * K));
    if (!resid || !b2 || !g2) { free(resid); free(b2); free(g2); return -1; }
    const double alpha = tp->scalef ? epssqr / (double)T : epssqr * (double)T2 / (double)T;
    const double bs = tp->scalef ? -1.0 / (double)T2 : -1.0;
    double L = DBL_MAX, L_old, deltaL = DBL_MAX;
    int it, spg_sum = 0, increased = 0;
    for (it = 0; it < tp->maxit; it++) {
        if (!tp->theta_fixed && !(it == 0 && skip_first))
            or_theta_update(x, gamma, theta, T, 1, K, xdim, heldout, nthreads);
        for (int64_t t = 0; t < T; t++)
            for (int k = 0; k < K; k++) {
                double q = 0.0;
                if (!(heldout && heldout[t]))
                    for (int c = 0; c < xdim; c++) {
                        const double dx = x[t * xdim + c] - theta[k * xdim + c];
                        q += dx * dx;
                    }
                resid[t * K + k] = q;
            }
        or_fem_reduce(resid, b2, T, T2, K, tp->fem_type);
        for (int64_t i = 0; i < T2 * K; i++) b2[i] *= bs;
        or_fem_reduce(gamma, g2, T, T2, K, tp->fem_type);
        or_spg_result r;
        if (or_spg_solve(g2, b2, T2, 1, K, alpha, NULL, NULL, sp, &r) != 0) {
            free(resid); free(b2); free(g2);
            return -1;
        }
        or_fem_prolongate(g2, gamma, T, T2, K, tp->fem_type);
        spg_sum += r.it;
        L_old = L;
        L = r.fx;
        deltaL = fabs(L - L_old);
        if (L - L_old > 0) increased++;
        if (deltaL < tp->eps) break;
    }
    o->it = it;
    o->spg = spg_sum;
    o->L = L;
    o->deltaL = deltaL;
    o->increased = increased;
    free(resid); free(b2); free(g2);
    return 0;
}

// C ABI of libpasc_hip.so (declared in include/pasc_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "device_common.h"

namespace pasc {

static thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

void clear_error() { g_err.clear(); }

}  // namespace pasc

using namespace pasc;

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(PASC_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

struct pasc_ctx {
    Ctx c;
};

struct pasc_problem {
    pasc_ctx *ctx = nullptr;
    int64_t T = 0, R = 0;
    int K = 0, xdim = 1, scalef = 1, precision = PASC_FP64;
    int32_t grid_width = 0;
    double epssqr = 1.0;
    bool theta_set = false;
    bool gamma_set = false;
    bool solved = false;
    bool b_set = false;  // explicit linear term instead of the data term
    bool has_data = true;  // false: created without x (only an explicit b can define the QP)
    uint8_t *d_heldout = nullptr;
    std::vector<uint8_t> h_heldout;
    std::vector<double> h_x;  // host copy of the data, [T][R][xdim] (for abserr etc.)
    Engine e;
    // FEM coarsening (1-D): the engine holds the coarse QP of T2 steps, the fine grid of T
    // steps lives here (data, gamma) and meets the engine through reduce / prolongate
    bool graph = false;        // node-wise graph stencil (R > 1, image grid or CSR adjacency)
    int fem_type = PASC_FEM_NONE;
    FemGeom fg{};
    double *d_x1 = nullptr;    // fine data rows [xdim][ld1]
    double *d_g1 = nullptr;    // fine gamma rows [K][ld1]
    double *d_tmp = nullptr;   // coarse rows [K][ld2]
    InstDesc fine{};           // descriptor of the fine rows (theta step, labels, abserr)
};

struct pasc_batch {
    pasc_ctx *ctx = nullptr;
    int n_series = 0, n_eps = 0;
    uint8_t *d_heldout = nullptr;  // shared data-term mask (single-problem batches)
    std::vector<int64_t> T;
    std::vector<int32_t> K;
    Engine e;
};

static int use_device(pasc_ctx *ctx) {
    HIP_TRY(hipSetDevice(ctx->c.device));
    return PASC_OK;
}

extern "C" {

const char *pasc_last_error(void) { return g_err.c_str(); }

int pasc_version(void) { return 100; }

int pasc_ctx_create(int device, pasc_ctx **out) {
    clear_error();
    if (!out) return fail(PASC_ERR_ARG, "ctx_create: out is NULL");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(PASC_ERR_ARG, "ctx_create: device %d of %d", device, n);
    HIP_TRY(hipSetDevice(device));
    pasc_ctx *c = new (std::nothrow) pasc_ctx();
    if (!c) return fail(PASC_ERR_ALLOC, "ctx_create: out of host memory");
    c->c.device = device;
    hipStream_t s;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(PASC_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    c->c.stream = s;
    c->c.own_stream = true;
    int ncu = 0;
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
    c->c.n_cu = ncu;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->c.arch = prop.gcnArchName;
    // the device address of the static scratch of pasc_theta_update_noalloc on this device
    const int rc = noalloc_init(device);
    if (rc) {
        (void)hipStreamDestroy(s);
        delete c;
        return rc;
    }
    *out = c;
    return PASC_OK;
}

int pasc_ctx_destroy(pasc_ctx *ctx) {
    if (!ctx) return PASC_OK;
    (void)hipSetDevice(ctx->c.device);
    if (ctx->c.own_stream && ctx->c.stream) (void)hipStreamDestroy((hipStream_t)ctx->c.stream);
    delete ctx;
    return PASC_OK;
}

int pasc_ctx_set_stream(pasc_ctx *ctx, void *stream) {
    if (!ctx) return fail(PASC_ERR_ARG, "ctx is NULL");
    if (ctx->c.own_stream && ctx->c.stream) (void)hipStreamDestroy((hipStream_t)ctx->c.stream);
    if (stream) {
        ctx->c.stream = stream;
        ctx->c.own_stream = false;
    } else {
        hipStream_t s;
        HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        ctx->c.stream = s;
        ctx->c.own_stream = true;
    }
    return PASC_OK;
}

int pasc_ctx_synchronize(pasc_ctx *ctx) {
    if (!ctx) return fail(PASC_ERR_ARG, "ctx is NULL");
    HIP_TRY(hipStreamSynchronize((hipStream_t)ctx->c.stream));
    return PASC_OK;
}

void pasc_spg_default_params(pasc_spg_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->maxit = 1000;
    p->eps = 1e-9;
    p->m = 20;
    p->gamma = 0.9;
    p->sigma1 = 0.0;
    p->sigma2 = 1.0;
    p->alphainit = 2.0;
    p->stop_mask = PASC_STOP_DIFFF;
    p->kkt_tol = 1e-6;
    p->check_every = 1;
    p->use_graph = 0;
    p->fp32_switch = 0.0;
    p->stall_window = 200;
    p->polish_maxit = 0;
    p->kkt_every = 1;
}

void pasc_ts_default_params(pasc_ts_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->maxit = 300;
    p->eps = 1e-6;
    p->annealing = 1;
    p->restart_seed = 13;
}

static int check_params(const pasc_spg_params *p) {
    if (!p) return fail(PASC_ERR_ARG, "params is NULL");
    if (p->maxit < 0) return fail(PASC_ERR_ARG, "maxit < 0");
    if (p->polish_maxit < -1) return fail(PASC_ERR_ARG, "polish_maxit < -1");
    if (p->m < 1 || p->m > PASC_MAX_M) return fail(PASC_ERR_ARG, "m must be in [1, %d]", PASC_MAX_M);
    if (!(p->sigma2 >= p->sigma1)) return fail(PASC_ERR_ARG, "sigma2 < sigma1");
    if (!(p->alphainit > 0)) return fail(PASC_ERR_ARG, "alphainit must be > 0");
    if (p->kkt_every < 1) return fail(PASC_ERR_ARG, "kkt_every must be >= 1");
    return PASC_OK;
}

// A = alpha*A0: graphh1fem.cpp:146-151 (with FEM reduction the coarse problem keeps the
// fine T in the scalef denominator, and without scalef scales by T2/T)
static double problem_alpha(const pasc_problem *p) {
    const double Tn = (double)p->T;
    if (p->fem_type != PASC_FEM_NONE)
        return p->scalef ? p->epssqr / Tn : p->epssqr * (double)p->fg.T2 / Tn;
    return p->scalef ? p->epssqr / Tn : p->epssqr;
}

static void fill_desc(pasc_problem *p, InstDesc &d) {
    const bool fem = p->fem_type != PASC_FEM_NONE;
    const int64_t n = fem ? p->fg.T2 : p->T * p->R;  // nodes of the QP's graph
    d.T = n;
    d.ld = (n + 63) / 64 * 64;
    d.goff = 0;
    d.xoff = 0;
    d.K = p->K;
    d.xdim = p->xdim;
    const double Tn = (double)p->T;  // the scalef denominator is the number of time steps
    d.alpha = problem_alpha(p);
    d.bscale = p->scalef ? 1.0 / Tn : 1.0;
    // KKT step: the time steps of the grid the QP lives on (the coarse one with FEM)
    d.s_kkt = p->scalef ? (fem ? (double)p->fg.T2 : Tn) : 1.0;
    d.heldout = fem ? nullptr : p->d_heldout;
}

// Sizes, precision and graph of a problem description (shared by problems and batches).
static int validate_desc(const pasc_problem_desc *desc) {
    if (desc->T < 1 || desc->R < 1) return fail(PASC_ERR_ARG, "T and R must be >= 1");
    if (desc->K < 1 || desc->K > PASC_MAX_K) return fail(PASC_ERR_ARG, "K must be in [1, %d]", PASC_MAX_K);
    if (desc->xdim < 1 || desc->xdim > 3) return fail(PASC_ERR_ARG, "xdim must be in [1, 3]");
    if (desc->precision != PASC_FP64 && desc->precision != PASC_FP32_BULK)
        return fail(PASC_ERR_ARG, "unknown precision %d", desc->precision);
    const bool grid = desc->grid_width > 0;
    const bool csr = desc->nbr_ptr != nullptr;
    if (grid && csr) return fail(PASC_ERR_ARG, "give either grid_width or a CSR adjacency, not both");
    if (grid && desc->R % desc->grid_width != 0)
        return fail(PASC_ERR_ARG, "grid graph needs R a multiple of grid_width");
    if (csr) {
        // the reference's graphs are undirected: symmetric adjacency, no self loops
        if (!desc->nbr_idx && desc->nbr_ptr[desc->R] > 0) return fail(PASC_ERR_ARG, "nbr_idx is NULL");
        const int64_t *pt = desc->nbr_ptr;
        const int32_t *ix = desc->nbr_idx;
        if (pt[0] != 0) return fail(PASC_ERR_ARG, "nbr_ptr[0] must be 0");
        std::vector<std::pair<int32_t, int32_t>> edges;
        for (int64_t r = 0; r < desc->R; r++) {
            if (pt[r + 1] < pt[r]) return fail(PASC_ERR_ARG, "nbr_ptr is not monotone");
            for (int64_t e = pt[r]; e < pt[r + 1]; e++) {
                if (ix[e] < 0 || ix[e] >= desc->R || ix[e] == r)
                    return fail(PASC_ERR_ARG, "nbr_idx[%lld] out of range or a self loop", (long long)e);
                edges.push_back({(int32_t)r, ix[e]});
            }
        }
        std::vector<std::pair<int32_t, int32_t>> rev(edges.size());
        for (size_t q = 0; q < edges.size(); q++) rev[q] = {edges[q].second, edges[q].first};
        std::sort(edges.begin(), edges.end());
        std::sort(rev.begin(), rev.end());
        if (edges != rev) return fail(PASC_ERR_ARG, "the adjacency is not symmetric");
        if (std::adjacent_find(edges.begin(), edges.end()) != edges.end())
            return fail(PASC_ERR_ARG, "the adjacency lists a neighbour twice");
    }
    return PASC_OK;
}

int pasc_problem_create(pasc_ctx *ctx, const pasc_problem_desc *desc, pasc_problem **out) {
    clear_error();
    if (!ctx || !desc || !out) return fail(PASC_ERR_ARG, "problem_create: NULL argument");
    int vrc = validate_desc(desc);
    if (vrc) return vrc;
    const bool grid = desc->grid_width > 0;
    const bool csr = desc->nbr_ptr != nullptr;
    const bool fem = desc->fem_reduce > 0.0 && desc->fem_reduce < 1.0 &&
                     desc->fem_type != PASC_FEM_NONE;
    int64_t T2 = 0;
    if (fem) {
        if (grid || desc->R != 1 || !desc->x)
            return fail(PASC_ERR_ARG, "FEM reduction: 1-D signals with data only");
        T2 = (int64_t)std::ceil((double)desc->T * desc->fem_reduce);  // fem.cpp:198
        int frc = fem_check_geom(desc->fem_type, desc->T, T2);
        if (frc) return frc;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    pasc_problem *p = new (std::nothrow) pasc_problem();
    if (!p) return fail(PASC_ERR_ALLOC, "problem_create: out of host memory");
    p->ctx = ctx;
    p->T = desc->T;
    p->R = desc->R;
    p->K = desc->K;
    p->xdim = desc->xdim;
    p->scalef = desc->scalef;
    p->precision = desc->precision;
    p->grid_width = desc->grid_width;
    p->graph = grid || csr || desc->R > 1;
    if (fem) {
        p->fem_type = desc->fem_type;
        const int64_t ld1 = (desc->T + 63) / 64 * 64, ld2 = (T2 + 63) / 64 * 64;
        p->fg = fem_geom(desc->fem_type, desc->T, T2, ld1, ld2);
    }
    const int64_t n = p->T * p->R;  // nodes
    p->has_data = desc->x != nullptr;
    if (p->has_data) p->h_x.assign(desc->x, desc->x + n * p->xdim);
    if (desc->heldout) {
        p->h_heldout.assign(desc->heldout, desc->heldout + n);
        hipError_t e = hipMalloc(&p->d_heldout, (size_t)n);
        if (e == hipSuccess) e = hipMemcpy(p->d_heldout, desc->heldout, (size_t)n, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (p->d_heldout) (void)hipFree(p->d_heldout);
            delete p;
            return fail(PASC_ERR_HIP, "heldout upload: %s", hipGetErrorString(e));
        }
    }
    InstDesc d;
    std::memset(&d, 0, sizeof(d));
    fill_desc(p, d);
    std::vector<InstDesc> descs{d};
    rc = engine_alloc(p->e, &ctx->c, p->precision, p->graph ? STENCIL_GRID : STENCIL_TIME,
                      p->grid_width, descs, d.ld * p->xdim);
    if (rc == PASC_OK && p->graph)
        rc = engine_set_graph(p->e, p->T, p->R, p->grid_width, desc->nbr_ptr, desc->nbr_idx);
    if (rc == PASC_OK && !fem) {
        // data rows: SoA, row c holds channel c
        std::vector<double> rows((size_t)d.ld * p->xdim, 0.0);
        for (int c = 0; c < p->xdim; c++)
            for (int64_t i = 0; p->has_data && i < n; i++)
                rows[(size_t)c * d.ld + i] = desc->x[i * p->xdim + c];
        rc = engine_upload_data(p->e, rows.data(), 0, (int64_t)rows.size());
    }
    if (rc == PASC_OK && fem) {
        // the fine grid: data rows, gamma rows and a coarse scratch; the engine itself runs
        // on the explicit linear term reduced from the fine residual before every solve
        const int64_t ld1 = p->fg.ld1;
        std::vector<double> rows((size_t)ld1 * p->xdim, 0.0);
        for (int c = 0; c < p->xdim; c++)
            for (int64_t i = 0; i < n; i++) rows[(size_t)c * ld1 + i] = desc->x[i * p->xdim + c];
        hipError_t he = hipMalloc(&p->d_x1, sizeof(double) * rows.size());
        if (he == hipSuccess)
            he = hipMemcpy(p->d_x1, rows.data(), sizeof(double) * rows.size(), hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMalloc(&p->d_g1, sizeof(double) * (size_t)p->K * ld1);
        if (he == hipSuccess) he = hipMemset(p->d_g1, 0, sizeof(double) * (size_t)p->K * ld1);
        if (he == hipSuccess) he = hipMalloc(&p->d_tmp, sizeof(double) * (size_t)p->K * p->fg.ld2);
        if (he != hipSuccess) rc = fail(PASC_ERR_HIP, "FEM buffers: %s", hipGetErrorString(he));
        std::memset(&p->fine, 0, sizeof(p->fine));
        p->fine.T = p->T;
        p->fine.ld = ld1;
        p->fine.K = p->K;
        p->fine.xdim = p->xdim;
        p->fine.heldout = p->d_heldout;
    }
    if (rc) {
        engine_free(p->e);
        if (p->d_heldout) (void)hipFree(p->d_heldout);
        if (p->d_x1) (void)hipFree(p->d_x1);
        if (p->d_g1) (void)hipFree(p->d_g1);
        if (p->d_tmp) (void)hipFree(p->d_tmp);
        delete p;
        return rc;
    }
    *out = p;
    return PASC_OK;
}

int pasc_problem_destroy(pasc_problem *p) {
    if (!p) return PASC_OK;
    (void)hipSetDevice(p->ctx->c.device);
    engine_free(p->e);
    if (p->d_heldout) (void)hipFree(p->d_heldout);
    if (p->d_x1) (void)hipFree(p->d_x1);
    if (p->d_g1) (void)hipFree(p->d_g1);
    if (p->d_tmp) (void)hipFree(p->d_tmp);
    delete p;
    return PASC_OK;
}

int pasc_set_theta(pasc_problem *p, const double *theta) {
    clear_error();
    if (!p || !theta) return fail(PASC_ERR_ARG, "set_theta: NULL argument");
    for (int i = 0; i < p->K * p->xdim; i++) {
        if (!std::isfinite(theta[i])) return fail(PASC_ERR_ARG, "set_theta: theta[%d] is not finite", i);
        p->e.h_desc[0].theta[i] = theta[i];
        p->fine.theta[i] = theta[i];
    }
    p->theta_set = true;
    int rc = use_device(p->ctx);
    if (rc) return rc;
    return engine_update_desc(p->e);
}

int pasc_get_theta(pasc_problem *p, double *theta) {
    if (!p || !theta) return fail(PASC_ERR_ARG, "get_theta: NULL argument");
    for (int i = 0; i < p->K * p->xdim; i++) theta[i] = p->e.h_desc[0].theta[i];
    return PASC_OK;
}

int pasc_set_gamma(pasc_problem *p, const double *gamma, int layout) {
    clear_error();
    if (!p || !gamma) return fail(PASC_ERR_ARG, "set_gamma: NULL argument");
    if (layout != PASC_LAYOUT_TMAJOR && layout != PASC_LAYOUT_KMAJOR) return fail(PASC_ERR_ARG, "bad layout");
    int rc = use_device(p->ctx);
    if (rc) return rc;
    if (p->fem_type != PASC_FEM_NONE) {
        // the fine gamma (TSData::gammavector); it is reduced to the coarse grid per solve
        const int64_t T = p->T, ld1 = p->fg.ld1;
        std::vector<double> rows((size_t)p->K * ld1, 0.0);
        for (int k = 0; k < p->K; k++)
            for (int64_t t = 0; t < T; t++)
                rows[k * ld1 + t] = layout == PASC_LAYOUT_TMAJOR ? gamma[t * p->K + k] : gamma[k * T + t];
        HIP_TRY(hipMemcpy(p->d_g1, rows.data(), sizeof(double) * rows.size(), hipMemcpyHostToDevice));
        p->gamma_set = true;
        return PASC_OK;
    }
    rc = engine_set_gamma(p->e, 0, gamma, layout);
    if (rc == PASC_OK) p->gamma_set = true;
    return rc;
}

int pasc_set_gamma_uniform(pasc_problem *p) {
    clear_error();
    if (!p) return fail(PASC_ERR_ARG, "set_gamma_uniform: NULL argument");
    int rc = use_device(p->ctx);
    if (rc) return rc;
    if (p->fem_type != PASC_FEM_NONE) {
        std::vector<double> g((size_t)p->T * p->K, 1.0 / p->K);
        return pasc_set_gamma(p, g.data(), PASC_LAYOUT_TMAJOR);
    }
    rc = engine_set_uniform(p->e, 0);
    if (rc == PASC_OK) p->gamma_set = true;
    return rc;
}

int pasc_get_gamma(pasc_problem *p, double *gamma, int layout) {
    clear_error();
    if (!p || !gamma) return fail(PASC_ERR_ARG, "get_gamma: NULL argument");
    if (layout != PASC_LAYOUT_TMAJOR && layout != PASC_LAYOUT_KMAJOR) return fail(PASC_ERR_ARG, "bad layout");
    if (!p->gamma_set) return fail(PASC_ERR_STATE, "get_gamma: gamma was never set");
    int rc = use_device(p->ctx);
    if (rc) return rc;
    if (p->fem_type != PASC_FEM_NONE) {
        const int64_t T = p->T, ld1 = p->fg.ld1;
        std::vector<double> rows((size_t)p->K * ld1);
        HIP_TRY(hipMemcpyAsync(rows.data(), p->d_g1, sizeof(double) * rows.size(),
                               hipMemcpyDeviceToHost, (hipStream_t)p->ctx->c.stream));
        HIP_TRY(hipStreamSynchronize((hipStream_t)p->ctx->c.stream));
        for (int k = 0; k < p->K; k++)
            for (int64_t t = 0; t < T; t++) {
                if (layout == PASC_LAYOUT_TMAJOR) gamma[t * p->K + k] = rows[k * ld1 + t];
                else gamma[k * T + t] = rows[k * ld1 + t];
            }
        return PASC_OK;
    }
    return engine_get_gamma(p->e, 0, gamma, layout);
}

int pasc_set_epssqr(pasc_problem *p, double epssqr) {
    clear_error();
    if (!p) return fail(PASC_ERR_ARG, "set_epssqr: NULL problem");
    if (!(epssqr >= 0) || !std::isfinite(epssqr)) return fail(PASC_ERR_ARG, "epssqr must be finite and >= 0");
    p->epssqr = epssqr;
    InstDesc &d = p->e.h_desc[0];
    d.alpha = problem_alpha(p);
    int rc = use_device(p->ctx);
    if (rc) return rc;
    return engine_update_desc(p->e);
}

// FEM: the coarse linear term b2 = -reduce((x - theta_k)^2) / T2 (scalef) and the coarse start
// reduce(gamma) (graphh1fem.cpp:261-286), before the coarse solve
static int fem_before_solve(pasc_problem *p) {
    void *s = p->ctx->c.stream;
    double *b2;
    int rc = engine_linear_term_rows(p->e, 0, &b2);
    if (rc) return rc;
    const double bs = p->scalef ? -1.0 / (double)p->fg.T2 : -1.0;
    rc = fem_residual_rows(s, p->fg, p->K, p->d_x1, p->xdim, p->fine.theta, p->d_heldout, b2, bs);
    if (rc) return rc;
    rc = engine_commit_linear_term(p->e, 0);
    if (rc) return rc;
    rc = fem_reduce_rows(s, p->fg, p->K, p->d_g1, p->d_tmp, 1.0);
    if (rc) return rc;
    return engine_set_gamma_rows(p->e, 0, p->d_tmp);
}

// the gamma step of GraphH1FEMModel: with FEM, reduce -> coarse solve -> prolongate
// (graphh1fem.cpp:236-301); otherwise the solve itself
static int problem_gamma_step(pasc_problem *p, const pasc_spg_params &sp, pasc_spg_result *r) {
    if (p->fem_type == PASC_FEM_NONE) return engine_solve(p->e, sp, r, nullptr, nullptr);
    int rc = fem_before_solve(p);
    if (rc) return rc;
    rc = engine_solve(p->e, sp, r, nullptr, nullptr);
    if (rc) return rc;
    bool f32 = false;
    const void *g2 = engine_current_gamma(p->e, 0, &f32);
    return fem_prolongate_rows(p->ctx->c.stream, p->fg, p->K, g2, f32, p->d_g1);
}

// the theta step: from the fine gamma with FEM
static int problem_theta_step(pasc_problem *p) {
    if (p->fem_type == PASC_FEM_NONE) return engine_theta_update(p->e, 0, nullptr);
    double th[PASC_MAX_K * 3];
    int rc = rows_theta(p->e, p->d_g1, false, p->fine, p->d_x1, th);
    if (rc) return rc;
    for (int i = 0; i < p->K * p->xdim; i++) {
        p->e.h_desc[0].theta[i] = th[i];
        p->fine.theta[i] = th[i];
    }
    return engine_update_desc(p->e);
}

int pasc_gamma_solve(pasc_problem *p, const pasc_spg_params *params, pasc_spg_result *res) {
    clear_error();
    if (!p) return fail(PASC_ERR_ARG, "gamma_solve: NULL problem");
    int rc = check_params(params);
    if (rc) return rc;
    if (!p->theta_set && !p->b_set)
        return fail(PASC_ERR_STATE, "gamma_solve: neither theta nor an explicit b was set");
    if (!p->has_data && !p->b_set)
        return fail(PASC_ERR_STATE, "gamma_solve: the problem has no data; set an explicit b");
    rc = use_device(p->ctx);
    if (rc) return rc;
    if (!p->gamma_set) {
        rc = pasc_set_gamma_uniform(p);
        if (rc) return rc;
    }
    pasc_spg_result r;
    std::memset(&r, 0, sizeof(r));
    rc = problem_gamma_step(p, *params, &r);
    if (res) *res = r;
    if (rc == PASC_OK) p->solved = true;
    return rc;
}

int pasc_set_linear_term(pasc_problem *p, const double *b, int layout) {
    clear_error();
    if (!p) return fail(PASC_ERR_ARG, "set_linear_term: NULL problem");
    if (p->fem_type != PASC_FEM_NONE)
        return fail(PASC_ERR_STATE, "set_linear_term: the FEM-reduced problem forms its own b");
    if (layout != PASC_LAYOUT_TMAJOR && layout != PASC_LAYOUT_KMAJOR) return fail(PASC_ERR_ARG, "bad layout");
    if (b) {
        const int64_t n = p->T * p->R * (int64_t)p->K;
        for (int64_t i = 0; i < n; i++)
            if (!std::isfinite(b[i])) return fail(PASC_ERR_ARG, "set_linear_term: b[%lld] is not finite", (long long)i);
    }
    int rc = use_device(p->ctx);
    if (rc) return rc;
    rc = engine_set_linear_term(p->e, 0, b, layout);
    if (rc == PASC_OK) {
        p->b_set = b != nullptr;
        p->solved = false;
    }
    return rc;
}

int pasc_set_data(pasc_problem *p, const double *x) {
    clear_error();
    if (!p || !x) return fail(PASC_ERR_ARG, "set_data: NULL argument");
    if (!p->has_data) return fail(PASC_ERR_STATE, "set_data: the problem was created without data");
    const int64_t n = p->T * p->R;
    for (int64_t i = 0; i < n * p->xdim; i++)
        if (!std::isfinite(x[i])) return fail(PASC_ERR_ARG, "set_data: x[%lld] is not finite", (long long)i);
    int rc = use_device(p->ctx);
    if (rc) return rc;
    // the data rows as at creation: SoA, row c holds channel c (the fine grid when FEM-reduced)
    const int64_t ld = p->fem_type != PASC_FEM_NONE ? p->fg.ld1 : p->e.h_desc[0].ld;
    std::vector<double> rows((size_t)ld * p->xdim, 0.0);
    for (int c = 0; c < p->xdim; c++)
        for (int64_t i = 0; i < n; i++) rows[(size_t)c * ld + i] = x[i * p->xdim + c];
    if (p->fem_type != PASC_FEM_NONE) {
        const hipError_t he = hipMemcpy(p->d_x1, rows.data(), sizeof(double) * rows.size(),
                                        hipMemcpyHostToDevice);
        if (he != hipSuccess) return fail(PASC_ERR_HIP, "set_data: %s", hipGetErrorString(he));
    } else {
        rc = engine_upload_data(p->e, rows.data(), 0, (int64_t)rows.size());
        if (rc) return rc;
    }
    p->h_x.assign(x, x + n * p->xdim);
    p->solved = false;
    return PASC_OK;
}

int pasc_theta_update(pasc_problem *p) {
    clear_error();
    if (!p) return fail(PASC_ERR_ARG, "theta_update: NULL problem");
    if (!p->gamma_set) return fail(PASC_ERR_STATE, "theta_update: gamma was never set");
    if (!p->has_data) return fail(PASC_ERR_STATE, "theta_update: the problem has no data");
    int rc = use_device(p->ctx);
    if (rc) return rc;
    rc = problem_theta_step(p);
    if (rc == PASC_OK) p->theta_set = true;
    return rc;
}

int pasc_theta_update_noalloc(pasc_problem *p) {
    clear_error();
    if (!p) return fail(PASC_ERR_ARG, "theta_update_noalloc: NULL problem");
    if (!p->gamma_set) return fail(PASC_ERR_STATE, "theta_update_noalloc: gamma was never set");
    if (!p->has_data) return fail(PASC_ERR_STATE, "theta_update_noalloc: the problem has no data");
    int rc = use_device(p->ctx);
    if (rc) return rc;
    double th[PASC_MAX_K * 3];
    InstDesc &d = p->e.h_desc[0];
    if (p->fem_type == PASC_FEM_NONE) {
        bool f32 = false;
        const void *g = engine_current_gamma(p->e, 0, &f32);
        rc = rows_theta_noalloc(p->e, g, f32, d, (const double *)p->e.d_x64, th);
    } else {
        rc = rows_theta_noalloc(p->e, p->d_g1, false, p->fine, p->d_x1, th);
    }
    if (rc) return rc;
    for (int i = 0; i < p->K * p->xdim; i++) {
        d.theta[i] = th[i];
        p->fine.theta[i] = th[i];
    }
    rc = engine_update_desc(p->e);
    if (rc == PASC_OK) p->theta_set = true;
    return rc;
}

// One run of the alternation (tssolver.h:542-650): theta step, gamma step, L = fx of the gamma
// solve (GraphH1FEMModel::get_L), stop at the first |L - L_old| < eps.  The reference breaks
// before its it++, so it_last is the index of the breaking iteration, or maxit.
struct TsRun {
    int it = 0, spg = 0, increased = 0;
    double L = 0.0, deltaL = 0.0;
};

static int ts_run(pasc_problem *p, const pasc_ts_params *tp, const pasc_spg_params *sp,
                  bool skip_first, TsRun &o) {
    double L = std::numeric_limits<double>::max(), L_old, deltaL = L;
    int it, spg_sum = 0, increased = 0;
    for (it = 0; it < tp->maxit; it++) {
        if (!tp->theta_fixed && !(it == 0 && skip_first)) {
            int rc = problem_theta_step(p);
            if (rc) return rc;
            p->theta_set = true;
        }
        if (!p->theta_set) return fail(PASC_ERR_STATE, "ts_solve: theta was never set");
        pasc_spg_result r;
        int rc = problem_gamma_step(p, *sp, &r);
        if (rc) return rc;
        spg_sum += r.it;
        L_old = L;
        L = r.fx;
        deltaL = std::fabs(L - L_old);
        if (L - L_old > 0) increased++;
        if (deltaL < tp->eps) break;
    }
    o.it = it;
    o.spg = spg_sum;
    o.L = L;
    o.deltaL = deltaL;
    o.increased = increased;
    return PASC_OK;
}

int pasc_ts_solve(pasc_problem *p, const pasc_ts_params *tp, const pasc_spg_params *sp,
                  pasc_ts_result *res) {
    clear_error();
    if (!p || !tp) return fail(PASC_ERR_ARG, "ts_solve: NULL argument");
    if (p->b_set) return fail(PASC_ERR_STATE, "ts_solve: the problem has an explicit b (theta would not enter it)");
    if (!p->has_data) return fail(PASC_ERR_STATE, "ts_solve: the problem has no data");
    if (tp->annealing < 0) return fail(PASC_ERR_ARG, "ts_solve: annealing must be >= 0");
    int rc = check_params(sp);
    if (rc) return rc;
    rc = use_device(p->ctx);
    if (rc) return rc;
    if (!p->gamma_set) {
        rc = pasc_set_gamma_uniform(p);
        if (rc) return rc;
    }
    // tssolver.h:491-715: `annealing` runs; each ends with AIC = 2 log L + K
    // (graphh1fem.h:443-447), and with more than one run the gamma and theta of the smallest AIC
    // are kept (tssolver.h:669-692).  Every run after the first starts from the random gamma of
    // set_random (tssolver.h:687-689), which re-creates its generator with the same seed at
    // every call (generalvector.cpp:7-29): every restart draws the same vector, here from
    // pasc_generate_uniform(restart_seed).  The best run is the best of this solve (the
    // reference compares against the AIC kept in TSData from earlier solves too).
    const int runs = tp->annealing > 1 ? tp->annealing : 1;
    const int64_t n = p->T * p->R * (int64_t)p->K;
    const int nth = p->K * p->xdim;
    std::vector<double> gbest, tbest, grand;
    if (runs > 1) {
        gbest.resize((size_t)n);
        tbest.resize((size_t)nth);
        grand.resize((size_t)n);
        rc = pasc_generate_uniform(tp->restart_seed, n, grand.data());
        if (rc) return rc;
    }
    double best_aic = std::numeric_limits<double>::max(), L_keep = 0.0, dL_keep = 0.0;
    int it_last = 0, it_sum = 0, spg_sum = 0, increased = 0, best_run = 0;
    for (int a = 0; a < runs; a++) {
        TsRun o;
        rc = ts_run(p, tp, sp, a == 0 && tp->skip_first_theta, o);
        if (rc) return rc;
        it_last = o.it;
        it_sum += o.it;
        spg_sum += o.spg;
        increased += o.increased;
        const double aic = 2.0 * std::log(o.L) + p->K;
        if (runs <= 1 || aic < best_aic) {
            if (runs > 1) {
                rc = pasc_get_gamma(p, gbest.data(), PASC_LAYOUT_TMAJOR);
                if (rc == PASC_OK) rc = pasc_get_theta(p, tbest.data());
                if (rc) return rc;
            }
            best_aic = aic;
            L_keep = o.L;
            dL_keep = o.deltaL;
            best_run = a;
        }
        if (a < runs - 1) {
            rc = pasc_set_gamma(p, grand.data(), PASC_LAYOUT_TMAJOR);
            if (rc) return rc;
        }
    }
    if (runs > 1) {
        rc = pasc_set_gamma(p, gbest.data(), PASC_LAYOUT_TMAJOR);
        if (rc == PASC_OK) rc = pasc_set_theta(p, tbest.data());
        if (rc) return rc;
    }
    p->solved = true;
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
    return PASC_OK;
}

int pasc_objective(pasc_problem *p, double *fx, double *kkt) {
    clear_error();
    if (!p) return fail(PASC_ERR_ARG, "objective: NULL problem");
    if (!p->gamma_set) return fail(PASC_ERR_STATE, "objective: gamma was never set");
    int rc = use_device(p->ctx);
    if (rc) return rc;
    return engine_eval(p->e, 0, fx, kkt);
}

int pasc_abserr_reconstructed(pasc_problem *p, const double *solution, double *abserr) {
    clear_error();
    if (!p || !solution || !abserr) return fail(PASC_ERR_ARG, "abserr: NULL argument");
    if (p->xdim != 1) return fail(PASC_ERR_ARG, "abserr needs xdim == 1");
    if (!p->gamma_set) return fail(PASC_ERR_STATE, "abserr: gamma was never set");
    int rc = use_device(p->ctx);
    if (rc) return rc;
    if (p->fem_type != PASC_FEM_NONE) return rows_abserr(p->e, p->d_g1, false, p->fine, solution, abserr);
    return engine_abserr(p->e, 0, solution, abserr);
}

int pasc_get_labels(pasc_problem *p, int32_t *labels) {
    clear_error();
    if (!p || !labels) return fail(PASC_ERR_ARG, "labels: NULL argument");
    if (!p->gamma_set) return fail(PASC_ERR_STATE, "labels: gamma was never set");
    int rc = use_device(p->ctx);
    if (rc) return rc;
    if (p->fem_type != PASC_FEM_NONE) return rows_labels(p->e, p->d_g1, false, p->fine, labels);
    return engine_labels(p->e, 0, labels);
}

int pasc_project_simplex(pasc_ctx *ctx, void *x_dev, int64_t nblocks, int32_t K, int dtype,
                         int layout, int64_t ld) {
    clear_error();
    if (!ctx || (!x_dev && nblocks > 0)) return fail(PASC_ERR_ARG, "project: NULL argument");
    if (nblocks < 0 || K < 1 || K > 65536) return fail(PASC_ERR_ARG, "project: bad sizes");
    if (dtype != PASC_DTYPE_F32 && dtype != PASC_DTYPE_F64) return fail(PASC_ERR_ARG, "bad dtype");
    if (layout != PASC_LAYOUT_TMAJOR && layout != PASC_LAYOUT_KMAJOR) return fail(PASC_ERR_ARG, "bad layout");
    if (layout == PASC_LAYOUT_KMAJOR && ld < nblocks) return fail(PASC_ERR_ARG, "project: ld < nblocks");
    int rc = use_device(ctx);
    if (rc) return rc;
    return kernel_project(&ctx->c, x_dev, nblocks, K, dtype, layout, ld);
}

int pasc_h1_matvec(pasc_ctx *ctx, const void *x_dev, void *y_dev, int64_t T, int64_t R,
                   int32_t K, int64_t ld, double alpha, int32_t grid_width, int dtype) {
    clear_error();
    if (!ctx || !x_dev || !y_dev) return fail(PASC_ERR_ARG, "matvec: NULL argument");
    if (K < 1 || T < 1 || R < 1) return fail(PASC_ERR_ARG, "matvec: bad sizes");
    if (grid_width > 0 && R % grid_width) return fail(PASC_ERR_ARG, "matvec: grid needs R % width == 0");
    if (ld < T * R) return fail(PASC_ERR_ARG, "matvec: ld < T*R");
    if (dtype != PASC_DTYPE_F32 && dtype != PASC_DTYPE_F64) return fail(PASC_ERR_ARG, "bad dtype");
    int rc = use_device(ctx);
    if (rc) return rc;
    return kernel_matvec(&ctx->c, x_dev, y_dev, T, R, K, ld, alpha, grid_width, nullptr, nullptr,
                         dtype);
}

int pasc_h1_matvec_csr(pasc_ctx *ctx, const void *x_dev, void *y_dev, int64_t T, int64_t R,
                       int32_t K, int64_t ld, double alpha, const int64_t *nbr_ptr,
                       const int32_t *nbr_idx, int dtype) {
    clear_error();
    if (!ctx || !x_dev || !y_dev || !nbr_ptr) return fail(PASC_ERR_ARG, "matvec_csr: NULL argument");
    if (K < 1 || T < 1 || R < 1) return fail(PASC_ERR_ARG, "matvec_csr: bad sizes");
    if (ld < T * R) return fail(PASC_ERR_ARG, "matvec_csr: ld < T*R");
    if (dtype != PASC_DTYPE_F32 && dtype != PASC_DTYPE_F64) return fail(PASC_ERR_ARG, "bad dtype");
    const int64_t nnz = nbr_ptr[R];
    if (nbr_ptr[0] != 0 || nnz < 0 || (nnz > 0 && !nbr_idx)) return fail(PASC_ERR_ARG, "matvec_csr: bad CSR");
    for (int64_t r = 0; r < R; r++) {
        if (nbr_ptr[r + 1] < nbr_ptr[r]) return fail(PASC_ERR_ARG, "matvec_csr: nbr_ptr is not monotone");
        for (int64_t e = nbr_ptr[r]; e < nbr_ptr[r + 1]; e++)
            if (nbr_idx[e] < 0 || nbr_idx[e] >= R) return fail(PASC_ERR_ARG, "matvec_csr: index out of range");
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    int64_t *dp = nullptr;
    int32_t *di = nullptr;
    HIP_TRY(hipMalloc(&dp, sizeof(int64_t) * (size_t)(R + 1)));
    hipError_t he = hipMalloc(&di, sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1));
    if (he == hipSuccess)
        he = hipMemcpy(dp, nbr_ptr, sizeof(int64_t) * (size_t)(R + 1), hipMemcpyHostToDevice);
    if (he == hipSuccess && nnz > 0)
        he = hipMemcpy(di, nbr_idx, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice);
    if (he == hipSuccess)
        rc = kernel_matvec(&ctx->c, x_dev, y_dev, T, R, K, ld, alpha, 0, dp, di, dtype);
    else
        rc = fail(PASC_ERR_HIP, "matvec_csr: %s", hipGetErrorString(he));
    (void)hipFree(dp);
    if (di) (void)hipFree(di);
    return rc;
}

// ---- batch ----
int pasc_batch_create(pasc_ctx *ctx, int32_t n_series, const pasc_series_desc *series,
                      int32_t n_eps, const double *eps, int32_t precision, pasc_batch **out) {
    clear_error();
    if (!ctx || !series || !eps || !out || n_series < 1 || n_eps < 1)
        return fail(PASC_ERR_ARG, "batch_create: bad arguments");
    if (precision != PASC_FP64 && precision != PASC_FP32_BULK) return fail(PASC_ERR_ARG, "bad precision");
    std::vector<InstDesc> descs;
    std::vector<int64_t> xoff(n_series);
    int64_t xo = 0, go = 0;
    for (int s = 0; s < n_series; s++) {
        const pasc_series_desc &sd = series[s];
        if (sd.T < 1 || sd.K < 1 || sd.K > PASC_MAX_K || !sd.x || !sd.theta)
            return fail(PASC_ERR_ARG, "batch_create: bad series %d", s);
        xoff[s] = xo;
        xo += (sd.T + 63) / 64 * 64;
    }
    for (int s = 0; s < n_series; s++) {
        const pasc_series_desc &sd = series[s];
        for (int j = 0; j < n_eps; j++) {
            if (!(eps[j] >= 0) || !std::isfinite(eps[j])) return fail(PASC_ERR_ARG, "bad eps[%d]", j);
            InstDesc d;
            std::memset(&d, 0, sizeof(d));
            d.T = sd.T;
            d.ld = (sd.T + 63) / 64 * 64;
            d.goff = go;
            d.xoff = xoff[s];
            d.K = sd.K;
            d.xdim = 1;
            d.alpha = eps[j] / (double)sd.T;
            d.bscale = 1.0 / (double)sd.T;
            d.s_kkt = (double)sd.T;
            d.heldout = nullptr;
            for (int k = 0; k < sd.K; k++) d.theta[k] = sd.theta[k];
            go += (int64_t)d.K * d.ld;
            descs.push_back(d);
        }
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    pasc_batch *b = new (std::nothrow) pasc_batch();
    if (!b) return fail(PASC_ERR_ALLOC, "batch_create: out of host memory");
    b->ctx = ctx;
    b->n_series = n_series;
    b->n_eps = n_eps;
    rc = engine_alloc(b->e, &ctx->c, precision, STENCIL_TIME, 0, descs, xo);
    for (int s = 0; s < n_series && rc == PASC_OK; s++)
        rc = engine_upload_data(b->e, series[s].x, xoff[s], series[s].T);
    if (rc) {
        engine_free(b->e);
        delete b;
        return rc;
    }
    *out = b;
    return PASC_OK;
}

int pasc_batch_create_problem(pasc_ctx *ctx, const pasc_problem_desc *desc,
                              const double *theta, int32_t n_eps, const double *eps,
                              pasc_batch **out) {
    clear_error();
    if (!ctx || !desc || !theta || !eps || !out || n_eps < 1)
        return fail(PASC_ERR_ARG, "batch_create_problem: bad arguments");
    int rc = validate_desc(desc);
    if (rc) return rc;
    if (!desc->x) return fail(PASC_ERR_ARG, "batch_create_problem: the batch needs data");
    if (desc->fem_reduce > 0.0 && desc->fem_reduce < 1.0 && desc->fem_type != PASC_FEM_NONE)
        return fail(PASC_ERR_ARG, "batch_create_problem: no FEM reduction in batches");
    const bool graph = desc->grid_width > 0 || desc->nbr_ptr || desc->R > 1;
    const int64_t n = desc->T * desc->R, ld = (n + 63) / 64 * 64;
    const double Tn = (double)desc->T;
    std::vector<InstDesc> descs;
    for (int j = 0; j < n_eps; j++) {
        if (!(eps[j] >= 0) || !std::isfinite(eps[j])) return fail(PASC_ERR_ARG, "bad eps[%d]", j);
        InstDesc d;
        std::memset(&d, 0, sizeof(d));
        d.T = n;
        d.ld = ld;
        d.goff = (int64_t)j * desc->K * ld;
        d.xoff = 0;  // every epsilon reads the same data rows
        d.K = desc->K;
        d.xdim = desc->xdim;
        d.alpha = desc->scalef ? eps[j] / Tn : eps[j];
        d.bscale = desc->scalef ? 1.0 / Tn : 1.0;
        d.s_kkt = desc->scalef ? Tn : 1.0;
        for (int i = 0; i < desc->K * desc->xdim; i++) {
            if (!std::isfinite(theta[i])) return fail(PASC_ERR_ARG, "theta[%d] is not finite", i);
            d.theta[i] = theta[i];
        }
        descs.push_back(d);
    }
    rc = use_device(ctx);
    if (rc) return rc;
    pasc_batch *b = new (std::nothrow) pasc_batch();
    if (!b) return fail(PASC_ERR_ALLOC, "batch_create_problem: out of host memory");
    b->ctx = ctx;
    b->n_series = 1;
    b->n_eps = n_eps;
    if (desc->heldout) {
        hipError_t he = hipMalloc(&b->d_heldout, (size_t)n);
        if (he == hipSuccess) he = hipMemcpy(b->d_heldout, desc->heldout, (size_t)n, hipMemcpyHostToDevice);
        if (he != hipSuccess) {
            if (b->d_heldout) (void)hipFree(b->d_heldout);
            delete b;
            return fail(PASC_ERR_HIP, "heldout upload: %s", hipGetErrorString(he));
        }
        for (auto &d : descs) d.heldout = b->d_heldout;
    }
    rc = engine_alloc(b->e, &ctx->c, desc->precision, graph ? STENCIL_GRID : STENCIL_TIME,
                      desc->grid_width, descs, ld * desc->xdim);
    if (rc == PASC_OK && graph)
        rc = engine_set_graph(b->e, desc->T, desc->R, desc->grid_width, desc->nbr_ptr, desc->nbr_idx);
    if (rc == PASC_OK) {
        std::vector<double> rows((size_t)ld * desc->xdim, 0.0);
        for (int c = 0; c < desc->xdim; c++)
            for (int64_t i = 0; i < n; i++) rows[(size_t)c * ld + i] = desc->x[i * desc->xdim + c];
        rc = engine_upload_data(b->e, rows.data(), 0, (int64_t)rows.size());
    }
    if (rc) {
        engine_free(b->e);
        if (b->d_heldout) (void)hipFree(b->d_heldout);
        delete b;
        return rc;
    }
    *out = b;
    return PASC_OK;
}

int pasc_batch_destroy(pasc_batch *b) {
    if (!b) return PASC_OK;
    (void)hipSetDevice(b->ctx->c.device);
    engine_free(b->e);
    if (b->d_heldout) (void)hipFree(b->d_heldout);
    delete b;
    return PASC_OK;
}

int pasc_batch_solve(pasc_batch *b, const pasc_spg_params *params, pasc_instance_result *res,
                     double *time_s) {
    clear_error();
    if (!b) return fail(PASC_ERR_ARG, "batch_solve: NULL batch");
    int rc = check_params(params);
    if (rc) return rc;
    rc = use_device(b->ctx);
    if (rc) return rc;
    rc = engine_fill_uniform(b->e);
    if (rc) return rc;
    return engine_solve(b->e, *params, nullptr, res, time_s);
}

int pasc_batch_get_gamma(pasc_batch *b, int32_t series, int32_t eps_index, double *gamma,
                         int layout) {
    clear_error();
    if (!b || !gamma) return fail(PASC_ERR_ARG, "batch_get_gamma: NULL argument");
    if (series < 0 || series >= b->n_series || eps_index < 0 || eps_index >= b->n_eps)
        return fail(PASC_ERR_ARG, "batch_get_gamma: index out of range");
    int rc = use_device(b->ctx);
    if (rc) return rc;
    return engine_get_gamma(b->e, series * b->n_eps + eps_index, gamma, layout);
}

int pasc_adjoint(pasc_problem *p, const uint8_t *mask, double cg_tol, int32_t cg_maxit,
                 double *loss, double *dlogeps, double *dtheta, int32_t *cg_it) {
    clear_error();
    if (!p || !mask || !loss || !dlogeps || !dtheta) return fail(PASC_ERR_ARG, "adjoint: NULL argument");
    if (!p->solved) return fail(PASC_ERR_STATE, "adjoint: solve the gamma problem first");
    if (p->b_set || !p->has_data) return fail(PASC_ERR_STATE, "adjoint: needs the data term (no explicit b)");
    if (p->fem_type != PASC_FEM_NONE) return fail(PASC_ERR_ARG, "adjoint: not for FEM-reduced problems");
    int rc = use_device(p->ctx);
    if (rc) return rc;
    int it = 0;
    rc = engine_adjoint(p->e, 0, mask, cg_tol, cg_maxit, loss, dlogeps, dtheta, &it);
    if (cg_it) *cg_it = it;
    return rc;
}

int64_t pasc_fem_reduced_T(int64_t T1, double fem_reduce) {
    if (!(fem_reduce > 0.0) || fem_reduce >= 1.0) return T1;
    return (int64_t)std::ceil((double)T1 * fem_reduce);
}

static int fem_kernel_args(pasc_ctx *ctx, const void *a, const void *b, int64_t T1, int64_t T2,
                           int32_t K, int64_t ld1, int64_t ld2, int32_t type) {
    if (!ctx || !a || !b) return fail(PASC_ERR_ARG, "fem: NULL argument");
    if (K < 1 || K > PASC_MAX_K) return fail(PASC_ERR_ARG, "fem: K must be in [1, %d]", PASC_MAX_K);
    if (ld1 < T1 || ld2 < T2) return fail(PASC_ERR_ARG, "fem: row stride shorter than the row");
    return fem_check_geom(type, T1, T2);
}

int pasc_fem_reduce(pasc_ctx *ctx, const double *fine_dev, double *coarse_dev, int64_t T1,
                    int64_t T2, int32_t K, int64_t ld1, int64_t ld2, int32_t fem_type) {
    clear_error();
    int rc = fem_kernel_args(ctx, fine_dev, coarse_dev, T1, T2, K, ld1, ld2, fem_type);
    if (rc) return rc;
    rc = use_device(ctx);
    if (rc) return rc;
    rc = fem_reduce_rows(ctx->c.stream, fem_geom(fem_type, T1, T2, ld1, ld2), K, fine_dev,
                         coarse_dev, 1.0);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize((hipStream_t)ctx->c.stream));
    return PASC_OK;
}

int pasc_fem_prolongate(pasc_ctx *ctx, const double *coarse_dev, double *fine_dev, int64_t T1,
                        int64_t T2, int32_t K, int64_t ld1, int64_t ld2, int32_t fem_type) {
    clear_error();
    int rc = fem_kernel_args(ctx, coarse_dev, fine_dev, T1, T2, K, ld1, ld2, fem_type);
    if (rc) return rc;
    rc = use_device(ctx);
    if (rc) return rc;
    rc = fem_prolongate_rows(ctx->c.stream, fem_geom(fem_type, T1, T2, ld1, ld2), K, coarse_dev,
                             false, fine_dev);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize((hipStream_t)ctx->c.stream));
    return PASC_OK;
}

int pasc_device_copy_bandwidth(pasc_ctx *ctx, int64_t bytes, int32_t reps, double *gbps) {
    clear_error();
    if (!ctx || !gbps || bytes < 16 || reps < 1) return fail(PASC_ERR_ARG, "copy_bandwidth: bad arguments");
    int rc = use_device(ctx);
    if (rc) return rc;
    return device_copy_bandwidth(&ctx->c, bytes, reps, gbps);
}

int pasc_set_profiling(pasc_problem *p, int enable) {
    if (!p) return fail(PASC_ERR_ARG, "set_profiling: NULL problem");
    p->e.profile = enable != 0;
    return PASC_OK;
}

static PassShape to_pass(const pasc_pass_shape &s) {
    return PassShape{s.block, s.chunk, s.prefetch, s.max_slots > 0 ? s.max_slots : 2048,
                     s.min_tiles > 0 ? s.min_tiles : 1};
}

static pasc_pass_shape from_pass(const PassShape &s, double us) {
    return pasc_pass_shape{s.block, s.chunk, s.prefetch, s.max_slots, s.min_tiles, us};
}

// FP32 / FP64 phase shapes of an engine from the ABI struct
static void shapes_of(const Engine &e, const pasc_launch_shape &sh, PassShape &s32, PassShape &s64) {
    if (e.precision == PASC_FP32_BULK) {
        s32 = to_pass(sh.main);
        if (sh.polish.block > 0) {
            s64 = to_pass(sh.polish);
        } else {
            s64 = s32;
            // 1-D: the same bytes per thread; images: the same rows per tile
            if (!engine_is_image(e)) s64.chunk = s32.chunk / 2;
        }
    } else {
        s32 = e.sh[0];
        s64 = to_pass(sh.main);
    }
}

static int set_shape(Engine &e, const pasc_launch_shape *shape) {
    PassShape s32, s64;
    shapes_of(e, *shape, s32, s64);
    return engine_set_shape(e, s32, s64);
}

// Autotune with the per-context and on-file caches (one JSON object per line, last match wins).
static int autotune_cached(Engine &e, Ctx &c, const std::string &key, int64_t size,
                           const char *cache_path, pasc_launch_shape *shape) {
    pasc_launch_shape sh;
    std::memset(&sh, 0, sizeof(sh));
    bool hit = false;
    for (auto &kv : c.tune)
        if (kv.first == key) {
            sh = kv.second;
            sh.main.us_per_pass = sh.polish.us_per_pass = 0.0;
            hit = true;
        }
    if (!hit && cache_path) {
        FILE *f = fopen(cache_path, "r");
        if (f) {
            char line[1024], want[200];
            snprintf(want, sizeof(want), "\"key\": \"%s\"", key.c_str());
            while (fgets(line, sizeof(line), f)) {
                if (!strstr(line, want)) continue;
                const char *ps = strstr(line, "\"shape\": [");
                int v[10];
                // ps + 9 is the '[' of the list
                if (ps && sscanf(ps + 9, "[%d, %d, %d, %d, %d, %d, %d, %d, %d, %d]", &v[0], &v[1],
                                 &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9]) == 10) {
                    sh.main = pasc_pass_shape{v[0], v[1], v[2], v[3], v[4], 0.0};
                    sh.polish = pasc_pass_shape{v[5], v[6], v[7], v[8], v[9], 0.0};
                    hit = true;
                }
            }
            fclose(f);
            if (hit) c.tune.push_back({key, sh});
        }
    }
    int rc;
    if (hit) {
        PassShape s32, s64;
        shapes_of(e, sh, s32, s64);
        rc = engine_set_shape(e, s32, s64);
    } else {
        PassShape s32, s64;
        double us[2];
        rc = engine_autotune(e, &s32, &s64, us);
        if (rc == PASC_OK) {
            if (e.precision == PASC_FP32_BULK) {
                sh.main = from_pass(s32, us[0]);
                sh.polish = from_pass(s64, us[1]);
            } else {
                sh.main = from_pass(s64, us[1]);
                sh.polish = from_pass(s64, us[1]);
            }
            c.tune.push_back({key, sh});
        }
    }
    if (rc) return rc;
    if (shape) *shape = sh;
    if (cache_path) {
        FILE *f = fopen(cache_path, "a");
        if (f) {
            fprintf(f, "{\"key\": \"%s\", \"size\": %lld, \"shape\": [%d, %d, %d, %d, %d, %d, %d, %d, %d, %d], "
                       "\"us_per_pass\": [%.3f, %.3f], \"cached\": %d}\n",
                    key.c_str(), (long long)size, sh.main.block, sh.main.chunk, sh.main.prefetch,
                    sh.main.max_slots, sh.main.min_tiles, sh.polish.block, sh.polish.chunk,
                    sh.polish.prefetch, sh.polish.max_slots, sh.polish.min_tiles,
                    sh.main.us_per_pass, sh.polish.us_per_pass, hit ? 1 : 0);
            fclose(f);
        }
    }
    return PASC_OK;
}

int pasc_set_launch_shape(pasc_problem *p, const pasc_launch_shape *shape) {
    clear_error();
    if (!p || !shape) return fail(PASC_ERR_ARG, "set_launch_shape: NULL argument");
    int rc = use_device(p->ctx);
    if (rc) return rc;
    return set_shape(p->e, shape);
}

int pasc_autotune(pasc_problem *p, const char *cache_path, pasc_launch_shape *shape) {
    clear_error();
    if (!p) return fail(PASC_ERR_ARG, "autotune: NULL problem");
    if (p->graph && !engine_is_image(p->e))
        return fail(PASC_ERR_ARG, "autotune: graphs other than images have a fixed shape");
    if (!p->theta_set && !p->b_set) return fail(PASC_ERR_STATE, "autotune: neither theta nor b was set");
    int rc = use_device(p->ctx);
    if (rc) return rc;
    if (p->fem_type != PASC_FEM_NONE) {
        if (!p->gamma_set) {
            rc = pasc_set_gamma_uniform(p);
            if (rc) return rc;
        }
        rc = fem_before_solve(p);  // the coarse QP the autotuner times
        if (rc) return rc;
    }
    const int64_t n = p->e.h_desc[0].T;
    // keyed by device, K, precision and log2(T) (images: also the width)
    char key[160];
    char kind[32] = "";
    if (p->graph) snprintf(kind, sizeof(kind), "img%d/", p->e.geo.width);
    snprintf(key, sizeof(key), "%s/%sK%d/p%d/T2^%d", p->ctx->c.arch.c_str(), kind, p->K,
             p->precision, (int)std::floor(std::log2((double)n)));
    return autotune_cached(p->e, p->ctx->c, key, n, cache_path, shape);
}

int pasc_batch_set_launch_shape(pasc_batch *b, const pasc_launch_shape *shape) {
    clear_error();
    if (!b || !shape) return fail(PASC_ERR_ARG, "batch_set_launch_shape: NULL argument");
    int rc = use_device(b->ctx);
    if (rc) return rc;
    return set_shape(b->e, shape);
}

int pasc_batch_autotune(pasc_batch *b, const char *cache_path, pasc_launch_shape *shape) {
    clear_error();
    if (!b) return fail(PASC_ERR_ARG, "batch_autotune: NULL batch");
    if (b->e.stencil != STENCIL_TIME && !engine_is_image(b->e))
        return fail(PASC_ERR_ARG, "batch_autotune: graphs other than images have a fixed shape");
    int rc = use_device(b->ctx);
    if (rc) return rc;
    int64_t elems = 0;
    for (const auto &d : b->e.h_desc) elems += d.T * d.K;
    // keyed by device, precision, instance count and log2 of the total elements
    char key[160];
    char kind[32] = "";
    if (b->e.stencil != STENCIL_TIME) snprintf(kind, sizeof(kind), "img%d/", b->e.geo.width);
    snprintf(key, sizeof(key), "%s/%sbatch%d/p%d/E2^%d", b->ctx->c.arch.c_str(), kind, b->e.n_inst,
             b->e.precision, (int)std::floor(std::log2((double)std::max<int64_t>(elems, 1))));
    rc = engine_fill_uniform(b->e);  // the scan starts from gamma0 = 1/K
    if (rc) return rc;
    return autotune_cached(b->e, b->ctx->c, key, elems, cache_path, shape);
}

}  // extern "C"

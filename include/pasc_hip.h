/*
 * pasc_hip.h -- C ABI of the MI355X (gfx950) FEM-H1 K-means gamma/theta path.
 *
 * This is the drop-in boundary for the reference's PETSc vector path.  Each entry point names
 * the reference interface it replaces (paths relative to the PASC_inference checkout).  All
 * functions return a pasc_status (0 = PASC_OK); on failure pasc_last_error() returns a
 * thread-local message.  No exceptions cross the ABI.  Host buffers belong to the caller;
 * device buffers belong to the library behind opaque handles, except for the two kernel-level
 * entry points (pasc_project_simplex, pasc_h1_matvec) which work on caller-owned device memory.
 *
 * Layouts: the reference keeps gamma t-major with K innermost (idx = t*R*K + r*K + k,
 * decomposition.h:488-492).  Internally the library stores gamma k-major ([K][T*R], each row
 * padded to a multiple of 64 elements) so that stencil rows are coalesced; the t-major layout is
 * accepted and returned at this boundary (PASC_LAYOUT_TMAJOR) and transposed on the device.
 */
#ifndef PASC_HIP_H
#define PASC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PASC_OK = 0,
    PASC_ERR_ARG = 1,      /* invalid argument (sizes, null pointers, unsupported K) */
    PASC_ERR_HIP = 2,      /* a HIP runtime call failed */
    PASC_ERR_ALLOC = 3,    /* device or host allocation failed */
    PASC_ERR_STATE = 4,    /* call order violated (e.g. solve before theta was set) */
    PASC_ERR_NUMERIC = 5   /* NaN/Inf met in the iteration */
} pasc_status;

typedef enum { PASC_LAYOUT_TMAJOR = 0, PASC_LAYOUT_KMAJOR = 1 } pasc_layout;

typedef enum {
    PASC_FP64 = 0,        /* storage, stencil, projection and reductions in FP64 */
    PASC_FP32_BULK = 1    /* FP32 storage and bulk arithmetic, FP64 reductions and scalars,
                             finished by FP64 polish iterations that certify the KKT target */
} pasc_precision;

typedef enum { PASC_DTYPE_F64 = 0, PASC_DTYPE_F32 = 1 } pasc_dtype;

/* stopping criteria (spgqpsolver.cpp:260-274) plus the natural KKT residual */
enum {
    PASC_STOP_DIFFF = 1u,          /* |f - f_old| < eps                  (spgqpsolver_stop_difff) */
    PASC_STOP_NORMGP = 2u,         /* dd < eps                           (spgqpsolver_stop_normgp) */
    PASC_STOP_NORMGP_NORMB = 4u,   /* dd < eps*||b||                     (..._stop_normgp_normb) */
    PASC_STOP_ANORMGP = 8u,        /* dAd < eps                          (..._stop_Anormgp) */
    PASC_STOP_ANORMGP_NORMB = 16u, /* dAd < eps*||b||                    (..._stop_Anormgp_normb) */
    PASC_STOP_KKT = 32u            /* max|gamma - P(gamma - s*g)| <= kkt_tol, s = T (scalef) */
};

/* stop_reason values in pasc_spg_result */
enum {
    PASC_REASON_MAXIT = 0, PASC_REASON_DIFFF = 1, PASC_REASON_NORMGP = 2,
    PASC_REASON_NORMGP_NORMB = 3, PASC_REASON_ANORMGP = 4, PASC_REASON_ANORMGP_NORMB = 5,
    PASC_REASON_KKT = 6, PASC_REASON_STATIONARY = 7, PASC_REASON_NUMERIC = 8
};

typedef struct pasc_ctx pasc_ctx;
typedef struct pasc_problem pasc_problem;
typedef struct pasc_batch pasc_batch;

/* Problem description (replaces TSData + Decomposition + GraphH1FEMModel construction:
 * tsdata.cpp:151-193, decomposition.cpp:6-67, graphh1fem.cpp:7-65). */
typedef struct {
    int64_t T;                 /* time steps */
    int64_t R;                 /* graph nodes (1 for a 1-D signal) */
    int32_t K;                 /* clusters, 1..PASC_MAX_K */
    int32_t xdim;              /* data dimension (reference: 1; xdim > 1 is an extension) */
    int32_t scalef;            /* graphh1femmodel_scalef (default 1): A = eps/T*A0, b /= T */
    int32_t precision;         /* pasc_precision */
    const double *x;           /* host data, [T][R][xdim]; NULL: no data term -- the QP is then
                                  defined by an explicit b (pasc_set_linear_term) only */
    const int64_t *nbr_ptr;    /* optional CSR adjacency of the R nodes (R+1 entries; symmetric,
                                  no self loops): the spatio-temporal graph of
                                  blockgraphsparse.cpp:48-109 on T steps of R nodes */
    const int32_t *nbr_idx;    /* neighbour ids */
    const uint8_t *heldout;    /* optional [T*R] mask: held-out nodes carry no data term */
    int32_t grid_width;        /* > 0: the R nodes form the 4-neighbour grid of width grid_width
                                  (R = width*height; T = 1 for an image, T > 1 a video);
                                  nbr_ptr/nbr_idx are then NULL.  Neither given and R > 1: R
                                  unconnected nodes, each with its own time chain */
    double fem_reduce;         /* graphh1femmodel_fem_reduce: in (0, 1) the gamma QP is solved on
                                  a coarse time grid of T2 = ceil(T*fem_reduce) steps
                                  (graphh1fem.cpp:115-214,236-301); 0 or >= 1: no reduction */
    int32_t fem_type;          /* PASC_FEM_CONSTANT (Fem) or PASC_FEM_HAT (FemHat); 1-D only */
} pasc_problem_desc;

#define PASC_MAX_K 8

/* FEM coarsening of the time axis (graphh1femmodel_fem_type / fem_reduce) */
enum { PASC_FEM_NONE = 0, PASC_FEM_CONSTANT = 1, PASC_FEM_HAT = 2 };
#define PASC_MAX_M 64

/* SPG-QP settings (spgqpsolver.h:19-33, set_settings_from_console :161-202) */
typedef struct {
    int32_t maxit;        /* default 1000 */
    double eps;           /* default 1e-9 */
    int32_t m;            /* non-monotone window, default 20, at most PASC_MAX_M */
    double gamma;         /* default 0.9 */
    double sigma1;        /* default 0 */
    double sigma2;        /* default 1 */
    double alphainit;     /* default 2 */
    uint32_t stop_mask;   /* default PASC_STOP_DIFFF */
    double kkt_tol;       /* default 1e-6 */
    int32_t check_every;  /* iterations between host convergence polls (default 1) */
    int32_t use_graph;    /* capture check_every iterations in a hipGraph (default 0) */
    double fp32_switch;   /* FP32-bulk: hand over to FP64 once the FP32 KKT <= this
                             (default 0 = 100 * kkt_tol; never below the level
                             1e-5 * epsilon where FP32 storage stalls SPG) */
    int32_t stall_window; /* FP32-bulk: iterations without a 10% improvement of the best KKT
                             residual that end the FP32 phase (default 200) */
    int32_t polish_maxit; /* FP32-bulk: maximum FP64 polish iterations; 0 (default): maxit;
                             -1: the FP32 phase and the polish share maxit (the polish stops
                             once FP32 + FP64 iterations reach maxit) */
    int32_t kkt_every;    /* evaluate the KKT residual every kkt_every iterations (default 1);
                             the KKT stop then fires at the first evaluated iteration */
} pasc_spg_params;

typedef struct {
    int32_t it;           /* SPG iterations of this solve (QPSolver::it_last) */
    int32_t hessmult;     /* Hessian products (1 + it, as spgqpsolver.cpp:125,167) */
    int32_t status;       /* pasc_status of this solve */
    int32_t stop_reason;  /* PASC_REASON_* */
    int32_t it_fp32;      /* iterations spent in the FP32 phase (FP32-bulk) */
    int32_t it_fp64;      /* FP64 iterations (all of them for PASC_FP64) */
    double fx;            /* final objective 1/2 x'Ax - b'x (FP64 reduction) */
    double dd;            /* gP = dot(d,d) of the last direction */
    double dAd;
    double kkt;           /* natural KKT residual of the returned gamma (FP64 arithmetic) */
    double normb;         /* ||b||_2 */
    double time_s;        /* wall time of the solve on the device (hipEvent) */
    double pass_time_fp32_s; /* with profiling on: summed device time of the fused SPG pass */
    double pass_time_fp64_s; /*   kernels of each phase (hipEvents around every launch) */
    int32_t passes_fp32;     /* fused passes launched in each phase */
    int32_t passes_fp64;
} pasc_spg_result;

/* Outer alternation settings (tssolver.h:13-15, TSSolver(tsdata, annealing) :81) */
typedef struct {
    int32_t maxit;             /* default 300 */
    double eps;                /* default 1e-6, |L - L_old| < eps */
    int32_t theta_fixed;       /* TSSolver::set_solution_theta: theta is never updated */
    int32_t skip_first_theta;  /* start with a gamma step from the given theta (first run only) */
    int32_t annealing;         /* TSSolver::set_annealing (tssolver.h:486-488): runs of the
                                  alternation, default 1; each run after the first restarts from
                                  a random gamma and the run of the smallest AIC is kept */
    uint64_t restart_seed;     /* seed of the restarts' random gamma (pasc_generate_uniform);
                                  default 13, the seed of GeneralVector::set_random */
} pasc_ts_params;

typedef struct {
    int32_t it;            /* TSSolver::it_last of the last run: the index of the iteration whose
                              |L - L_old| < eps broke the loop (tssolver.h:642-650), or maxit */
    int32_t spg_it_sum;    /* QPSolver::it_sum: SPG iterations of every gamma solve of the call */
    double L;              /* model value of the kept run (GraphH1FEMModel::get_L = gamma fx) */
    double deltaL;
    int32_t increased;     /* times L increased ("ERROR: objective function increased") */
    double aic;            /* GraphH1FEMModel::get_aic of the kept run: 2 log L + K */
    int32_t it_sum;        /* TSSolver::it_sum: it_last summed over the runs */
    int32_t best_run;      /* index of the kept run (0 without annealing) */
} pasc_ts_result;

/* ---- errors / version ---- */
const char *pasc_last_error(void);
int pasc_version(void);

/* ---- context (replaces Initialize<PetscVector>, initialize.cpp:11-57) ---- */
int pasc_ctx_create(int device, pasc_ctx **out);
int pasc_ctx_destroy(pasc_ctx *ctx);
/* Use an external HIP stream (hipStream_t cast to void*), or NULL for the context's own. */
int pasc_ctx_set_stream(pasc_ctx *ctx, void *stream);
int pasc_ctx_synchronize(pasc_ctx *ctx);

/* ---- problem (TSData/QPData/GraphH1FEMModel) ---- */
int pasc_problem_create(pasc_ctx *ctx, const pasc_problem_desc *desc, pasc_problem **out);
int pasc_problem_destroy(pasc_problem *p);
/* TSSolver::set_solution_theta (tssolver.h:777-790) writes theta; theta is [K][xdim] */
int pasc_set_theta(pasc_problem *p, const double *theta);
int pasc_get_theta(pasc_problem *p, double *theta);
/* TSData::load_gammavector (tsdata.cpp:548-627) / the gamma vector read-out */
int pasc_set_gamma(pasc_problem *p, const double *gamma, int layout);
int pasc_set_gamma_uniform(pasc_problem *p);
int pasc_get_gamma(pasc_problem *p, double *gamma, int layout);
/* GraphH1FEMModel::set_epssqr (graphh1fem.h:286-306) */
int pasc_set_epssqr(pasc_problem *p, double epssqr);

/* ---- the gamma step ----
 * updatebeforesolve_gammasolver (graphh1fem.cpp:236-289) + SPGQPSolver<PetscVector>::solve
 * (spgqpsolver.cpp:55-312) + updateaftersolve_gammasolver (:291-301).  b is fused on the fly. */
void pasc_spg_default_params(pasc_spg_params *p);
int pasc_gamma_solve(pasc_problem *p, const pasc_spg_params *params, pasc_spg_result *res);

/* ---- explicit linear term: the b of QPData (qpdata->get_b(), spgqpsolver.cpp:60,104-106) ----
 * With b set, pasc_gamma_solve minimises 1/2 g'Ag - b'g over the simplices with this b in place
 * of the data term -(x - theta_k)^2/T formed from x and theta (graphh1fem.cpp:236-289); A is
 * still alpha*A0 with alpha from pasc_set_epssqr.  b has T*R*K entries in the given layout;
 * b == NULL returns to the data term.  pasc_ts_solve and pasc_adjoint reject a problem with an
 * explicit b. */
int pasc_set_linear_term(pasc_problem *p, const double *b, int layout);

/* ---- replace the data: the datavector edits of TSData::cutdata / scaledata / unscaledata
 * (tsdata.cpp:384-478,511-545) made on the host and uploaded.  x has the layout and size of
 * pasc_problem_desc.x ([T][R][xdim], finite); theta, gamma and epssqr are kept, and later solves
 * and theta steps use the new data. */
int pasc_set_data(pasc_problem *p, const double *x);

/* ---- the theta step: updatebeforesolve_thetasolver (graphh1fem.cpp:305-396), FP64 sums ---- */
int pasc_theta_update(pasc_problem *p);

/* ---- the theta step without device allocations ----
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  The same theta as pasc_theta_update, bit for bit,
 * but the partial sums live in a static device array of 4 MiB per device (set up by
 * pasc_ctx_create, held until the process ends) instead of the problem's scratch buffer, which
 * pasc_theta_update grows with hipMalloc.  Problems of T*R above
 * 256 * (524288 / (K (1 + xdim)) - 1) steps (22.4 million at K = 3, xdim = 1) are rejected with
 * PASC_ERR_ARG.  All calls on one device share its array: calls that overlap on one device
 * race. */
int pasc_theta_update_noalloc(pasc_problem *p);

/* ---- outer alternation: TSSolver<PetscVector>::solve (tssolver.h:491-715) ---- */
void pasc_ts_default_params(pasc_ts_params *p);
int pasc_ts_solve(pasc_problem *p, const pasc_ts_params *tp, const pasc_spg_params *sp,
                  pasc_ts_result *res);

/* ---- model read-outs ----
 * objective f = 1/2 g'Ag - b'g of the current gamma, explicit FP64 (get_fx, spgqpsolver.cpp:316)
 * and its KKT residual */
int pasc_objective(pasc_problem *p, double *fx, double *kkt);
/* Signal1DData::compute_abserr_reconstructed (signal1Ddata.cpp:133-191): ||sum theta_k gamma_k
 * - solution||_1; solution is a host [T*R] array (xdim = 1) */
int pasc_abserr_reconstructed(pasc_problem *p, const double *solution, double *abserr);
/* argmax labels [T*R] (int32); ties resolve to the lowest k (cf. tsdata.cpp:195-235) */
int pasc_get_labels(pasc_problem *p, int32_t *labels);

/* ---- kernel-level entry points on caller-owned device memory ----
 * SimplexFeasibleSet_Local<PetscVector>::project (simplex_local.cpp:37-65, kernel_project
 * simplex_local.cu:70-139): project every K-block (t-major: K consecutive values; k-major:
 * stride ld between the K values) of x in place.  K <= 32 uses one thread per block; larger K
 * uses one wavefront per block. */
int pasc_project_simplex(pasc_ctx *ctx, void *x_dev, int64_t nblocks, int32_t K, int dtype,
                         int layout, int64_t ld);
/* BlockGraphSparseMatrix<PetscVector>::matmult (blockgraphsparse.cpp:151-194): y = alpha*A0*x
 * for the spatio-temporal graph of T steps of R nodes (node n = t*R + r): the 1-D time graph
 * (R = 1, grid_width = 0), R unconnected nodes (grid_width = 0), or the 4-neighbour grid of
 * width grid_width (R = width*height), any T.  x, y are k-major device arrays, each row holding
 * the T*R nodes with stride ld >= T*R. */
int pasc_h1_matvec(pasc_ctx *ctx, const void *x_dev, void *y_dev, int64_t T, int64_t R,
                   int32_t K, int64_t ld, double alpha, int32_t grid_width, int dtype);
/* The same for a general graph given as a CSR adjacency of the R nodes (host arrays,
 * nbr_ptr[R+1], nbr_idx[nbr_ptr[R]]; blockgraphsparse.cpp:48-109 with BGraph's neighbours). */
int pasc_h1_matvec_csr(pasc_ctx *ctx, const void *x_dev, void *y_dev, int64_t T, int64_t R,
                       int32_t K, int64_t ld, double alpha, const int64_t *nbr_ptr,
                       const int32_t *nbr_idx, int dtype);

/* ---- FEM coarsening, kernel level (caller-owned device memory, k-major FP64 rows) ----
 * Fem<PetscVector>::reduce_gamma / prolongate_gamma (fem.cpp:28-166) for PASC_FEM_CONSTANT and
 * FemHat<PetscVector>::reduce_gamma / prolongate_gamma (femhat.cpp:26-217) for PASC_FEM_HAT,
 * one rank: fine rows in[k*ld1 + t1] (t1 < T1) <-> coarse rows out[k*ld2 + t2] (t2 < T2).
 * pasc_fem_reduced_T gives T2 = ceil(T1*fem_reduce) (fem.cpp:198), T1 for fem_reduce >= 1. */
int64_t pasc_fem_reduced_T(int64_t T1, double fem_reduce);
int pasc_fem_reduce(pasc_ctx *ctx, const double *fine_dev, double *coarse_dev, int64_t T1,
                    int64_t T2, int32_t K, int64_t ld1, int64_t ld2, int32_t fem_type);
int pasc_fem_prolongate(pasc_ctx *ctx, const double *coarse_dev, double *fine_dev, int64_t T1,
                        int64_t T2, int32_t K, int64_t ld1, int64_t ld2, int32_t fem_type);

/* ---- batched epsilon scan (new capability; cf. the sequential scan test_signal1D.cpp:268-353) --
 * n_series series (each its own T, K, data and theta), each solved for every epsilon in eps[]:
 * n_series*n_eps independent gamma QPs in one graph-captured launch sequence with per-instance
 * convergence masks.  Series data are host arrays; theta is [K_i] per series (xdim = 1). */
typedef struct {
    int64_t T;
    int32_t K;
    const double *x;
    const double *theta;
} pasc_series_desc;

typedef struct {
    int32_t it;
    int32_t it_fp32;
    int32_t it_fp64;
    int32_t stop_reason;
    int32_t status;
    double fx;
    double kkt;
} pasc_instance_result;

int pasc_batch_create(pasc_ctx *ctx, int32_t n_series, const pasc_series_desc *series,
                      int32_t n_eps, const double *eps, int32_t precision, pasc_batch **out);
int pasc_batch_destroy(pasc_batch *b);
/* solve all instances from gamma0 = 1/K; results are per instance, index = s*n_eps + e */
int pasc_batch_solve(pasc_batch *b, const pasc_spg_params *params, pasc_instance_result *res,
                     double *time_s);
int pasc_batch_get_gamma(pasc_batch *b, int32_t series, int32_t eps_index, double *gamma,
                         int layout);
/* One problem (any graph, xdim, held-out mask; no FEM reduction) solved for n_eps values of
 * epsilon at once: the same batched engine with a single series (series index 0). */
int pasc_batch_create_problem(pasc_ctx *ctx, const pasc_problem_desc *desc,
                              const double *theta, int32_t n_eps, const double *eps,
                              pasc_batch **out);

/* ---- adjoint (new capability) ----
 * Gradient of the held-out reconstruction loss  J = sum_{t in M} (x_t - sum_k theta_k gamma_kt)^2
 * with respect to log(epssqr) and theta, through the converged gamma of the last
 * pasc_gamma_solve (implicit differentiation on the active face of the simplex, projected CG).
 * Any graph of the problem; with xdim > 1 the loss sums over the channels,
 * J = sum_{n in M} sum_c (x_nc - sum_k theta_kc gamma_kn)^2.  mask is [T*R] (1 = held out).
 * dtheta has K*xdim entries (theta layout). */
int pasc_adjoint(pasc_problem *p, const uint8_t *mask, double cg_tol, int32_t cg_maxit,
                 double *loss, double *dlogeps, double *dtheta, int32_t *cg_it);

/* ---- profiling: record hipEvents around every fused-pass launch of later solves (graph
 * capture is disabled meanwhile); the totals are returned in pasc_spg_result. ---- */
int pasc_set_profiling(pasc_problem *p, int enable);

/* ---- device copy bandwidth (STREAM-style copy kernel over two buffers of `bytes` each): the
 * achievable HBM rate of this device, and the calibration kernel for traffic counters ---- */
int pasc_device_copy_bandwidth(pasc_ctx *ctx, int64_t bytes, int32_t reps, double *gbps);

/* ---- launch shape of the fused 1-D pass and of the image pass ----
 * Images (T = 1 on the implicit 4-neighbour grid) run the fused image pass: block must be 64,
 * chunk is the number of image rows per one-wave tile (1..256; default: the rows in 6..24 that
 * spread the tiles most evenly over the SIMDs, 18 on a 1024 x 1024 image) and prefetch the
 * number of raw rows loaded ahead besides the next one (1..3, 0 = 1); max_slots and
 * min_tiles are not used.  Other graphs have a fixed shape. */
typedef struct {
    int32_t block;       /* threads per workgroup: 64 (one wave), 128, 256 or 512 */
    int32_t chunk;       /* time steps per thread: FP32 passes 4 or 8, FP64 passes 2 or 4 */
    int32_t prefetch;    /* 1: asynchronous LDS staging of a workgroup's next tile */
    int32_t max_slots;   /* partial-sum slots per instance at most (0 = 2048); each slot is one
                            workgroup's run of consecutive tiles */
    int32_t min_tiles;   /* tiles per slot at least (0 = 1); staging overlaps from 2 on */
    double us_per_pass;  /* pasc_autotune: measured average pass time (0 when from a cache) */
} pasc_pass_shape;

typedef struct {
    pasc_pass_shape main;    /* FP64 problems: every pass; FP32-bulk: the FP32 phase */
    pasc_pass_shape polish;  /* FP32-bulk: the FP64 polish passes (ignored for FP64 problems);
                                polish.block == 0 derives it from main (chunk / 2) */
} pasc_launch_shape;

/* Pin the launch shapes of later solves (the iterate is kept). */
int pasc_set_launch_shape(pasc_problem *p, const pasc_launch_shape *shape);

/* ---- autotuner ----
 * Times the fused SPG pass for the problem's (K, precision, T) over the candidate launch
 * shapes (block x chunk x prefetch x slots; images: the default rows and 6, 8, 12, 16, 24 rows
 * per tile) and keeps the fastest for later solves; for an FP32-bulk problem the FP32 phase and
 * the FP64 polish are tuned separately.
 * The choice is cached per context.  With a non-NULL cache_path, a choice recorded there for
 * the same key by an earlier process is reused without timing, and every call appends its
 * choice to the file (one JSON object per line). */
int pasc_autotune(pasc_problem *p, const char *cache_path, pasc_launch_shape *shape);
/* the same for a batch (all instances share one shape per phase) */
int pasc_batch_autotune(pasc_batch *b, const char *cache_path, pasc_launch_shape *shape);
int pasc_batch_set_launch_shape(pasc_batch *b, const pasc_launch_shape *shape);

/* ---- synthetic data (deterministic SplitMix64 + Box-Muller generator) ---- */
int pasc_generate_markov(uint64_t seed, int64_t T, int32_t K, const double *levels,
                         double p_switch, double sigma, double *x, int32_t *labels);
int pasc_generate_mask(uint64_t seed, int64_t n, double fraction, uint8_t *mask);
/* n uniforms in [0, 1) (SplitMix64 z, (z >> 11) * 2^-53): the random gamma of the annealing
 * restarts, in place of GeneralVector<PetscVector>::set_random (generalvector.cpp:7-29) */
int pasc_generate_uniform(uint64_t seed, int64_t n, double *out);
int pasc_generate_rect_image(uint64_t seed, int32_t width, int32_t height, int32_t K,
                             int32_t n_rect, double sigma, double *x, double *colours,
                             int32_t *labels);
/* per-series parameters of the uneven batch: T_i = round(10^U[lg_tmin,lg_tmax]), K_i uniform
 * in {kmin..kmax}, levels U[0,level_max], p U[pmin,pmax], sigma U[smin,smax], series seed */
int pasc_generate_batch_params(uint64_t seed, int32_t n_series, double lg_tmin, double lg_tmax,
                               int32_t kmin, int32_t kmax, double level_max, double pmin,
                               double pmax, double smin, double smax, int64_t *T, int32_t *K,
                               double *levels /* n_series*PASC_MAX_K */, double *p_switch,
                               double *sigma, uint64_t *series_seed);

#ifdef __cplusplus
}
#endif

#endif

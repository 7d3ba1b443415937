// Internal declarations shared by the HIP sources of libpasc_hip.so.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pasc_hip.h"

namespace pasc {

// Sets the thread-local error message and returns code (printf-style message).
int fail(int code, const char *fmt, ...);
void clear_error();

constexpr int NPART = 5;  // per-tile partial sums: dd, dAd, gd, f, kkt

// Spatio-temporal graph of blockgraphsparse.cpp:48-109: node n = t*R + r (t < T, r < R); the
// spatial neighbours of r come from a symmetric CSR adjacency (ptr/idx, device arrays) or, with
// ptr == nullptr and width > 0, from the 4-neighbour grid of that width (none otherwise).
struct GraphGeo {
    int64_t T, R;
    int32_t width;
    const int64_t *ptr;
    const int32_t *idx;
};

// Read-only description of one gamma QP instance on the device.
struct InstDesc {
    int64_t T;        // time steps (1-D) or nodes (grid graph, T == 1 in the reference sense)
    int64_t ld;       // row stride of the k-major gamma rows (multiple of 64 elements)
    int64_t goff;     // offset of row 0 in the gamma state buffers, in elements
    int64_t xoff;     // offset of the instance's data in the data buffer, in elements
    int32_t K;
    int32_t tile0;    // first partial slot of this instance
    int32_t ntiles;   // partial slots used per pass (each covers msub consecutive tiles)
    int32_t xdim;
    int32_t msub;     // tiles per partial slot
    double alpha;     // A = alpha * A0
    double bscale;    // b = -bscale * ||x - theta_k||^2
    double s_kkt;     // step of the natural KKT residual
    const uint8_t *heldout;  // device mask or nullptr
    // explicit linear term b (k-major rows of stride ld, like gamma) or nullptr: the data term
    // of the iteration is then -b instead of bscale*||x - theta_k||^2
    const double *b64;
    const float *b32;
    double theta[PASC_MAX_K * 3];  // theta[k*xdim + c], xdim <= 3
};

// SPG state of one instance, advanced on the device by step_scalar (in the step kernel, or in
// the 1-D pass's last workgroup for single problems).
struct InstState {
    double f, f_prev;
    double dd_prev, dAd_prev;  // dots of the direction of the last completed iteration
    double dd, dAd, gd;        // dots of the current direction
    double alpha_dir;          // BB step defining the current direction d_i
    double alpha_apply;        // BB step that defined d_{i-1} (recomputed in the next pass)
    double beta_apply;         // beta_{i-1}, applied to d_{i-1} at the start of the next pass
    double kkt, best_kkt;
    double normb;
    double fs[PASC_MAX_M];
    int32_t fs_last;
    int32_t it;                // completed reference iterations
    int32_t first;             // next pass starts from an unprojected gamma_0
    int32_t active;
    int32_t reason;
    int32_t status;
    int32_t stall;
    int32_t parity;            // state buffer holding the current iterate
    int32_t it_base;           // iterations of an earlier phase counted against maxit
};

// Completed-slot counters of the fused step (single 1-D problems): slot j counts on sub-counter
// j % DONE_SUB (its group), the last arrival of each sub-counter on the top counter.  One
// counter of all slots would serialise the atomics (~10 ns each on one address, 2048 slots per
// pass); each counter sits on its own 128-byte line.  The group partials (DONE_SUB x NPART
// doubles) follow the counters in the same allocation.
constexpr int DONE_SUB = 64;
constexpr int DONE_STRIDE = 32;  // int32 words per counter
constexpr int DONE_WORDS = (DONE_SUB + 1) * DONE_STRIDE;
constexpr size_t DONE_BYTES = DONE_WORDS * sizeof(int32_t) + DONE_SUB * NPART * sizeof(double);

// Solver constants of the step (kernel arguments of the step kernel and of the fused pass).
struct StepParams {
    int32_t maxit;
    int32_t m;
    double eps;
    double gamma, sigma1, sigma2, alphainit;
    uint32_t stop_mask;
    double kkt_tol;
    int32_t fp32_phase;   // 1: FP32 phase (stagnation / switch rules apply)
    int32_t stall_window;
    double fp32_switch;
    int32_t fx_identity;  // f_i = f_{i-1} + beta (gd + beta dAd / 2) instead of the FP32 sum
};

// STENCIL_TIME: 1-D time graph (R = 1, fused pass); STENCIL_GRID: any other graph of
// blockgraphsparse.cpp (4-neighbour image grid or CSR adjacency, T >= 1), node-wise kernels
enum StencilKind { STENCIL_TIME = 0, STENCIL_GRID = 1 };

// Launch shape of the fused 1-D pass for one phase
struct PassShape {
    int block = 256;       // threads per block
    int chunk = 2;         // time steps per thread (of the phase's type)
    int prefetch = 0;      // 1: asynchronous LDS staging of the next tile
    int max_slots = 2048;  // partial slots per instance at most
    int min_tiles = 1;     // tiles per partial slot at least (lengthens the staging pipeline)
};

struct Ctx;

// Engine: one or many instances sharing state buffers (k-major rows, padded).
struct Engine {
    Ctx *ctx = nullptr;
    int precision = PASC_FP64;  // storage type of the main state buffers
    int stencil = STENCIL_TIME;
    int32_t grid_width = 0;
    int n_inst = 0;
    int64_t state_elems = 0;     // elements per state buffer
    int64_t data_elems = 0;
    int total_slots = 0;
    std::vector<InstDesc> h_desc;
    std::vector<int> h_parity;   // buffer holding each instance's current iterate
    bool cur64 = false;          // FP32-bulk: the current iterates live in the FP64 buffers
    // device buffers
    void *d_gamma[2] = {nullptr, nullptr};    // ping-pong state buffers (precision type)
    void *d_gamma64[2] = {nullptr, nullptr};  // FP64 polish buffers (FP32-bulk only)
    GraphGeo geo{};                           // graph stencil: T, R, grid width / CSR
    int64_t *d_nbr_ptr = nullptr;             // CSR adjacency of the R nodes (device)
    int32_t *d_nbr_idx = nullptr;
    void *d_dir[2] = {nullptr, nullptr};      // grid stencil: direction ping-pong
    void *d_dir64[2] = {nullptr, nullptr};
    void *d_q[2] = {nullptr, nullptr};        // image pass: data term rows, [0] FP32, [1] FP64
    void *d_x64 = nullptr;                    // data, FP64, rows of stride ld
    void *d_x32 = nullptr;                    // data, FP32 copy (FP32-bulk only)
    void *d_b64 = nullptr;                    // explicit linear term (single instance), FP64
    void *d_b32 = nullptr;                    //   and its FP32 copy (FP32-bulk only)
    InstDesc *d_desc = nullptr;
    InstState *d_state = nullptr;
    double *d_part = nullptr;                 // total_slots * NPART
    void *d_scratch = nullptr;
    size_t scratch_bytes = 0;
    // K groups for the fused passes
    struct Group {
        int K = 0;
        int n = 0;
        int64_t max_tiles = 0;
        int32_t *d_ids = nullptr;       // instance ids of the group
        int32_t *d_off = nullptr;       // n+1 exclusive scan of active tile counts
        int32_t *d_hint = nullptr;      // tile >> 6 -> first candidate position
        int hint_len = 0;
    };
    std::vector<Group> groups;
    int32_t *d_nactive = nullptr;        // device counter of active instances
    int32_t *d_done = nullptr;           // completed-slot counters and group partials of the
                                         // fused step (DONE_BYTES, single problems only)
    int32_t *h_nactive = nullptr;        // pinned mirror
    PassShape sh[2];                     // [0] FP32 phase of FP32-bulk, [1] every FP64 pass
    int grid_cap = 0;                    // persistent grid cap (0: 32 blocks per CU)
    int cur_tile = 0;                    // time steps per tile the partial slots are sized for
    int cur_phase = -1;                  // shape index the partial slots are sized for
    bool profile = false;                // hipEvents around every pass launch
    double prof_ms[2] = {0.0, 0.0};      // summed pass time per phase (FP32, FP64)
    int prof_n[2] = {0, 0};              // passes launched per phase
};

struct Ctx {
    int device = 0;
    void *stream = nullptr;     // hipStream_t
    bool own_stream = false;
    int n_cu = 0;
    std::string arch;
    // autotune cache: key -> launch shapes of the two phases
    std::vector<std::pair<std::string, pasc_launch_shape>> tune;
};

// engine.hip
int engine_alloc(Engine &e, Ctx *ctx, int precision, int stencil, int32_t grid_width,
                 const std::vector<InstDesc> &descs, int64_t data_elems);
void engine_free(Engine &e);
int engine_upload_data(Engine &e, const double *x_host, int64_t offset, int64_t n);
// graph stencil geometry: T time steps of R nodes, grid width or CSR adjacency (host arrays)
int engine_set_graph(Engine &e, int64_t T, int64_t R, int32_t width, const int64_t *nbr_ptr,
                     const int32_t *nbr_idx);
int engine_set_linear_term(Engine &e, int inst, const double *b_host, int layout);
int engine_set_uniform(Engine &e, int inst);
int engine_fill_uniform(Engine &e);
int engine_set_gamma(Engine &e, int inst, const double *gamma_host, int layout);
int engine_get_gamma(Engine &e, int inst, double *gamma_host, int layout);
int engine_solve(Engine &e, const pasc_spg_params &p, pasc_spg_result *res_single,
                 pasc_instance_result *res_multi, double *time_s);
int engine_eval(Engine &e, int inst, double *fx, double *kkt);
int engine_eval_raw(Engine &e, int inst, double *fx, double *kkt, double *nb2);
int engine_update_desc(Engine &e);
int engine_theta_update(Engine &e, int inst, double *theta_out);
int engine_labels(Engine &e, int inst, int32_t *labels);
int engine_abserr(Engine &e, int inst, const double *solution, double *abserr);
int engine_adjoint(Engine &e, int inst, const uint8_t *mask, double cg_tol, int cg_maxit,
                   double *loss, double *dlogeps, double *dtheta, int *cg_it);
// us[0] / us[1]: measured pass time of the chosen FP32 / FP64 shape (0 when not measured)
int engine_autotune(Engine &e, PassShape *s32, PassShape *s64, double us[2]);
int engine_set_shape(Engine &e, const PassShape &s32, const PassShape &s64);
// the graph is an image that the fused image pass serves
bool engine_is_image(const Engine &e);

int device_copy_bandwidth(Ctx *ctx, int64_t bytes, int reps, double *gbps);

// device-side hand-over between the engine and code outside it (FEM coarsening)
int engine_linear_term_rows(Engine &e, int inst, double **rows);  // FP64 rows to fill
int engine_commit_linear_term(Engine &e, int inst);               // FP32 copy + descriptor
int engine_set_gamma_rows(Engine &e, int inst, const double *rows_dev);
const void *engine_current_gamma(Engine &e, int inst, bool *f32);
// theta step, labels and abserr on gamma rows described by d (gamma g of the given type,
// data rows x of stride d.ld); theta_out receives theta[k*xdim + c]
int rows_theta(Engine &e, const void *g, bool f32, const InstDesc &d, const double *x,
               double *theta_out);
int rows_labels(Engine &e, const void *g, bool f32, const InstDesc &d, int32_t *labels);
int rows_abserr(Engine &e, const void *g, bool f32, const InstDesc &d, const double *solution,
                double *abserr);

// noalloc.hip: the theta step with its partial sums in a static device array
int noalloc_init(int device);  // once per device, with the device current
int64_t noalloc_theta_max_T(int K, int xdim);
int rows_theta_noalloc(Engine &e, const void *g, bool f32, const InstDesc &d, const double *x,
                       double *theta_out);

// fem.hip
struct FemGeom {
    int type;            // PASC_FEM_CONSTANT or PASC_FEM_HAT
    int64_t T1, T2;      // fine and coarse steps
    int64_t ld1, ld2;    // row strides
    double diff;
    int64_t right_idx;   // FemHat: fine steps at or beyond it leave the right half of a hat
};
FemGeom fem_geom(int type, int64_t T1, int64_t T2, int64_t ld1, int64_t ld2);
int fem_check_geom(int type, int64_t T1, int64_t T2);
int fem_reduce_rows(void *stream, const FemGeom &g, int K, const double *in, double *out,
                    double scale);
int fem_residual_rows(void *stream, const FemGeom &g, int K, const double *x, int xdim,
                      const double *theta, const uint8_t *heldout, double *out, double scale);
int fem_prolongate_rows(void *stream, const FemGeom &g, int K, const void *in, bool in_f32,
                        double *out);
int kernel_project(Ctx *ctx, void *x, int64_t nblocks, int K, int dtype, int layout, int64_t ld);
int kernel_matvec(Ctx *ctx, const void *x, void *y, int64_t T, int64_t R, int K, int64_t ld,
                  double alpha, int32_t grid_width, const int64_t *nbr_ptr_dev,
                  const int32_t *nbr_idx_dev, int dtype);

}  // namespace pasc

// Host side of the SPG engine: buffers, kernel dispatch, the convergence loop (optionally
// hipGraph-captured), FP32 -> FP64 polish, and the auxiliary device computations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "img_kernels.h"
#include "spg_kernels.h"
#include "step_kernels.h"
#include "util_kernels.h"

namespace pasc {

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(PASC_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                              \
    } while (0)

static hipStream_t S_(Engine &e) { return (hipStream_t)e.ctx->stream; }

static int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// ------------------------------------------------------------------------------------------
// kernel dispatch tables
// ------------------------------------------------------------------------------------------
// instantiated in pass_inst.hip / grid_inst.hip (one object per storage type and block size)
int launch_pass_1d_float_64_(int K, const PassArgs &a, int grid, int chunk, int pf, hipStream_t s);
int launch_pass_1d_float_128_(int K, const PassArgs &a, int grid, int chunk, int pf, hipStream_t s);
int launch_pass_1d_float_256_(int K, const PassArgs &a, int grid, int chunk, int pf, hipStream_t s);
int launch_pass_1d_float_512_(int K, const PassArgs &a, int grid, int chunk, int pf, hipStream_t s);
int launch_pass_1d_double_64_(int K, const PassArgs &a, int grid, int chunk, int pf, hipStream_t s);
int launch_pass_1d_double_128_(int K, const PassArgs &a, int grid, int chunk, int pf, hipStream_t s);
int launch_pass_1d_double_256_(int K, const PassArgs &a, int grid, int chunk, int pf, hipStream_t s);
int launch_pass_1d_double_512_(int K, const PassArgs &a, int grid, int chunk, int pf, hipStream_t s);
int launch_pass_grid_float(int K, const GridArgs &a, int grid, hipStream_t s);
int launch_pass_grid_double(int K, const GridArgs &a, int grid, hipStream_t s);
int launch_pass_img_float(int K, const ImgArgs &a, int grid, int pd, hipStream_t s);
int launch_pass_img_double(int K, const ImgArgs &a, int grid, int pd, hipStream_t s);

// Images (T == 1 on the implicit 4-neighbour grid, whole rows, fewer than 2^31 nodes, every
// instance's K rows within 4 GiB: the pass addresses them with 32-bit byte offsets) run the
// fused image pass; its wave tiles are IMG_OWN columns x sh.chunk image rows, with sh.prefetch
// raw rows in flight besides the next one.  Larger images run the node-wise graph kernels.
bool engine_is_image(const Engine &e) {
    const GraphGeo &g = e.geo;
    if (!(e.stencil == STENCIL_GRID && g.T == 1 && g.width > 0 && !g.ptr && g.R % g.width == 0 &&
          g.R < (int64_t(1) << 31)))
        return false;
    for (const auto &d : e.h_desc)
        if ((int64_t)d.K * d.ld * 8 >= (int64_t(1) << 32)) return false;
    return true;
}

static int img_ncol(const Engine &e) { return (e.geo.width + IMG_OWN - 1) / IMG_OWN; }
static int img_nrow(const Engine &e, int rows) {
    const int64_t h = e.geo.R / e.geo.width;
    return (int)((h + rows - 1) / rows);
}

// Image rows per wave tile that spread the one-wave tiles evenly over the SIMDs (4 per CU).  A
// SIMD that holds m tiles walks m x (rows + 1) image rows (the last one forms the halo row),
// and the pass lasts about as long as the longest such walk: the rows in [6, 24] with the least
// m (rows + 1), the most rows on a tie (fewer partial slots).  On the 1024 x 1024 image (17
// column tiles, 1024 SIMDs) that is 18 rows (969 tiles, cost 19); at a fixed 600 FP64 passes
// 18 / 9 / 10 / 11 / 12 rows (cost 19 / 20 / 22 / 24 / 26) took 27.5 / 27.7 / 27.9 / 28.4 /
// 28.6 us per pass, the order of the costs.  Small images (at most one tile per SIMD at any
// rows) get 6 rows: shorter walks.
static int img_balanced_rows(const Engine &e) {
    const int64_t nsimd = (int64_t)std::max(e.ctx->n_cu, 1) * 4;
    int best = 6;
    int64_t best_cost = std::numeric_limits<int64_t>::max();
    for (int r = 6; r <= 24; r++) {
        const int64_t tiles = (int64_t)img_ncol(e) * img_nrow(e, r);
        const int64_t cost = (tiles + nsimd - 1) / nsimd * (r + 1);
        if (cost <= best_cost) {
            best_cost = cost;
            best = r;
        }
    }
    return best;
}

static int launch_pass_1d(bool f64, int K, const PassArgs &a, int grid, int block, int chunk,
                          int pf, hipStream_t s) {
    if (f64) {
        if (block == 64) return launch_pass_1d_double_64_(K, a, grid, chunk, pf, s);
        if (block == 128) return launch_pass_1d_double_128_(K, a, grid, chunk, pf, s);
        if (block == 256) return launch_pass_1d_double_256_(K, a, grid, chunk, pf, s);
        if (block == 512) return launch_pass_1d_double_512_(K, a, grid, chunk, pf, s);
    } else {
        if (block == 64) return launch_pass_1d_float_64_(K, a, grid, chunk, pf, s);
        if (block == 128) return launch_pass_1d_float_128_(K, a, grid, chunk, pf, s);
        if (block == 256) return launch_pass_1d_float_256_(K, a, grid, chunk, pf, s);
        if (block == 512) return launch_pass_1d_float_512_(K, a, grid, chunk, pf, s);
    }
    return fail(PASC_ERR_ARG, "no fused pass for block %d", block);
}

// ------------------------------------------------------------------------------------------
// allocation
// ------------------------------------------------------------------------------------------
// time steps per tile of the pass kernels in a phase (shape index 1: FP64 passes)
static int tile_len(const Engine &e, bool f64_phase) {
    if (engine_is_image(e)) return -e.sh[f64_phase ? 1 : 0].chunk;  // image rows per tile
    if (e.stencil != STENCIL_TIME) return 256;
    const PassShape &sh = e.sh[f64_phase ? 1 : 0];
    const int ht = (3 + sh.chunk - 1) / sh.chunk;  // halo threads per side (see tile_steps)
    return (sh.block - 2 * ht) * sh.chunk;
}

static int engine_retile(Engine &e, bool f64_phase) {
    const PassShape &sh = e.sh[f64_phase ? 1 : 0];
    const int tl = tile_len(e, f64_phase);
    e.cur_tile = tl;
    e.cur_phase = f64_phase ? 1 : 0;
    int slots = 0;
    const bool img = engine_is_image(e);
    for (auto &d : e.h_desc) {
        if (img) {
            // one partial slot per wave tile
            d.msub = 1;
            d.tile0 = slots;
            d.ntiles = img_ncol(e) * img_nrow(e, sh.chunk);
            slots += d.ntiles;
            continue;
        }
        // at most max_slots (default 2048) partial slots per instance: enough workgroups to fill
        // the device, few enough that the step reduces them quickly; at least min_tiles
        // tiles per slot
        const int64_t raw = (d.T + tl - 1) / tl;
        const int64_t ms = std::max(sh.max_slots, 1);
        d.msub = (int32_t)std::max<int64_t>(std::max(sh.min_tiles, 1), (raw + ms - 1) / ms);
        d.msub = (int32_t)std::min<int64_t>(d.msub, std::max<int64_t>(raw, 1));
        d.tile0 = slots;
        d.ntiles = (int32_t)((raw + d.msub - 1) / d.msub);
        slots += d.ntiles;
    }
    if (slots != e.total_slots || !e.d_part) {
        if (e.d_part) (void)hipFree(e.d_part);
        e.d_part = nullptr;
        e.total_slots = slots;
        HIP_TRY(hipMalloc(&e.d_part, sizeof(double) * (size_t)std::max(slots, 1) * NPART));
        // every partial starts unpublished (the fused step's hand-over, take_published)
        const int64_t nw = (int64_t)std::max(slots, 1) * NPART;
        hipLaunchKernelGGL(fill_u64, dim3((unsigned)std::min<int64_t>((nw + 255) / 256, 1024)),
                           dim3(256), 0, S_(e), (uint64_t *)e.d_part, nw, PART_EMPTY);
        HIP_TRY(hipGetLastError());
    }
    for (auto &g : e.groups) {
        std::vector<int32_t> ids;
        int64_t tiles = 0;
        for (int i = 0; i < e.n_inst; i++)
            if (e.h_desc[i].K == g.K) { ids.push_back(i); tiles += e.h_desc[i].ntiles; }
        g.n = (int)ids.size();
        g.max_tiles = tiles;
        const int hl = (int)(tiles / 64 + 2);
        if (hl > g.hint_len) {
            if (g.d_hint) (void)hipFree(g.d_hint);
            g.d_hint = nullptr;
            HIP_TRY(hipMalloc(&g.d_hint, sizeof(int32_t) * hl));
            g.hint_len = hl;
        }
        HIP_TRY(hipMemsetAsync(g.d_hint, 0, sizeof(int32_t) * g.hint_len, S_(e)));
    }
    HIP_TRY(hipMemcpyAsync(e.d_desc, e.h_desc.data(), sizeof(InstDesc) * e.n_inst,
                           hipMemcpyHostToDevice, S_(e)));
    HIP_TRY(hipStreamSynchronize(S_(e)));
    return PASC_OK;
}

int engine_update_desc(Engine &e) {
    HIP_TRY(hipMemcpyAsync(e.d_desc, e.h_desc.data(), sizeof(InstDesc) * e.n_inst,
                           hipMemcpyHostToDevice, S_(e)));
    HIP_TRY(hipStreamSynchronize(S_(e)));
    return PASC_OK;
}

int engine_alloc(Engine &e, Ctx *ctx, int precision, int stencil, int32_t grid_width,
                 const std::vector<InstDesc> &descs, int64_t data_elems) {
    e.ctx = ctx;
    e.precision = precision;
    e.stencil = stencil;
    e.grid_width = grid_width;
    e.h_desc = descs;
    e.n_inst = (int)descs.size();
    e.data_elems = data_elems;
    e.sh[0] = PassShape{256, 4, 0, 2048, 1};
    e.sh[1] = PassShape{256, 2, 0, 2048, 1};
    if (stencil != STENCIL_TIME) {
        // graph problems (fixed shape): at most 512 partial slots per instance, i.e. each
        // workgroup walks 8 consecutive 256-node tiles of a 1024^2 image (two image rows).
        // Measured on that image: 59.5 us per iteration against 60.8 (2048) and 63.0 (4096).
        e.sh[0].max_slots = e.sh[1].max_slots = 512;
    }
    int64_t se = 0;
    for (auto &d : descs) se = std::max(se, d.goff + (int64_t)d.K * d.ld);
    e.state_elems = round_up(std::max<int64_t>(se, 64), 64);
    const size_t es = precision == PASC_FP32_BULK ? 4 : 8;
    for (int b = 0; b < 2; b++) {
        HIP_TRY(hipMalloc(&e.d_gamma[b], es * (size_t)e.state_elems));
        HIP_TRY(hipMemsetAsync(e.d_gamma[b], 0, es * (size_t)e.state_elems, S_(e)));
        if (stencil == STENCIL_GRID) {
            HIP_TRY(hipMalloc(&e.d_dir[b], es * (size_t)e.state_elems));
            HIP_TRY(hipMemsetAsync(e.d_dir[b], 0, es * (size_t)e.state_elems, S_(e)));
        }
    }
    const int64_t de = round_up(std::max<int64_t>(data_elems, 64), 64);
    HIP_TRY(hipMalloc(&e.d_x64, 8 * (size_t)de));
    HIP_TRY(hipMemsetAsync(e.d_x64, 0, 8 * (size_t)de, S_(e)));
    if (precision == PASC_FP32_BULK) {
        HIP_TRY(hipMalloc(&e.d_x32, 4 * (size_t)de));
        HIP_TRY(hipMemsetAsync(e.d_x32, 0, 4 * (size_t)de, S_(e)));
    }
    HIP_TRY(hipMalloc(&e.d_desc, sizeof(InstDesc) * std::max(e.n_inst, 1)));
    HIP_TRY(hipMalloc(&e.d_state, sizeof(InstState) * std::max(e.n_inst, 1)));
    HIP_TRY(hipMalloc(&e.d_nactive, 2 * sizeof(int32_t)));  // [0] active count, [1] dirty flag
    HIP_TRY(hipMalloc(&e.d_done, DONE_BYTES));
    HIP_TRY(hipMemsetAsync(e.d_done, 0, DONE_BYTES, S_(e)));
    // the group partials after the counters start unpublished
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(256), 0, S_(e),
                       (uint64_t *)(e.d_done + DONE_WORDS), (int64_t)DONE_SUB * NPART, PART_EMPTY);
    HIP_TRY(hipGetLastError());
    // pinned poll area: [0] active instances, [1 + g] active tiles of K group g
    HIP_TRY(hipHostMalloc((void **)&e.h_nactive, sizeof(int32_t) * (1 + PASC_MAX_K),
                          hipHostMallocDefault));
    e.h_parity.assign(e.n_inst, 0);
    // K groups
    for (int K = 1; K <= PASC_MAX_K; K++) {
        std::vector<int32_t> ids;
        for (int i = 0; i < e.n_inst; i++)
            if (descs[i].K == K) ids.push_back(i);
        if (ids.empty()) continue;
        Engine::Group g;
        g.K = K;
        g.n = (int)ids.size();
        HIP_TRY(hipMalloc(&g.d_ids, sizeof(int32_t) * ids.size()));
        HIP_TRY(hipMemcpyAsync(g.d_ids, ids.data(), sizeof(int32_t) * ids.size(),
                               hipMemcpyHostToDevice, S_(e)));
        HIP_TRY(hipMalloc(&g.d_off, sizeof(int32_t) * (ids.size() + 1)));
        e.groups.push_back(g);
    }
    if (e.groups.size() > (size_t)MAX_GROUPS) return fail(PASC_ERR_ARG, "too many K groups");
    return engine_retile(e, precision != PASC_FP32_BULK);
}

void engine_free(Engine &e) {
    for (int b = 0; b < 2; b++) {
        if (e.d_gamma[b]) (void)hipFree(e.d_gamma[b]);
        if (e.d_gamma64[b]) (void)hipFree(e.d_gamma64[b]);
        if (e.d_dir[b]) (void)hipFree(e.d_dir[b]);
        if (e.d_dir64[b]) (void)hipFree(e.d_dir64[b]);
        if (e.d_q[b]) (void)hipFree(e.d_q[b]);
        e.d_gamma[b] = e.d_gamma64[b] = e.d_dir[b] = e.d_dir64[b] = e.d_q[b] = nullptr;
    }
    if (e.d_x64) (void)hipFree(e.d_x64);
    if (e.d_x32) (void)hipFree(e.d_x32);
    if (e.d_b64) (void)hipFree(e.d_b64);
    if (e.d_b32) (void)hipFree(e.d_b32);
    e.d_b64 = e.d_b32 = nullptr;
    if (e.d_nbr_ptr) (void)hipFree(e.d_nbr_ptr);
    if (e.d_nbr_idx) (void)hipFree(e.d_nbr_idx);
    e.d_nbr_ptr = nullptr;
    e.d_nbr_idx = nullptr;
    if (e.d_desc) (void)hipFree(e.d_desc);
    if (e.d_state) (void)hipFree(e.d_state);
    if (e.d_part) (void)hipFree(e.d_part);
    if (e.d_nactive) (void)hipFree(e.d_nactive);
    if (e.d_done) (void)hipFree(e.d_done);
    e.d_done = nullptr;
    if (e.h_nactive) (void)hipHostFree(e.h_nactive);
    if (e.d_scratch) (void)hipFree(e.d_scratch);
    for (auto &g : e.groups) {
        if (g.d_ids) (void)hipFree(g.d_ids);
        if (g.d_off) (void)hipFree(g.d_off);
        if (g.d_hint) (void)hipFree(g.d_hint);
    }
    e.groups.clear();
    e.d_x64 = e.d_x32 = nullptr;
    e.d_desc = nullptr;
    e.d_state = nullptr;
    e.d_part = nullptr;
    e.d_nactive = nullptr;
    e.h_nactive = nullptr;
    e.d_scratch = nullptr;
}

static int scratch(Engine &e, size_t bytes, void **p) {
    if (bytes > e.scratch_bytes) {
        if (e.d_scratch) (void)hipFree(e.d_scratch);
        e.d_scratch = nullptr;
        e.scratch_bytes = 0;
        HIP_TRY(hipMalloc(&e.d_scratch, bytes));
        e.scratch_bytes = bytes;
    }
    *p = e.d_scratch;
    return PASC_OK;
}

static int alloc_linear_term(Engine &e);

int engine_set_graph(Engine &e, int64_t T, int64_t R, int32_t width, const int64_t *nbr_ptr,
                     const int32_t *nbr_idx) {
    e.geo = GraphGeo{T, R, width, nullptr, nullptr};
    if (nbr_ptr) {
        const int64_t nnz = nbr_ptr[R];
        HIP_TRY(hipMalloc(&e.d_nbr_ptr, sizeof(int64_t) * (size_t)(R + 1)));
        HIP_TRY(hipMalloc(&e.d_nbr_idx, sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)));
        HIP_TRY(hipMemcpy(e.d_nbr_ptr, nbr_ptr, sizeof(int64_t) * (size_t)(R + 1), hipMemcpyHostToDevice));
        if (nnz > 0)
            HIP_TRY(hipMemcpy(e.d_nbr_idx, nbr_idx, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice));
        e.geo.ptr = e.d_nbr_ptr;
        e.geo.idx = e.d_nbr_idx;
        e.geo.width = 0;
    }
    // images: one-wave tiles of img_balanced_rows image rows, one row loaded ahead (a second
    // row ahead measured 0.1 - 0.8 us slower on the 1024 x 1024 image), in both phases;
    // set_launch_shape changes them
    if (engine_is_image(e)) {
        const int rows = img_balanced_rows(e);
        e.sh[0] = e.sh[1] = PassShape{WAVE, rows, 1, 0, 1};
    }
    return PASC_OK;
}

int engine_upload_data(Engine &e, const double *x_host, int64_t offset, int64_t n) {
    HIP_TRY(hipMemcpyAsync((double *)e.d_x64 + offset, x_host, 8 * (size_t)n,
                           hipMemcpyHostToDevice, S_(e)));
    if (e.d_x32) {
        const int64_t blocks = (n + 255) / 256;
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned)blocks), dim3(256), 0,
                           S_(e), (const double *)e.d_x64 + offset, (float *)e.d_x32 + offset, n);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(S_(e)));
    return PASC_OK;
}

// host copy of one instance's k-major rows (as double)
static int download_rows(Engine &e, int inst, std::vector<double> &rows) {
    const InstDesc &d = e.h_desc[inst];
    const int par = e.h_parity[inst];
    const size_t n = (size_t)d.K * d.ld;
    rows.resize(n);
    if (e.cur64) {
        HIP_TRY(hipMemcpyAsync(rows.data(), (double *)e.d_gamma64[par] + d.goff, 8 * n,
                               hipMemcpyDeviceToHost, S_(e)));
    } else if (e.precision == PASC_FP32_BULK) {
        std::vector<float> tmp(n);
        HIP_TRY(hipMemcpyAsync(tmp.data(), (float *)e.d_gamma[par] + d.goff, 4 * n,
                               hipMemcpyDeviceToHost, S_(e)));
        HIP_TRY(hipStreamSynchronize(S_(e)));
        for (size_t i = 0; i < n; i++) rows[i] = tmp[i];
    } else {
        HIP_TRY(hipMemcpyAsync(rows.data(), (double *)e.d_gamma[par] + d.goff, 8 * n,
                               hipMemcpyDeviceToHost, S_(e)));
    }
    HIP_TRY(hipStreamSynchronize(S_(e)));
    return PASC_OK;
}

static int upload_rows(Engine &e, int inst, const std::vector<double> &rows) {
    const InstDesc &d = e.h_desc[inst];
    const size_t n = (size_t)d.K * d.ld;
    e.cur64 = false;
    e.h_parity[inst] = 0;
    if (e.precision == PASC_FP32_BULK) {
        std::vector<float> tmp(n);
        for (size_t i = 0; i < n; i++) tmp[i] = (float)rows[i];
        HIP_TRY(hipMemcpyAsync((float *)e.d_gamma[0] + d.goff, tmp.data(), 4 * n,
                               hipMemcpyHostToDevice, S_(e)));
        HIP_TRY(hipStreamSynchronize(S_(e)));
    } else {
        HIP_TRY(hipMemcpyAsync((double *)e.d_gamma[0] + d.goff, rows.data(), 8 * n,
                               hipMemcpyHostToDevice, S_(e)));
        HIP_TRY(hipStreamSynchronize(S_(e)));
    }
    return PASC_OK;
}

int engine_set_uniform(Engine &e, int inst) {
    const InstDesc &d = e.h_desc[inst];
    std::vector<double> rows((size_t)d.K * d.ld, 0.0);
    for (int k = 0; k < d.K; k++)
        for (int64_t t = 0; t < d.T; t++) rows[k * d.ld + t] = 1.0 / d.K;
    return upload_rows(e, inst, rows);
}

int engine_set_gamma(Engine &e, int inst, const double *g, int layout) {
    const InstDesc &d = e.h_desc[inst];
    std::vector<double> rows((size_t)d.K * d.ld, 0.0);
    for (int k = 0; k < d.K; k++)
        for (int64_t t = 0; t < d.T; t++)
            rows[k * d.ld + t] = layout == PASC_LAYOUT_TMAJOR ? g[t * d.K + k] : g[k * d.T + t];
    return upload_rows(e, inst, rows);
}

// Explicit linear term of one instance (the QP's b as the reference's QPData holds it): stored
// k-major like gamma, FP64 plus an FP32 copy for the FP32 phase.  b_host == nullptr returns the
// instance to the data term formed from x and theta.
int engine_set_linear_term(Engine &e, int inst, const double *b_host, int layout) {
    InstDesc &d = e.h_desc[inst];
    if (!b_host) {
        d.b64 = nullptr;
        d.b32 = nullptr;
        return engine_update_desc(e);
    }
    int arc = alloc_linear_term(e);
    if (arc) return arc;
    const int64_t n = (int64_t)d.K * d.ld;
    std::vector<double> rows((size_t)n, 0.0);
    for (int k = 0; k < d.K; k++)
        for (int64_t t = 0; t < d.T; t++)
            rows[k * d.ld + t] = layout == PASC_LAYOUT_TMAJOR ? b_host[t * d.K + k] : b_host[k * d.T + t];
    double *b64 = (double *)e.d_b64 + d.goff;
    HIP_TRY(hipMemcpyAsync(b64, rows.data(), 8 * (size_t)n, hipMemcpyHostToDevice, S_(e)));
    float *b32 = nullptr;
    if (e.d_b32) {
        b32 = (float *)e.d_b32 + d.goff;
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned)((n + 255) / 256)),
                           dim3(256), 0, S_(e), (const double *)b64, b32, n);
        HIP_TRY(hipGetLastError());
    }
    d.b64 = b64;
    d.b32 = b32;
    return engine_update_desc(e);  // synchronises: rows may go out of scope
}

int engine_get_gamma(Engine &e, int inst, double *g, int layout) {
    const InstDesc &d = e.h_desc[inst];
    std::vector<double> rows;
    int rc = download_rows(e, inst, rows);
    if (rc) return rc;
    for (int k = 0; k < d.K; k++)
        for (int64_t t = 0; t < d.T; t++) {
            if (layout == PASC_LAYOUT_TMAJOR) g[t * d.K + k] = rows[k * d.ld + t];
            else g[k * d.T + t] = rows[k * d.ld + t];
        }
    return PASC_OK;
}

// ------------------------------------------------------------------------------------------
// the iteration loop
// ------------------------------------------------------------------------------------------
struct PhaseCfg {
    bool f64;
    StepParams sp;
    int kkt_every;
};

static int enqueue_iteration(Engine &e, const PhaseCfg &ph, const std::vector<int> &grids,
                             std::vector<hipEvent_t> *ev = nullptr) {
    hipStream_t s = S_(e);
    // single 1-D and image problems run the step inside the pass.  (Other graphs keep the step
    // kernel: run in their dAd kernel, the step's serial tail cost ~8 us per iteration against
    // ~6 us for the separate step and scan launches under graph replay; single graph problems
    // launch no scan.)
    const bool img = engine_is_image(e);
    const bool fused = e.n_inst == 1 && (e.stencil == STENCIL_TIME || img);
    if (ev) {
        hipEvent_t a;
        HIP_TRY(hipEventCreate(&a));
        ev->push_back(a);  // owned by the caller's list from here on
        HIP_TRY(hipEventRecord(a, s));
    }
    void *g0 = ph.f64 && e.precision == PASC_FP32_BULK ? e.d_gamma64[0] : e.d_gamma[0];
    void *g1 = ph.f64 && e.precision == PASC_FP32_BULK ? e.d_gamma64[1] : e.d_gamma[1];
    const void *x = ph.f64 ? (const void *)e.d_x64 : (const void *)e.d_x32;
    for (size_t gi = 0; gi < e.groups.size(); gi++) {
        const auto &g = e.groups[gi];
        const GroupView gv{g.d_ids, g.d_off, g.d_hint, g.n};
        int rc;
        if (e.stencil == STENCIL_TIME) {
            const void *bb = ph.f64 ? e.d_b64 : e.d_b32;
            PassArgs a{g0, g1, x, e.d_desc, e.d_state, e.d_part, gv, ph.kkt_every,
                       ph.sp.fx_identity ? 0 : 1, bb, ph.sp, fused ? e.d_done : nullptr,
                       e.d_nactive, e.d_nactive + 1};
            const PassShape &sh = e.sh[ph.f64 ? 1 : 0];
            rc = launch_pass_1d(ph.f64, g.K, a, grids[gi], sh.block, sh.chunk, sh.prefetch, s);
        } else if (img) {
            void *d0 = ph.f64 && e.precision == PASC_FP32_BULK ? e.d_dir64[0] : e.d_dir[0];
            void *d1 = ph.f64 && e.precision == PASC_FP32_BULK ? e.d_dir64[1] : e.d_dir[1];
            const PassShape &sh = e.sh[ph.f64 ? 1 : 0];
            const int rows = sh.chunk;
            ImgArgs a{g0, g1, d0, d1, e.d_q[ph.f64 ? 1 : 0], e.d_desc, e.d_state, e.d_part, gv,
                      e.geo.width, (int32_t)(e.geo.R / e.geo.width), rows, img_ncol(e),
                      ph.kkt_every, ph.sp, fused ? e.d_done : nullptr, e.d_nactive,
                      e.d_nactive + 1};
            rc = ph.f64 ? launch_pass_img_double(g.K, a, grids[gi], sh.prefetch, s)
                        : launch_pass_img_float(g.K, a, grids[gi], sh.prefetch, s);
        } else {
            void *d0 = ph.f64 && e.precision == PASC_FP32_BULK ? e.d_dir64[0] : e.d_dir[0];
            void *d1 = ph.f64 && e.precision == PASC_FP32_BULK ? e.d_dir64[1] : e.d_dir[1];
            GridArgs a{g0, g1, d0, d1, x, e.d_desc, e.d_state, e.d_part, gv, e.geo};
            rc = ph.f64 ? launch_pass_grid_double(g.K, a, grids[gi], s)
                        : launch_pass_grid_float(g.K, a, grids[gi], s);
        }
        if (rc) return rc;
    }
    if (ev) {
        hipEvent_t b;
        HIP_TRY(hipEventCreate(&b));
        ev->push_back(b);
        HIP_TRY(hipEventRecord(b, s));
    }
    // A single 1-D or image problem: the pass ran the step in its last workgroup, and its tiles
    // never need compacting (the slots of a stopped instance are skipped), so nothing else is
    // launched.
    if (fused) {
        HIP_TRY(hipGetLastError());
        return PASC_OK;
    }
    // A single graph problem: the step kernel counts the instance out when it stops, and the
    // graph kernels skip the tiles of a stopped instance, so no scan is launched.
    const bool single = e.n_inst == 1;
    hipLaunchKernelGGL(spg_step, dim3(std::min(e.n_inst, std::max(e.ctx->n_cu, 1) * 4)), dim3(256),
                       0, s, e.d_desc, e.d_state, e.d_part, e.n_inst, ph.sp, e.d_nactive + 1,
                       single ? e.d_nactive : nullptr);
    if (single) {
        HIP_TRY(hipGetLastError());
        return PASC_OK;
    }
    ScanArgs sa{};
    sa.n_groups = (int)e.groups.size();
    for (size_t gi = 0; gi < e.groups.size(); gi++)
        sa.g[gi] = ScanGroup{e.groups[gi].d_ids, e.groups[gi].d_off, e.groups[gi].d_hint, e.groups[gi].n};
    sa.desc = e.d_desc;
    sa.state = e.d_state;
    sa.nactive = e.d_nactive;
    sa.dirty = e.d_nactive + 1;
    hipLaunchKernelGGL(spg_scan, dim3(1), dim3(1024), 0, s, sa);
    HIP_TRY(hipGetLastError());
    return PASC_OK;
}

// Owners of the graph, its executable and the profiling events of a phase: released on every
// exit path, error returns included.
struct GraphGuard {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~GraphGuard() { reset(); }
};

struct EventsGuard {
    std::vector<hipEvent_t> ev;
    ~EventsGuard() {
        for (auto x : ev) (void)hipEventDestroy(x);
    }
};

// The image pass reads the data term q (constant while theta is) from rows laid out like
// gamma, formed here at the start of every phase in the phase's type.
template <typename S>
static int img_data_terms(Engine &e, const void *x, S *q) {
    int64_t tmax = 1;
    for (int i = 0; i < e.n_inst; i++) tmax = std::max<int64_t>(tmax, e.h_desc[i].T);
    const dim3 grid((unsigned)std::min<int64_t>((tmax + 255) / 256, 4096),
                    (unsigned)std::min(std::max(e.n_inst, 1), 65535));
    for (auto &g : e.groups) {
        switch (g.K) {
#define PASC_QK(KK) \
    case KK: hipLaunchKernelGGL((img_data_term<S, KK>), grid, dim3(256), 0, S_(e), e.d_desc, e.n_inst, x, q); break;
            PASC_QK(1) PASC_QK(2) PASC_QK(3) PASC_QK(4) PASC_QK(5) PASC_QK(6) PASC_QK(7) PASC_QK(8)
#undef PASC_QK
            default: return fail(PASC_ERR_ARG, "unsupported K %d", g.K);
        }
    }
    HIP_TRY(hipGetLastError());
    return PASC_OK;
}

static int run_phase(Engine &e, const PhaseCfg &ph, const pasc_spg_params &p, int max_passes) {
    hipStream_t s = S_(e);
    if (e.cur_phase != (ph.f64 ? 1 : 0) || tile_len(e, ph.f64) != e.cur_tile) {
        int rc0 = engine_retile(e, ph.f64);
        if (rc0) return rc0;
    }
    if (engine_is_image(e)) {
        const int b = ph.f64 ? 1 : 0;
        if (!e.d_q[b]) HIP_TRY(hipMalloc(&e.d_q[b], (ph.f64 ? 8 : 4) * (size_t)e.state_elems));
        const void *x = ph.f64 ? (const void *)e.d_x64 : (const void *)e.d_x32;
        const int rc0 = ph.f64 ? img_data_terms<double>(e, x, (double *)e.d_q[1])
                               : img_data_terms<float>(e, x, (float *)e.d_q[0]);
        if (rc0) return rc0;
    }
    // initial scan (active tiles of every group)
    {
        ScanArgs sa{};
        sa.n_groups = (int)e.groups.size();
        for (size_t gi = 0; gi < e.groups.size(); gi++)
            sa.g[gi] = ScanGroup{e.groups[gi].d_ids, e.groups[gi].d_off, e.groups[gi].d_hint, e.groups[gi].n};
        sa.desc = e.d_desc;
        sa.state = e.d_state;
        sa.nactive = e.d_nactive;
        sa.dirty = e.d_nactive + 1;
        // force the initial rebuild (any non-zero dirty flag)
        HIP_TRY(hipMemsetAsync(e.d_nactive + 1, 1, sizeof(int32_t), s));
        hipLaunchKernelGGL(spg_scan, dim3(1), dim3(1024), 0, s, sa);
        HIP_TRY(hipGetLastError());
    }
    std::vector<int> grids;
    for (auto &g : e.groups) {
        int64_t cap = e.grid_cap > 0 ? e.grid_cap : (int64_t)std::max(e.ctx->n_cu, 1) * 64;
        grids.push_back((int)std::max<int64_t>(1, std::min<int64_t>(g.max_tiles, cap)));
    }
    const int every = std::max(1, p.check_every);
    std::vector<int32_t *> d_totals;
    for (auto &g : e.groups) d_totals.push_back(g.d_off + g.n);
    GraphGuard gg;
    EventsGuard eg;
    std::vector<hipEvent_t> &evs = eg.ev;
    int passes = 0;
    int rc = PASC_OK;
    while (passes < max_passes) {
        const int n = std::min(every, max_passes - passes);
        if (p.use_graph && n == every && !e.profile) {
            if (!gg.exec) {
                HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < n && rc == PASC_OK; i++) rc = enqueue_iteration(e, ph, grids);
                hipError_t ec = hipStreamEndCapture(s, &gg.graph);
                if (rc) return rc;
                if (ec != hipSuccess) return fail(PASC_ERR_HIP, "graph capture: %s", hipGetErrorString(ec));
                ec = hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0);
                if (ec != hipSuccess) return fail(PASC_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ec));
            }
            HIP_TRY(hipGraphLaunch(gg.exec, s));
        } else {
            for (int i = 0; i < n; i++) {
                rc = enqueue_iteration(e, ph, grids, e.profile ? &evs : nullptr);
                if (rc) return rc;
            }
        }
        passes += n;
        HIP_TRY(hipMemcpyAsync(e.h_nactive, e.d_nactive, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        for (size_t gi = 0; gi < d_totals.size(); gi++)
            HIP_TRY(hipMemcpyAsync(e.h_nactive + 1 + gi, d_totals[gi], sizeof(int32_t),
                                   hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (*e.h_nactive == 0) break;
        // shrink the persistent grids once most tiles have converged (idle workgroups still cost
        // a dispatch each); the graph is re-captured with the new launch shapes
        bool shrink = false;
        for (size_t gi = 0; gi < grids.size(); gi++) {
            const int need = std::max(1, (int)e.h_nactive[1 + gi]);
            if (grids[gi] > 2 * need && grids[gi] > 64) {
                grids[gi] = std::max(need, 64);
                shrink = true;
            }
        }
        if (shrink) gg.reset();
    }
    gg.reset();
    {
        // the passes stored K - 1 rows: write the last one of every current iterate
        int64_t tmax = 1;
        for (int i = 0; i < e.n_inst; i++) tmax = std::max<int64_t>(tmax, e.h_desc[i].T);
        const dim3 grid((unsigned)std::min<int64_t>((tmax + 255) / 256, 2048),
                        (unsigned)std::min(std::max(e.n_inst, 1), 65535));
        const bool g64 = ph.f64 && e.precision == PASC_FP32_BULK;
        if (ph.f64)
            hipLaunchKernelGGL(complete_last_row<double>, grid, dim3(256), 0, s, e.d_desc, e.d_state,
                               e.n_inst, (double *)(g64 ? e.d_gamma64[0] : e.d_gamma[0]),
                               (double *)(g64 ? e.d_gamma64[1] : e.d_gamma[1]));
        else
            hipLaunchKernelGGL(complete_last_row<float>, grid, dim3(256), 0, s, e.d_desc, e.d_state,
                               e.n_inst, (float *)e.d_gamma[0], (float *)e.d_gamma[1]);
        HIP_TRY(hipGetLastError());
    }
    if (!evs.empty()) {
        HIP_TRY(hipStreamSynchronize(s));
        // Only the passes that ran work count: it + 1 of the longest-running instance.  The
        // launches after the last instance stopped within a poll interval find no active tile
        // (a few microseconds each) and would pull the average down.
        std::vector<InstState> hs(e.n_inst);
        HIP_TRY(hipMemcpy(hs.data(), e.d_state, sizeof(InstState) * e.n_inst,
                          hipMemcpyDeviceToHost));
        const int launched = (int)(evs.size() / 2);
        int real = 1;
        for (const auto &x : hs) real = std::max(real, x.it + 1);
        real = std::min(real, launched);
        double ms = 0.0;
        for (int i = 0; i < real; i++) {
            float m = 0.f;
            HIP_TRY(hipEventElapsedTime(&m, evs[2 * i], evs[2 * i + 1]));
            ms += m;
        }
        e.prof_ms[ph.f64 ? 1 : 0] += ms;
        e.prof_n[ph.f64 ? 1 : 0] += real;
    }
    return PASC_OK;
}

static void init_states(Engine &e, std::vector<InstState> &st, const pasc_spg_params &p,
                        const std::vector<char> &mask, bool keep_alpha) {
    for (int i = 0; i < e.n_inst; i++) {
        InstState &s = st[i];
        if (!mask[i]) {
            s.active = 0;
            s.first = 1;  // runs no pass: its rows are left as they are
            continue;
        }
        const double a = keep_alpha ? s.alpha_dir : p.alphainit;
        const double nb = s.normb;
        std::memset(&s, 0, sizeof(s));
        s.normb = nb;
        s.alpha_dir = (a > 0 && std::isfinite(a)) ? a : p.alphainit;
        s.alpha_apply = s.alpha_dir;
        s.best_kkt = std::numeric_limits<double>::infinity();
        s.kkt = std::numeric_limits<double>::infinity();
        s.first = 1;
        s.active = 1;
        s.parity = e.h_parity[i];
        s.f = std::numeric_limits<double>::max();
    }
}

static StepParams make_step(const pasc_spg_params &p, bool fp32_phase) {
    StepParams sp{};
    sp.maxit = p.maxit;
    sp.m = std::min(std::max(p.m, 1), PASC_MAX_M);
    sp.eps = p.eps;
    sp.gamma = p.gamma;
    sp.sigma1 = p.sigma1;
    sp.sigma2 = p.sigma2;
    sp.alphainit = p.alphainit;
    sp.stop_mask = p.stop_mask;
    sp.kkt_tol = p.kkt_tol;
    sp.fp32_phase = fp32_phase ? 1 : 0;
    sp.stall_window = std::max(1, p.stall_window);
    // hand over to FP64 once the FP32 residual is within 100x of the target (the FP32 floor of
    // the residual is ~1e-5 at epsilon 100), or on stagnation
    sp.fp32_switch = p.fp32_switch > 0.0 ? p.fp32_switch : 100.0 * p.kkt_tol;
    sp.fx_identity = fp32_phase ? 1 : 0;
    return sp;
}

static int compute_normb(Engine &e, std::vector<InstState> &st) {
    // ||b||_2 of every instance (only needed by the *_normb criteria)
    for (int i = 0; i < e.n_inst; i++) {
        double f, kkt, nb2;
        int rc = engine_eval_raw(e, i, &f, &kkt, &nb2);
        if (rc) return rc;
        st[i].normb = std::sqrt(nb2);
    }
    return PASC_OK;
}

int engine_solve(Engine &e, const pasc_spg_params &p, pasc_spg_result *res_single,
                 pasc_instance_result *res_multi, double *time_s) {
    hipStream_t s = S_(e);
    std::vector<InstState> st(e.n_inst);
    std::memset(st.data(), 0, sizeof(InstState) * st.size());
    if (e.cur64 && e.precision == PASC_FP32_BULK) {
        // warm start (x0 aliases x, graphh1fem.cpp:136-137): the last iterate is FP64
        for (int i = 0; i < e.n_inst; i++) {
            const InstDesc &d = e.h_desc[i];
            const int64_t n = (int64_t)d.K * d.ld;
            hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned)((n + 255) / 256)),
                               dim3(256), 0, s, (const double *)e.d_gamma64[e.h_parity[i]] + d.goff,
                               (float *)e.d_gamma[0] + d.goff, n);
            e.h_parity[i] = 0;
        }
        HIP_TRY(hipGetLastError());
    }
    e.cur64 = false;
    e.prof_ms[0] = e.prof_ms[1] = 0.0;
    e.prof_n[0] = e.prof_n[1] = 0;
    if (p.stop_mask & (PASC_STOP_NORMGP_NORMB | PASC_STOP_ANORMGP_NORMB)) {
        int rc = compute_normb(e, st);
        if (rc) return rc;
    }
    std::vector<char> all(e.n_inst, 1);
    init_states(e, st, p, all, false);
    HIP_TRY(hipMemcpyAsync(e.d_state, st.data(), sizeof(InstState) * e.n_inst,
                           hipMemcpyHostToDevice, s));
    EventsGuard eg;
    eg.ev.assign(2, nullptr);
    HIP_TRY(hipEventCreate(&eg.ev[0]));
    HIP_TRY(hipEventCreate(&eg.ev[1]));
    hipEvent_t ev0 = eg.ev[0], ev1 = eg.ev[1];
    HIP_TRY(hipEventRecord(ev0, s));

    const bool bulk32 = e.precision == PASC_FP32_BULK;
    std::vector<int> it32(e.n_inst, 0), it64(e.n_inst, 0);
    int rc;
    if (!bulk32) {
        PhaseCfg ph{true, make_step(p, false), p.kkt_every};
        rc = run_phase(e, ph, p, p.maxit + 2);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(st.data(), e.d_state, sizeof(InstState) * e.n_inst,
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int i = 0; i < e.n_inst; i++) { it64[i] = st[i].it; e.h_parity[i] = st[i].parity; }
    } else {
        PhaseCfg ph{false, make_step(p, true), p.kkt_every};
        rc = run_phase(e, ph, p, p.maxit + 2);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(st.data(), e.d_state, sizeof(InstState) * e.n_inst,
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        // FP64 polish: promote the FP32 iterate and certify the target in FP64
        if (!e.d_gamma64[0]) {
            for (int b = 0; b < 2; b++) {
                HIP_TRY(hipMalloc(&e.d_gamma64[b], 8 * (size_t)e.state_elems));
                HIP_TRY(hipMemsetAsync(e.d_gamma64[b], 0, 8 * (size_t)e.state_elems, s));
                if (e.stencil == STENCIL_GRID) {
                    HIP_TRY(hipMalloc(&e.d_dir64[b], 8 * (size_t)e.state_elems));
                    HIP_TRY(hipMemsetAsync(e.d_dir64[b], 0, 8 * (size_t)e.state_elems, s));
                }
            }
        }
        std::vector<char> pol(e.n_inst, 0);
        for (int i = 0; i < e.n_inst; i++) {
            it32[i] = st[i].it;
            e.h_parity[i] = st[i].parity;
            pol[i] = st[i].status == PASC_OK;
            const InstDesc &d = e.h_desc[i];
            const int64_t n = (int64_t)d.K * d.ld;
            hipLaunchKernelGGL((convert_kernel<float, double>), dim3((unsigned)((n + 255) / 256)),
                               dim3(256), 0, s, (const float *)e.d_gamma[st[i].parity] + d.goff,
                               (double *)e.d_gamma64[0] + d.goff, n);
            e.h_parity[i] = 0;
        }
        HIP_TRY(hipGetLastError());
        init_states(e, st, p, pol, true);
        // polish_maxit -1: the polish gets what the FP32 phase left of maxit
        const bool shared_cap = p.polish_maxit < 0;
        for (int i = 0; i < e.n_inst; i++) st[i].it_base = shared_cap ? it32[i] : 0;
        HIP_TRY(hipMemcpyAsync(e.d_state, st.data(), sizeof(InstState) * e.n_inst,
                               hipMemcpyHostToDevice, s));
        pasc_spg_params pp = p;
        pp.maxit = p.polish_maxit > 0 ? p.polish_maxit : p.maxit;
        PhaseCfg ph2{true, make_step(pp, false), p.kkt_every};
        rc = run_phase(e, ph2, pp, pp.maxit + 2);
        if (rc) return rc;
        e.cur64 = true;
        HIP_TRY(hipMemcpyAsync(st.data(), e.d_state, sizeof(InstState) * e.n_inst,
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int i = 0; i < e.n_inst; i++) {
            it64[i] = pol[i] ? st[i].it : 0;
            e.h_parity[i] = st[i].parity;
        }
    }
    HIP_TRY(hipEventRecord(ev1, s));
    HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    if (time_s) *time_s = ms * 1e-3;
    int worst = PASC_OK;
    for (int i = 0; i < e.n_inst; i++) {
        const InstState &x = st[i];
        if (x.status) worst = x.status;
        if (res_multi) {
            pasc_instance_result &r = res_multi[i];
            r.it = it32[i] + it64[i];
            r.it_fp32 = it32[i];
            r.it_fp64 = it64[i];
            r.stop_reason = x.reason;
            r.status = x.status;
            r.fx = x.f;
            r.kkt = x.kkt;
        }
    }
    if (res_single && e.n_inst >= 1) {
        const InstState &x = st[0];
        pasc_spg_result &r = *res_single;
        r.it = it32[0] + it64[0];
        r.hessmult = r.it + 1;
        r.status = x.status;
        r.stop_reason = x.reason;
        r.it_fp32 = it32[0];
        r.it_fp64 = it64[0];
        r.fx = x.f;
        r.dd = x.dd;
        r.dAd = x.dAd;
        r.kkt = x.kkt;
        r.normb = x.normb;
        r.time_s = ms * 1e-3;
        r.pass_time_fp32_s = e.prof_ms[0] * 1e-3;
        r.pass_time_fp64_s = e.prof_ms[1] * 1e-3;
        r.passes_fp32 = e.prof_n[0];
        r.passes_fp64 = e.prof_n[1];
    }
    if (worst == PASC_ERR_NUMERIC) return fail(PASC_ERR_NUMERIC, "NaN/Inf in the SPG iteration");
    return PASC_OK;
}

// ------------------------------------------------------------------------------------------
// auxiliary computations on one instance (current iterate)
// ------------------------------------------------------------------------------------------
static const void *cur_gamma(Engine &e, int inst, bool &f32) {
    const int par = e.h_parity[inst];
    if (e.cur64) { f32 = false; return e.d_gamma64[par]; }
    f32 = e.precision == PASC_FP32_BULK;
    return e.d_gamma[par];
}

template <int K>
static void launch_eval(Engine &e, const InstDesc &d, const void *g, bool f32, double *part,
                        int blocks) {
    hipStream_t s = S_(e);
    // the 1-D time graph of this instance, or the engine's graph
    const GraphGeo G = e.stencil == STENCIL_TIME ? GraphGeo{d.T, 1, 0, nullptr, nullptr} : e.geo;
    if (f32)
        hipLaunchKernelGGL((eval_kernel<float, K>), dim3(blocks), dim3(256), 0, s,
                           (const float *)g + d.goff, (const double *)e.d_x64, d, G, part);
    else
        hipLaunchKernelGGL((eval_kernel<double, K>), dim3(blocks), dim3(256), 0, s,
                           (const double *)g + d.goff, (const double *)e.d_x64, d, G, part);
}

int engine_eval_raw(Engine &e, int inst, double *fx, double *kkt, double *nb2) {
    const InstDesc &d = e.h_desc[inst];
    bool f32;
    const void *g = cur_gamma(e, inst, f32);
    const int blocks = (int)((d.T + 255) / 256);
    void *part;
    int rc = scratch(e, sizeof(double) * 3 * (size_t)blocks, &part);
    if (rc) return rc;
    switch (d.K) {
        case 1: launch_eval<1>(e, d, g, f32, (double *)part, blocks); break;
        case 2: launch_eval<2>(e, d, g, f32, (double *)part, blocks); break;
        case 3: launch_eval<3>(e, d, g, f32, (double *)part, blocks); break;
        case 4: launch_eval<4>(e, d, g, f32, (double *)part, blocks); break;
        case 5: launch_eval<5>(e, d, g, f32, (double *)part, blocks); break;
        case 6: launch_eval<6>(e, d, g, f32, (double *)part, blocks); break;
        case 7: launch_eval<7>(e, d, g, f32, (double *)part, blocks); break;
        case 8: launch_eval<8>(e, d, g, f32, (double *)part, blocks); break;
        default: return fail(PASC_ERR_ARG, "unsupported K");
    }
    HIP_TRY(hipGetLastError());
    std::vector<double> h(3 * (size_t)blocks);
    HIP_TRY(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, S_(e)));
    HIP_TRY(hipStreamSynchronize(S_(e)));
    double f = 0, n2 = 0, k = 0;
    for (int b = 0; b < blocks; b++) {
        f += h[3 * b];
        n2 += h[3 * b + 1];
        const double m = h[3 * b + 2];
        if (m > k || m != m) k = m;
    }
    if (fx) *fx = f;
    if (kkt) *kkt = k;
    if (nb2) *nb2 = n2;
    return PASC_OK;
}

int engine_eval(Engine &e, int inst, double *fx, double *kkt) {
    return engine_eval_raw(e, inst, fx, kkt, nullptr);
}

int rows_theta(Engine &e, const void *g, bool f32, const InstDesc &d, const double *x,
               double *theta_out) {
    const int blocks = (int)((d.T + 255) / 256);
    const int nv = d.K * (1 + d.xdim);
    void *part;
    int rc = scratch(e, sizeof(double) * (size_t)nv * blocks, &part);
    if (rc) return rc;
    if (f32)
        hipLaunchKernelGGL((theta_partials<float>), dim3(blocks), dim3(256), 0, S_(e),
                           (const float *)g, x, d, (double *)part);
    else
        hipLaunchKernelGGL((theta_partials<double>), dim3(blocks), dim3(256), 0, S_(e),
                           (const double *)g, x, d, (double *)part);
    HIP_TRY(hipGetLastError());
    std::vector<double> h((size_t)nv * blocks);
    HIP_TRY(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, S_(e)));
    HIP_TRY(hipStreamSynchronize(S_(e)));
    std::vector<double> tot(nv, 0.0);
    for (int b = 0; b < blocks; b++)
        for (int v = 0; v < nv; v++) tot[v] += h[(size_t)b * nv + v];
    // graphh1fem.cpp:374-379: theta_k = <x, gamma_k> / sum(gamma_k), 0 if the sum is 0
    for (int k = 0; k < d.K; k++)
        for (int c = 0; c < d.xdim; c++) {
            const double gs = tot[k];
            theta_out[k * d.xdim + c] = gs != 0 ? tot[d.K + k * d.xdim + c] / gs : 0.0;
        }
    return PASC_OK;
}

int engine_theta_update(Engine &e, int inst, double *theta_out) {
    InstDesc &d = e.h_desc[inst];
    bool f32;
    const void *g = cur_gamma(e, inst, f32);
    double th[PASC_MAX_K * 3];
    const size_t es = f32 ? 4 : 8;
    int rc = rows_theta(e, (const char *)g + es * d.goff, f32, d, (const double *)e.d_x64, th);
    if (rc) return rc;
    for (int i = 0; i < d.K * d.xdim; i++) {
        d.theta[i] = th[i];
        if (theta_out) theta_out[i] = th[i];
    }
    return engine_update_desc(e);
}

int rows_labels(Engine &e, const void *g, bool f32, const InstDesc &d, int32_t *labels) {
    const int blocks = (int)((d.T + 255) / 256);
    void *buf;
    int rc = scratch(e, sizeof(int32_t) * (size_t)d.T, &buf);
    if (rc) return rc;
    if (f32)
        hipLaunchKernelGGL((labels_kernel<float>), dim3(blocks), dim3(256), 0, S_(e),
                           (const float *)g, d, (int32_t *)buf);
    else
        hipLaunchKernelGGL((labels_kernel<double>), dim3(blocks), dim3(256), 0, S_(e),
                           (const double *)g, d, (int32_t *)buf);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(labels, buf, sizeof(int32_t) * d.T, hipMemcpyDeviceToHost, S_(e)));
    HIP_TRY(hipStreamSynchronize(S_(e)));
    return PASC_OK;
}

int engine_labels(Engine &e, int inst, int32_t *labels) {
    const InstDesc &d = e.h_desc[inst];
    bool f32;
    const void *g = cur_gamma(e, inst, f32);
    return rows_labels(e, (const char *)g + (f32 ? 4 : 8) * d.goff, f32, d, labels);
}

int rows_abserr(Engine &e, const void *g, bool f32, const InstDesc &d, const double *solution,
                double *abserr) {
    const int blocks = (int)((d.T + 255) / 256);
    void *buf;
    const size_t need = sizeof(double) * ((size_t)d.T + blocks);
    int rc = scratch(e, need, &buf);
    if (rc) return rc;
    double *sol = (double *)buf;
    double *part = sol + d.T;
    HIP_TRY(hipMemcpyAsync(sol, solution, sizeof(double) * d.T, hipMemcpyHostToDevice, S_(e)));
    if (f32)
        hipLaunchKernelGGL((abserr_kernel<float>), dim3(blocks), dim3(256), 0, S_(e),
                           (const float *)g, d, (const double *)sol, part);
    else
        hipLaunchKernelGGL((abserr_kernel<double>), dim3(blocks), dim3(256), 0, S_(e),
                           (const double *)g, d, (const double *)sol, part);
    HIP_TRY(hipGetLastError());
    std::vector<double> h(blocks);
    HIP_TRY(hipMemcpyAsync(h.data(), part, sizeof(double) * blocks, hipMemcpyDeviceToHost, S_(e)));
    HIP_TRY(hipStreamSynchronize(S_(e)));
    double a = 0;
    for (double v : h) a += v;
    *abserr = a;
    return PASC_OK;
}

int engine_abserr(Engine &e, int inst, const double *solution, double *abserr) {
    const InstDesc &d = e.h_desc[inst];
    bool f32;
    const void *g = cur_gamma(e, inst, f32);
    return rows_abserr(e, (const char *)g + (f32 ? 4 : 8) * d.goff, f32, d, solution, abserr);
}

const void *engine_current_gamma(Engine &e, int inst, bool *f32) {
    bool f;
    const void *g = cur_gamma(e, inst, f);
    if (f32) *f32 = f;
    return (const char *)g + (f ? 4 : 8) * e.h_desc[inst].goff;
}

int engine_set_gamma_rows(Engine &e, int inst, const double *rows_dev) {
    const InstDesc &d = e.h_desc[inst];
    const int64_t n = (int64_t)d.K * d.ld;
    e.cur64 = false;
    e.h_parity[inst] = 0;
    if (e.precision == PASC_FP32_BULK) {
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned)((n + 255) / 256)),
                           dim3(256), 0, S_(e), rows_dev, (float *)e.d_gamma[0] + d.goff, n);
        HIP_TRY(hipGetLastError());
    } else {
        HIP_TRY(hipMemcpyAsync((double *)e.d_gamma[0] + d.goff, rows_dev, 8 * (size_t)n,
                               hipMemcpyDeviceToDevice, S_(e)));
    }
    return PASC_OK;
}

static int alloc_linear_term(Engine &e) {
    if (e.d_b64) return PASC_OK;
    HIP_TRY(hipMalloc(&e.d_b64, 8 * (size_t)e.state_elems));
    HIP_TRY(hipMemsetAsync(e.d_b64, 0, 8 * (size_t)e.state_elems, S_(e)));
    if (e.precision == PASC_FP32_BULK) {
        HIP_TRY(hipMalloc(&e.d_b32, 4 * (size_t)e.state_elems));
        HIP_TRY(hipMemsetAsync(e.d_b32, 0, 4 * (size_t)e.state_elems, S_(e)));
    }
    return PASC_OK;
}

int engine_linear_term_rows(Engine &e, int inst, double **rows) {
    int rc = alloc_linear_term(e);
    if (rc) return rc;
    *rows = (double *)e.d_b64 + e.h_desc[inst].goff;
    return PASC_OK;
}

int engine_commit_linear_term(Engine &e, int inst) {
    InstDesc &d = e.h_desc[inst];
    const int64_t n = (int64_t)d.K * d.ld;
    double *b64 = (double *)e.d_b64 + d.goff;
    float *b32 = nullptr;
    if (e.d_b32) {
        b32 = (float *)e.d_b32 + d.goff;
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned)((n + 255) / 256)),
                           dim3(256), 0, S_(e), (const double *)b64, b32, n);
        HIP_TRY(hipGetLastError());
    }
    d.b64 = b64;
    d.b32 = b32;
    return engine_update_desc(e);
}

template <typename S>
__global__ __launch_bounds__(256) void fill_uniform_kernel(const InstDesc *desc, S *g) {
    const InstDesc &d = desc[blockIdx.y];
    const int64_t n = (int64_t)d.K * d.ld;
    const S v = S(1.0 / d.K);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = i % d.ld;
        g[d.goff + i] = t < d.T ? v : S(0);
    }
}

int engine_fill_uniform(Engine &e) {
    e.cur64 = false;
    for (int i = 0; i < e.n_inst; i++) e.h_parity[i] = 0;
    const dim3 grid(64, (unsigned)e.n_inst);
    if (e.precision == PASC_FP32_BULK)
        hipLaunchKernelGGL((fill_uniform_kernel<float>), grid, dim3(256), 0, S_(e), e.d_desc, (float *)e.d_gamma[0]);
    else
        hipLaunchKernelGGL((fill_uniform_kernel<double>), grid, dim3(256), 0, S_(e), e.d_desc, (double *)e.d_gamma[0]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(S_(e)));
    return PASC_OK;
}

static int check_shape(const PassShape &s, bool f32) {
    const bool ok_chunk = f32 ? (s.chunk == 4 || s.chunk == 8) : (s.chunk == 2 || s.chunk == 4);
    if (!(s.block == 64 || s.block == 128 || s.block == 256 || s.block == 512) || !ok_chunk ||
        (s.prefetch != 0 && s.prefetch != 1))
        return fail(PASC_ERR_ARG, "no fused %s pass for block %d chunk %d prefetch %d",
                    f32 ? "FP32" : "FP64", s.block, s.chunk, s.prefetch);
    if (s.max_slots < 1 || s.min_tiles < 1) return fail(PASC_ERR_ARG, "max_slots and min_tiles must be >= 1");
    return PASC_OK;
}

// image pass: one-wave tiles of 1..256 image rows, 1..3 rows loaded ahead (prefetch); the
// other fields are not used
static int check_img_shape(const PassShape &s) {
    if (s.block != WAVE || s.chunk < 1 || s.chunk > 256 || s.prefetch < 1 || s.prefetch > 3)
        return fail(PASC_ERR_ARG, "image pass: block must be %d, chunk (image rows per tile) "
                    "1..256 and prefetch 1..3, got block %d chunk %d prefetch %d", WAVE, s.block,
                    s.chunk, s.prefetch);
    return PASC_OK;
}

int engine_set_shape(Engine &e, const PassShape &s32_in, const PassShape &s64_in) {
    const bool f32 = e.precision == PASC_FP32_BULK;
    const bool img = engine_is_image(e);
    PassShape s32 = s32_in, s64 = s64_in;
    if (img) {
        // images: prefetch 0 asks for one row ahead
        if (s32.prefetch == 0) s32.prefetch = 1;
        if (s64.prefetch == 0) s64.prefetch = 1;
    }
    if (!img && e.stencil != STENCIL_TIME)
        return fail(PASC_ERR_ARG, "set_launch_shape: graphs other than images have a fixed shape");
    int rc = !f32 ? PASC_OK : img ? check_img_shape(s32) : check_shape(s32, true);
    if (rc == PASC_OK) rc = img ? check_img_shape(s64) : check_shape(s64, false);
    if (rc) return rc;
    std::vector<std::vector<double>> saved(e.n_inst);
    for (int i = 0; i < e.n_inst; i++) {
        rc = download_rows(e, i, saved[i]);
        if (rc) return rc;
    }
    if (f32) e.sh[0] = s32;
    e.sh[1] = s64;
    rc = engine_retile(e, !f32);
    for (int i = 0; i < e.n_inst && rc == PASC_OK; i++) rc = upload_rows(e, i, saved[i]);
    return rc;
}

// Self-written autotuner: times the fused pass of this problem (or batch) for every compiled
// (threads per block, time steps per thread, prefetch) triple, and with prefetch also for
// 8192 / 4096 / 2048 / 1024 / 512 partial slots per instance (more slots put more one-wave
// workgroups on each SIMD) and for at least 4 tiles per slot (more tiles per workgroup lengthen
// its pipeline) -- 20 passes each from the current iterate with the KKT
// residual on every 4th pass, hipEvents around every launch.  An FP32-bulk problem runs 20 FP32
// and 20 FP64 polish passes per candidate, and the fastest shape of each phase is kept (the
// preferred shape when it is within 5% of the fastest, see below).  The iterate is restored
// afterwards.
// Images: times 20 passes of the image pass (per phase) for the balanced rows
// (img_balanced_rows) and 6, 8, 12, 16 and 24 image rows per wave tile and keeps the fastest, or
// the balanced rows when within 5% of it (the rows fix the order of the sums, so a choice that
// flips between near-equal shapes would change the iterates).
static int img_autotune(Engine &e, const std::vector<std::vector<double>> &saved,
                        PassShape *s32, PassShape *s64, double us[2]) {
    const bool f32 = e.precision == PASC_FP32_BULK;
    pasc_spg_params p;
    pasc_spg_default_params(&p);
    p.maxit = 20;
    p.stop_mask = 0;
    p.polish_maxit = f32 ? 20 : 1;
    p.kkt_every = 4;
    const bool prof = e.profile;
    e.profile = true;
    const int PREF = img_balanced_rows(e);
    std::vector<int> rows = {PREF};
    for (int r : {6, 8, 12, 16, 24})
        if (r != PREF) rows.push_back(r);
    double best[2] = {1e300, 1e300}, pref_t[2] = {1e300, 1e300};
    int best_rows[2] = {PREF, PREF};
    for (size_t ri = 0; ri < rows.size(); ri++) {
        e.sh[0] = e.sh[1] = PassShape{WAVE, rows[ri], 1, 0, 1};
        int rc = engine_retile(e, !f32);
        for (int i = 0; i < e.n_inst && rc == PASC_OK; i++) rc = upload_rows(e, i, saved[i]);
        pasc_spg_result r;
        if (rc == PASC_OK) rc = engine_solve(e, p, &r, nullptr, nullptr);
        if (rc) {
            e.profile = prof;
            return rc;
        }
        for (int ph = 0; ph < 2; ph++) {
            const double t = e.prof_n[ph] ? e.prof_ms[ph] / e.prof_n[ph] : 1e300;
            if (t < best[ph]) { best[ph] = t; best_rows[ph] = rows[ri]; }
            if (rows[ri] == PREF) pref_t[ph] = t;
        }
    }
    for (int ph = 0; ph < 2; ph++)
        if (pref_t[ph] <= 1.05 * best[ph]) { best[ph] = pref_t[ph]; best_rows[ph] = PREF; }
    e.profile = prof;
    e.sh[0] = PassShape{WAVE, best_rows[0], 1, 0, 1};
    e.sh[1] = PassShape{WAVE, best_rows[1], 1, 0, 1};
    int rc = engine_retile(e, !f32);
    for (int i = 0; i < e.n_inst && rc == PASC_OK; i++) rc = upload_rows(e, i, saved[i]);
    if (rc) return rc;
    *s32 = e.sh[0];
    *s64 = e.sh[1];
    us[0] = f32 ? best[0] * 1e3 : 0.0;
    us[1] = best[1] * 1e3;
    return PASC_OK;
}

int engine_autotune(Engine &e, PassShape *s32, PassShape *s64, double us[2]) {
    if (e.stencil != STENCIL_TIME && !engine_is_image(e))
        return fail(PASC_ERR_ARG, "autotune: 1-D time graphs and images only");
    std::vector<std::vector<double>> saved(e.n_inst);
    for (int i = 0; i < e.n_inst; i++) {
        int rc = download_rows(e, i, saved[i]);
        if (rc) return rc;
    }
    if (engine_is_image(e)) return img_autotune(e, saved, s32, s64, us);
    const bool f32 = e.precision == PASC_FP32_BULK;
    const int chunks32[2] = {4, 8}, chunks64[2] = {2, 4};
    const int blocks[4] = {64, 128, 256, 512};
    constexpr int NSLOT = 6;
    const int slot_opts[NSLOT][2] = {{2048, 1}, {1024, 1}, {512, 1}, {2048, 4}, {4096, 1}, {8192, 1}};
    pasc_spg_params p;
    pasc_spg_default_params(&p);
    p.maxit = 20;
    p.stop_mask = 0;
    p.polish_maxit = f32 ? 20 : 1;
    p.kkt_every = 4;
    const bool prof = e.profile;
    e.profile = true;
    double best[2] = {1e300, 1e300};
    PassShape bs[2] = {e.sh[0], e.sh[1]};
    // The preferred shape of each phase: one-wave workgroups with the longer chunk, staged,
    // 4096 (FP32) / 8192 (FP64) slots -- the fastest for a long series in most sessions.  It is
    // kept whenever it is within TUNE_TIE of the fastest candidate: every shape sums the
    // partials in its own order, so a choice that flips between near-equal shapes from run to
    // run would change the iterates (at rounding level) and with them the iteration count.
    // (5%: on a slower box the 20-pass timings of near-equal shapes differ by 3-4% from run to
    // run, and a 3% band let the choice flip there)
    constexpr double TUNE_TIE = 1.05;
    const PassShape pref32{64, 8, 1, 4096, 1}, pref64{64, 4, 1, 8192, 1};
    double pref_t[2] = {1e300, 1e300};
    auto same = [](const PassShape &a, const PassShape &b) {
        return a.block == b.block && a.chunk == b.chunk && a.prefetch == b.prefetch &&
               a.max_slots == b.max_slots && a.min_tiles == b.min_tiles;
    };
    for (int pi = 0; pi < 2; pi++)
    for (int si = 0; si < (pi ? NSLOT : 1); si++)
    for (int ci = 0; ci < 2; ci++)
    for (int bi = 0; bi < 4; bi++) {
        const PassShape c32{blocks[bi], chunks32[ci], pi, slot_opts[si][0], slot_opts[si][1]};
        const PassShape c64{blocks[bi], chunks64[ci], pi, slot_opts[si][0], slot_opts[si][1]};
        e.sh[0] = c32;
        e.sh[1] = c64;
        int rc = engine_retile(e, !f32);
        for (int i = 0; i < e.n_inst && rc == PASC_OK; i++) rc = upload_rows(e, i, saved[i]);
        pasc_spg_result r;
        if (rc == PASC_OK) rc = engine_solve(e, p, &r, nullptr, nullptr);
        if (rc) {
            e.profile = prof;
            return rc;
        }
        const double t32 = e.prof_n[0] ? e.prof_ms[0] / e.prof_n[0] : 1e300;
        const double t64 = e.prof_n[1] ? e.prof_ms[1] / e.prof_n[1] : 1e300;
        if (f32 && t32 < best[0]) { best[0] = t32; bs[0] = c32; }
        if (t64 < best[1]) { best[1] = t64; bs[1] = c64; }
        if (same(c32, pref32)) pref_t[0] = t32;
        if (same(c64, pref64)) pref_t[1] = t64;
    }
    if (f32 && pref_t[0] <= TUNE_TIE * best[0]) { best[0] = pref_t[0]; bs[0] = pref32; }
    if (pref_t[1] <= TUNE_TIE * best[1]) { best[1] = pref_t[1]; bs[1] = pref64; }
    e.profile = prof;
    e.sh[0] = bs[0];
    e.sh[1] = bs[1];
    int rc = engine_retile(e, !f32);
    for (int i = 0; i < e.n_inst && rc == PASC_OK; i++) rc = upload_rows(e, i, saved[i]);
    if (rc) return rc;
    *s32 = bs[0];
    *s64 = bs[1];
    us[0] = f32 ? best[0] * 1e3 : 0.0;
    us[1] = best[1] * 1e3;
    return PASC_OK;
}

__global__ __launch_bounds__(256) void pasc_copy_kernel(const float4 *__restrict__ a,
                                                         float4 *__restrict__ b, int64_t n) {
    // 4 vectors per thread, issued before any store (one block covers 16 KiB)
    const int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    float4 v[4];
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (base + q * 256 < n) v[q] = a[base + q * 256];
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (base + q * 256 < n) b[base + q * 256] = v[q];
}

struct DevBuffers {
    std::vector<void *> p;
    ~DevBuffers() {
        for (auto x : p)
            if (x) (void)hipFree(x);
    }
};

int device_copy_bandwidth(Ctx *ctx, int64_t bytes, int reps, double *gbps) {
    hipStream_t s = (hipStream_t)ctx->stream;
    const int64_t n = bytes / 16;
    DevBuffers buf;
    buf.p.assign(2, nullptr);
    HIP_TRY(hipMalloc(&buf.p[0], 16 * (size_t)n));
    HIP_TRY(hipMalloc(&buf.p[1], 16 * (size_t)n));
    void *a = buf.p[0], *b = buf.p[1];
    HIP_TRY(hipMemsetAsync(a, 0, 16 * (size_t)n, s));
    EventsGuard eg;
    eg.ev.assign(2, nullptr);
    HIP_TRY(hipEventCreate(&eg.ev[0]));
    HIP_TRY(hipEventCreate(&eg.ev[1]));
    hipEvent_t e0 = eg.ev[0], e1 = eg.ev[1];
    const unsigned grid = (unsigned)((n + 1023) / 1024);
    float best = 1e30f;
    for (int r = 0; r < reps; r++) {
        HIP_TRY(hipEventRecord(e0, s));
        hipLaunchKernelGGL(pasc_copy_kernel, dim3(grid), dim3(256), 0, s, (const float4 *)a,
                           (float4 *)b, n);
        HIP_TRY(hipEventRecord(e1, s));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) best = ms;
    }
    *gbps = 2.0 * 16.0 * (double)n / (best * 1e-3) / 1e9;
    return PASC_OK;
}

// ------------------------------------------------------------------------------------------
// kernel-level entry points
// ------------------------------------------------------------------------------------------
template <typename S>
static int project_dispatch(Ctx *ctx, S *x, int64_t nb, int K, int layout, int64_t ld) {
    hipStream_t s = (hipStream_t)ctx->stream;
    if (K <= 8) {
        const unsigned blocks = (unsigned)((nb + 255) / 256);
        switch (K) {
#define PCASE(KK) case KK: hipLaunchKernelGGL((project_thread<S, S, KK>), dim3(blocks), dim3(256), 0, s, x, nb, layout, ld); break;
            PCASE(1) PCASE(2) PCASE(3) PCASE(4) PCASE(5) PCASE(6) PCASE(7) PCASE(8)
#undef PCASE
        }
    } else {
        const unsigned blocks = (unsigned)((nb + 3) / 4);
        hipLaunchKernelGGL((project_wave<S>), dim3(blocks), dim3(256), 0, s, x, nb, K, layout, ld);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return PASC_OK;
}

int kernel_project(Ctx *ctx, void *x, int64_t nblocks, int K, int dtype, int layout, int64_t ld) {
    if (nblocks == 0) return PASC_OK;
    if (dtype == PASC_DTYPE_F32) return project_dispatch<float>(ctx, (float *)x, nblocks, K, layout, ld);
    return project_dispatch<double>(ctx, (double *)x, nblocks, K, layout, ld);
}

int kernel_matvec(Ctx *ctx, const void *x, void *y, int64_t T, int64_t R, int K, int64_t ld,
                  double alpha, int32_t grid_width, const int64_t *nbr_ptr_dev,
                  const int32_t *nbr_idx_dev, int dtype) {
    hipStream_t s = (hipStream_t)ctx->stream;
    const int64_t n = T * R;
    if (n == 0) return PASC_OK;
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)K);
    if (R == 1 && grid_width == 0 && !nbr_ptr_dev) {
        if (dtype == PASC_DTYPE_F32)
            hipLaunchKernelGGL((matvec_time<float>), grid, dim3(256), 0, s, (const float *)x, (float *)y, T, K, ld, alpha);
        else
            hipLaunchKernelGGL((matvec_time<double>), grid, dim3(256), 0, s, (const double *)x, (double *)y, T, K, ld, alpha);
    } else {
        const GraphGeo G{T, R, nbr_ptr_dev ? 0 : grid_width, nbr_ptr_dev, nbr_idx_dev};
        if (dtype == PASC_DTYPE_F32)
            hipLaunchKernelGGL((matvec_graph<float>), grid, dim3(256), 0, s, (const float *)x, (float *)y, G, K, ld, alpha);
        else
            hipLaunchKernelGGL((matvec_graph<double>), grid, dim3(256), 0, s, (const double *)x, (double *)y, G, K, ld, alpha);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return PASC_OK;
}

}  // namespace pasc

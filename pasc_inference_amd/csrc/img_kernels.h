// Fused SPG pass for images: T = 1 step of R = W x H nodes on the implicit 4-neighbour grid
// (blockgraphsparse.cpp:48-109 with T = 1, the graph test_image.cpp:156-168 builds), one launch
// per iteration with the step in the last workgroup, as in the 1-D pass.
//
// A wave owns a tile of IMG_OWN = 61 image columns x `rows` image rows and walks along it one
// image row at a time.  Lane l holds column c0 - 1 + l: lanes 1..61 are the owned columns, lane
// 0 and lane 63 only rebuild gamma_i (the left and right neighbours of the owned edge columns),
// lane 62 also forms d_i (the right end of the last owned column's edge).  Left and right
// neighbours come from the adjacent lanes (DPP lane shifts), the row above and below from the
// registers of the walk, so each wave reads its rows once and needs neither LDS nor barriers.
// Per image row y (even bands walk down, odd bands up) the wave
//   * rebuilds gamma_i on the next row of the walk: gamma_{i-1} + beta_{i-1} (p_{i-1} -
//     gamma_{i-1}) from the K - 1 stored rows of both, or, when beta_{i-1} = 1 (most passes),
//     p_{i-1} itself, without reading gamma_{i-1},
//   * forms g_i = alpha L gamma_i + q, p_i = P(gamma_i - alpha_i g_i), d_i = p_i - gamma_i and,
//     on passes that want it, the natural KKT residual on row y,
//   * stores gamma_i and p_i (K - 1 rows) at its owned nodes and adds |d|^2, <g, d>, f and the
//     residual of those nodes,
//   * adds alpha (d_a - d_b)^2 for the right edges of row y and the vertical edges between row
//     y and the row before it in the walk (each vertical edge counted by the tile of its upper
//     end), one row past the tile (the first row of the tile below) being formed for the last
//     vertical edges.
// The data term q (K rows, constant while theta is) is formed once per phase by img_data_term.
// The raw rows of gamma_{i-1}, p_{i-1} and q are loaded PD rows before gamma_i is formed on
// them.  Plain stores: nontemporal ones measured slower, although no lane reads gamma_i or p_i
// again in the pass.
// Every node's Laplacian (walk order up, left, right, down, a missing neighbour's term
// multiplied by 0) and projection are those of the node-wise graph kernels; gamma_i = p_{i-1}
// when beta_{i-1} = 1 and the sums over the nodes in tile order differ from those kernels' in
// rounding only.
#pragma once

#include "spg_kernels.h"

namespace pasc {

constexpr int IMG_OWN = WAVE - 3;  // owned columns of a wave tile

struct ImgArgs {
    void *gamma0;
    void *gamma1;
    void *dir0;  // p_{i-1} / p_i ping-pong, indexed like gamma by the instance parity
    void *dir1;
    const void *q;  // data term rows, laid out like gamma
    const InstDesc *desc;
    InstState *state;
    double *part;
    GroupView grp;
    int32_t width, height;  // image
    int32_t rows;           // image rows per wave tile
    int32_t ncol;           // wave tiles per band of rows: ceil(width / IMG_OWN)
    int32_t kkt_every;      // evaluate the KKT residual on passes with it % kkt_every == 0
    StepParams sp;
    int32_t *done;          // fused step (single problems, as PassArgs::done) or null
    int32_t *nactive;
    int32_t *dirty;
};

// Data term of every node of the instances with K clusters, q_k = bscale ||x - theta_k||^2 (0
// held out) or -b, with grid_dataterm_u's arithmetic, into rows laid out like gamma.
template <typename S, int K>
__global__ __launch_bounds__(256) void img_data_term(const InstDesc *desc, int n_inst,
                                                     const void *x, S *q) {
    for (int i = blockIdx.y; i < n_inst; i += gridDim.y) {
        const InstDesc &D = desc[i];
        if (D.K != K) continue;
        const GridData gd = grid_data(D);
        S *qi = q + D.goff;
        for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < D.T;
             r += (int64_t)gridDim.x * 256) {
            S v[K];
            grid_dataterm_u<S, K>(gd, x, r, sizeof(S) == 4, v);
#pragma unroll
            for (int k = 0; k < K; k++) qi[k * D.ld + r] = v[k];
        }
    }
}

// How a pass forms gamma_i from the previous pass's rows: IMG_FIRST projects gamma_0 (K rows);
// IMG_BETA1 takes gamma_i = p_{i-1}, the previous pass's projected point (beta_{i-1} = 1, the
// common case: gamma_{i-1} is not read); IMG_GENERAL forms gamma_{i-1} + beta_{i-1} (p_{i-1} -
// gamma_{i-1}).
constexpr int IMG_FIRST = 0, IMG_BETA1 = 1, IMG_GENERAL = 2;

// The loads of one image row at this lane's column: the rows of gamma_{i-1} and p_{i-1} the mode
// needs (K - 1 stored rows each; K rows of gamma_0 for IMG_FIRST) and the K rows of the data
// term.
template <typename S, int K, int MODE>
struct ImgRaw {
    static constexpr int KS = stored_rows<K>();
    static constexpr int NG = MODE == IMG_FIRST ? K : MODE == IMG_GENERAL ? KS : 0;
    static constexpr int NP = MODE == IMG_FIRST ? 0 : KS;
    S g[NG > 0 ? NG : 1];
    S p[NP > 0 ? NP : 1];
    S q[K];
};

// Per-instance constants of a pass in scalar registers.
template <typename S, typename C, int K>
struct ImgCtx {
    const S *src;    // gamma_{i-1}
    S *dst;          // gamma_i
    const S *pprev;  // p_{i-1} = P(gamma_{i-1} - alpha_{i-1} g_{i-1})
    S *pout;         // p_i
    const S *q;      // data term
    uint32_t ld;     // K * ld * sizeof(S) < 4 GiB (engine_is_image)
    C alpha, a_dir, s_kkt, beta;
    double alpha_d, a_dir_d;
    bool want_kkt;
};

template <typename S, typename C, int K>
__device__ __forceinline__ void make_img_ctx(const ImgArgs &A, const InstDesc &D,
                                             const InstState &st, ImgCtx<S, C, K> &c) {
    const int par = uni(st.parity);
    const int64_t goff = uni(D.goff);
    c.ld = (uint32_t)uni(D.ld);
    c.src = reinterpret_cast<const S *>(par ? A.gamma1 : A.gamma0) + goff;
    c.dst = reinterpret_cast<S *>(par ? A.gamma0 : A.gamma1) + goff;
    c.pprev = reinterpret_cast<const S *>(par ? A.dir1 : A.dir0) + goff;
    c.pout = reinterpret_cast<S *>(par ? A.dir0 : A.dir1) + goff;
    c.q = reinterpret_cast<const S *>(A.q) + goff;
    c.alpha = C(uni(D.alpha));
    c.alpha_d = uni(D.alpha);
    c.a_dir = C(uni(st.alpha_dir));
    c.a_dir_d = (double)c.a_dir;
    c.s_kkt = C(uni(D.s_kkt));
    c.beta = C(uni(st.beta_apply));
    const int it = uni(st.it);
    c.want_kkt = (it % (A.kkt_every > 0 ? A.kkt_every : 1)) == 0;
}

// Element i of an array whose elements all lie within 4 GiB of its base, addressed as the base
// (scalar registers) plus a 32-bit byte offset (one vector register).
template <typename T>
__device__ __forceinline__ const T &at32(const T *base, uint32_t i) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                        (uint32_t)(i * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ T &at32(T *base, uint32_t i) {
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) +
                                  (uint32_t)(i * (uint32_t)sizeof(T)));
}

template <typename S, typename C, int K, int MODE>
__device__ __forceinline__ void img_load(const ImgCtx<S, C, K> &c, uint32_t n,
                                         ImgRaw<S, K, MODE> &r) {
    using R = ImgRaw<S, K, MODE>;
    const uint32_t ld = c.ld;
#pragma unroll
    for (int k = 0; k < R::NG; k++) r.g[k] = at32(c.src, k * ld + n);
#pragma unroll
    for (int k = 0; k < R::NP; k++) r.p[k] = at32(c.pprev, k * ld + n);
#pragma unroll
    for (int k = 0; k < K; k++) r.q[k] = at32(c.q, k * ld + n);
}

// gamma_i from the raw row (IMG_GENERAL: grid_gamma_u's arithmetic, with d_{i-1} = p_{i-1} -
// gamma_{i-1} formed in registers)
template <typename S, typename C, int K, int MODE>
__device__ __forceinline__ void img_gamma(const ImgRaw<S, K, MODE> &r, C beta, C (&v)[K]) {
    constexpr int KS = stored_rows<K>();
    if constexpr (MODE == IMG_FIRST) {
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = C(r.g[k]);
        project_simplex<K, C>(v);
    } else {
        if constexpr (MODE == IMG_BETA1) {
#pragma unroll
            for (int k = 0; k < KS; k++) v[k] = C(r.p[k]);
        } else {
#pragma unroll
            for (int k = 0; k < KS; k++) v[k] = C(r.g[k]) + beta * (C(r.p[k]) - C(r.g[k]));
        }
        if (KS < K) v[K - 1] = last_row(v);
    }
}

// One wave tile: `slot` = band * ncol + column tile.  Even bands walk down, odd bands up, so
// two adjacent bands reach the rows they share at the same time (both at the end of their
// walks, or both at the start), while those rows are still in L2.  acc[2] collects
// sum(|d|^2 + off) and acc[3] sum((g + q) gamma); the caller scales them.
template <typename S, typename C, int K, int MODE, int PD, bool DOWN>
__device__ __forceinline__ void img_tile(const ImgCtx<S, C, K> &c, int slot, int W, int H,
                                         int rows, int ncol, double (&acc)[NPART]) {
    constexpr int KS = stored_rows<K>();
    constexpr int NB = PD + 1;
    constexpr int STEP = DOWN ? 1 : -1;
    using Raw = ImgRaw<S, K, MODE>;
    const int lane = threadIdx.x & (WAVE - 1);
    const int band = slot / ncol, ct = slot - band * ncol;
    const int col = ct * IMG_OWN - 1 + lane;
    const int cc = col < 0 ? 0 : (col >= W ? W - 1 : col);  // clamped: loads stay in the image
    const bool own_col = lane >= 1 && lane <= IMG_OWN && col < W;
    // weights of the left and right neighbour terms and of this lane's right edge (0 or 1)
    const C wl = col > 0 ? C(1) : C(0), wr = col < W - 1 ? C(1) : C(0);
    const double we = (own_col && col < W - 1) ? 1.0 : 0.0, wo = own_col ? 1.0 : 0.0;
    const int r0 = band * rows;
    const int r1 = r0 + rows < H ? r0 + rows : H;  // owned rows [r0, r1)
    const int ylast = r1 < H ? r1 : H - 1;          // rows [r0, ylast] get d_i (r1: halo)
    const int ys = DOWN ? r0 : ylast;  // first row of the walk
    const int yb = ys - STEP;          // the row behind the first
    const uint32_t ld = c.ld;
    auto node = [&](int y) {
        const int yc = y < 0 ? 0 : (y < H ? y : H - 1);
        return (uint32_t)(yc * W + cc);
    };

    // gamma_i on the rows behind (gB), at (gC) and ahead of (gA) the walk; d_i behind (dB)
    C gB[K], gC[K], gA[K], dB[K];
#pragma unroll
    for (int k = 0; k < K; k++) { gB[k] = C(0); dB[k] = C(0); gA[k] = C(0); }
    if (yb >= 0 && yb < H) {
        Raw behind;
        img_load<S, C, K, MODE>(c, node(yb), behind);
        img_gamma<S, C, K, MODE>(behind, c.beta, gB);
    }
    Raw buf[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) img_load<S, C, K, MODE>(c, node(ys + i * STEP), buf[i]);
    img_gamma<S, C, K, MODE>(buf[0], c.beta, gC);

    const int nsteps = ylast - r0 + 1;
    for (int i0 = 0; i0 < nsteps; i0 += NB) {
#pragma unroll
        for (int u = 0; u < NB; u++) {
            if (i0 + u >= nsteps) break;
            const int y = ys + (i0 + u) * STEP;
            // the data term of row y, then the slot refilled with row y + NB steps on
            C q[K];
#pragma unroll
            for (int k = 0; k < K; k++) q[k] = C(buf[u].q[k]);
            img_load<S, C, K, MODE>(c, node(y + NB * STEP), buf[u]);
            const int ya = y + STEP;
            if (ya >= 0 && ya < H) img_gamma<S, C, K, MODE>(buf[(u + 1) % NB], c.beta, gA);
            const C wu = y > 0 ? C(1) : C(0), wd = y + 1 < H ? C(1) : C(0);
            // Laplacian in the walk order up, left, right, down
            C lap[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const C gl = dpp_shift<DPP_WAVE_SHR1>(gC[k]);
                const C gr = dpp_shift<DPP_WAVE_SHL1>(gC[k]);
                const C gup = DOWN ? gB[k] : gA[k], gdn = DOWN ? gA[k] : gB[k];
                C l = wu * (gC[k] - gup);
                l = l + wl * (gC[k] - gl);
                l = l + wr * (gC[k] - gr);
                l = l + wd * (gC[k] - gdn);
                lap[k] = l;
            }
            C g[K], yv[K], p[K], z[K], d[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                g[k] = c.alpha * lap[k] + q[k];
                yv[k] = gC[k] - c.a_dir * g[k];
                p[k] = yv[k];
                z[k] = gC[k] - c.s_kkt * g[k];
            }
            const C tau = project_simplex<K, C>(p);
#pragma unroll
            for (int k = 0; k < K; k++) d[k] = p[k] - gC[k];
            const bool own = own_col && y < r1;
            if (own) {
                const uint32_t n = (uint32_t)(y * W + col);
#pragma unroll
                for (int k = 0; k < KS; k++) {
                    at32(c.dst, k * ld + n) = S(gC[k]);
                    at32(c.pout, k * ld + n) = S(p[k]);
                }
            }
            const double wn = own ? 1.0 : 0.0;
            double dd = 0.0, off = 0.0, fq = 0.0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                dd += (double)d[k] * (double)d[k];
                // the off-support term of <g, d>: (tau - y_k) gamma_k where y_k <= tau
                const C o = tau - yv[k];
                off += (double)(o > C(0) ? o : C(0)) * (double)gC[k];
                fq += ((double)g[k] + (double)q[k]) * (double)gC[k];
            }
            acc[0] += wn * dd;
            acc[2] += wn * (dd + off);
            acc[3] += wn * fq;
            if (c.want_kkt) {
                project_simplex<K, C>(z);
                double e = 0.0;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double ek = fabs((double)gC[k] - (double)z[k]);
                    e = (ek > e || ek != ek) ? ek : e;
                }
                e = own ? e : 0.0;
                acc[4] = (e > acc[4] || e != e) ? e : acc[4];
            }
            // d'A d: the right edge of (y, col) and the vertical edge to the row behind, owned
            // by its upper end: (y - 1, y) walking down, (y, y + 1) walking up
            double eh = 0.0, ev = 0.0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const C dr = dpp_shift<DPP_WAVE_SHL1>(d[k]);
                const double e1 = (double)d[k] - (double)dr;
                const double e2 = (double)dB[k] - (double)d[k];
                eh += e1 * e1;
                ev += e2 * e2;
            }
            const bool vown = DOWN ? y > r0 : (y < r1 && y + 1 <= ylast);
            acc[1] += c.alpha_d * ((y < r1 ? we : 0.0) * eh + (vown ? wo : 0.0) * ev);
#pragma unroll
            for (int k = 0; k < K; k++) {
                gB[k] = gC[k];
                gC[k] = gA[k];
                dB[k] = d[k];
            }
        }
    }
}

template <typename S, typename C, int K, int PD>
__global__ __launch_bounds__(WAVE) void spg_pass_img(ImgArgs A) {
    __shared__ double lds_red[NPART];
    __shared__ uint64_t groups_done;  // fused step: bit g set if this workgroup completed group g
    __shared__ int32_t step_due;      // fused step: this workgroup completed the last group
    __shared__ InstState lds_state;   // fused step
    if (threadIdx.x == 0) {
        groups_done = 0;
        step_due = 0;
    }
    const int total = A.grp.off[A.grp.n];
    const int W = A.width, H = A.height, rows = A.rows, ncol = A.ncol;
    for (int j = blockIdx.x; j < total; j += gridDim.x) {
        int inst, slot;
        find_tile(A.grp, xcd_tile(j, total), inst, slot);
        inst = uni(inst);
        slot = uni(slot);
        if (uni(A.state[inst].active) == 0) continue;
        const InstDesc &D = A.desc[inst];
        const InstState &st = A.state[inst];
        ImgCtx<S, C, K> cx;
        make_img_ctx<S, C, K>(A, D, st, cx);
        double acc[NPART] = {0.0, 0.0, 0.0, 0.0, cx.want_kkt ? 0.0 : -1.0};
        const bool down = ((slot / ncol) & 1) == 0;
        const int mode = uni(st.first) ? IMG_FIRST
                                       : (uni(st.beta_apply) == 1.0 ? IMG_BETA1 : IMG_GENERAL);
#define PASC_IMG_TILE(M)                                                                   \
    do {                                                                                   \
        if (down) img_tile<S, C, K, M, PD, true>(cx, slot, W, H, rows, ncol, acc);         \
        else img_tile<S, C, K, M, PD, false>(cx, slot, W, H, rows, ncol, acc);             \
    } while (0)
        if (mode == IMG_FIRST) PASC_IMG_TILE(IMG_FIRST);
        else if (mode == IMG_BETA1) PASC_IMG_TILE(IMG_BETA1);
        else PASC_IMG_TILE(IMG_GENERAL);
#undef PASC_IMG_TILE
        // <g, d> in the sign-safe form of the projected-gradient step (pass_tile_1d), f = 1/2
        // <g + q, gamma>
        acc[2] = -acc[2] / cx.a_dir_d;
        acc[3] *= 0.5;
        block_reduce_part<WAVE>(acc, lds_red);
        if (threadIdx.x == 0) {
            const int32_t slot_out = uni(D.tile0) + slot;
            double *o = A.part + (int64_t)slot_out * NPART;
            if (A.done) {
                // published and counted as in spg_pass_1d
                const int g = slot % DONE_SUB;
#pragma unroll
                for (int i = 0; i < NPART; i++) store_coherent(o + i, acc[i]);
                if (count_slot(A.done, g, uni(D.ntiles))) groups_done |= 1ull << g;
            } else {
#pragma unroll
                for (int i = 0; i < NPART; i++) o[i] = acc[i];
            }
        }
    }
    if (A.done) {
        __syncthreads();
        uint64_t gmask = (uint64_t)uni((int64_t)groups_done);
        if (gmask) {
            const int ns = uni(A.desc[0].ntiles);
            double *part0 = A.part + (int64_t)uni(A.desc[0].tile0) * NPART;
            while (gmask) {
                const int g = __builtin_ctzll(gmask);
                gmask &= gmask - 1;
                if (reduce_group<WAVE>(A.done, part0, g, ns, lds_red)) step_due = 1;
            }
            __syncthreads();
            if (step_due)
                fused_step<WAVE>(A.desc, A.state, A.done, A.sp, A.dirty, A.nactive, 0, &lds_state,
                                 lds_red);
        }
    }
}

}  // namespace pasc

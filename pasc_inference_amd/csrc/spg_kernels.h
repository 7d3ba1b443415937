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
// (neighbouring threads load the same cache lines; HBM sees each element once).  Only K - 1
// rows of gamma are read and written (the last is 1 - the others), and for a single problem
// the workgroup that completes the pass also runs the step, so an iteration is one launch.
#pragma once

#include "device_common.h"
#include "step_scalar.h"

namespace pasc {

struct PassArgs {
    void *gamma0;
    void *gamma1;
    const void *x;
    const InstDesc *desc;
    InstState *state;
    double *part;
    GroupView grp;
    int32_t kkt_every;  // evaluate the KKT residual on passes with it % kkt_every == 0
    int32_t need_f;     // 0: f comes from the quadratic identity (only the first pass sums it)
    const void *b;      // explicit linear terms of the phase's type, indexed like gamma (or null)
    // fused step (single problems): the slots are reduced in DONE_SUB groups as they complete,
    // and the workgroup that completes the last group reduces the group partials and runs the
    // SPG step (no separate step launch)
    StepParams sp;
    int32_t *done;      // DONE_WORDS counters of completed slots in this pass (back to 0
                        // after), then DONE_SUB group partials; null: the step kernel runs
                        // the step
    int32_t *nactive;   // active instances (decremented when one stops)
    int32_t *dirty;     // set when an instance stops (the batch's tile scan rebuilds)
};

template <int CH>
constexpr int halo_threads() { return (3 + CH - 1) / CH; }

// Rows of gamma the 1-D pass reads and writes.  Every iterate after the first pass of a phase
// lies on the simplex, so its last row is 1 - (sum of the others): the pass stores K - 1 rows
// and rebuilds the last one in registers (clamped at 0, so that it stays exactly feasible; the
// sum is then 1 up to rounding, as for the projection's own output).  complete_last_row writes
// the full row once a phase ends, so that every other reader sees K rows.
template <int K>
constexpr int stored_rows() { return K >= 2 ? K - 1 : K; }

template <int NT, int CH>
constexpr int tile_steps() { return (NT - 2 * halo_threads<CH>()) * CH; }

// Loads CH consecutive steps [t0, t0+CH) of one row (t0 a multiple of the vector width);
// steps outside [0, T) read as 0.  INT: the caller guarantees [t0, t0+CH) inside [0, T).
template <typename S, typename C, int CH, bool INT = false>
__device__ __forceinline__ void load_chunk(const S *__restrict__ row, int64_t t0, int64_t T,
                                           int64_t ld, C (&v)[CH]) {
    using VT = V16<S>;
    constexpr int VW = VT::n;
    static_assert(CH % VW == 0, "chunk must be whole vectors");
#pragma unroll
    for (int q = 0; q < CH / VW; q++) {
        const int64_t t = t0 + q * VW;
        if (INT) {
            const typename VT::type vec = *reinterpret_cast<const typename VT::type *>(row + t);
#pragma unroll
            for (int i = 0; i < VW; i++) v[q * VW + i] = C(VT::get(vec, i));
        } else if (t >= 0 && t < ld) {
            const typename VT::type vec = *reinterpret_cast<const typename VT::type *>(row + t);
#pragma unroll
            for (int i = 0; i < VW; i++) v[q * VW + i] = (t + i < T) ? C(VT::get(vec, i)) : C(0);
        } else {
#pragma unroll
            for (int i = 0; i < VW; i++) v[q * VW + i] = C(0);
        }
    }
}

// Whole-wave lane shifts by one lane (DPP wave_shr:1 / wave_shl:1, VALU moves): lane i gets
// lane i-1 (from_left) or lane i+1 (from_right); the lane without a source gets 0 (the
// bound-control bit: no register has to be set to 0 before each move).
template <int CTRL>
__device__ __forceinline__ float dpp_shift(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffll), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)(unsigned)hi << 32) | (long long)(unsigned)lo);
}
constexpr int DPP_WAVE_SHR1 = 0x138;  // lane i <- lane i-1
constexpr int DPP_WAVE_SHL1 = 0x130;  // lane i <- lane i+1

// Neighbour values across chunk edges: every thread publishes its first and last step of each
// row; thread i then reads step t0-1 from thread i-1 and step t0+CH from thread i+1.  A
// one-wave workgroup (NT = 64) moves them with DPP lane shifts: no LDS and no barrier, so its
// waves never wait for each other.
template <typename C, int K, int CH, int NT>
__device__ __forceinline__ void exchange_edges(const C (&v)[K][CH], C *lds, C (&left)[K],
                                               C (&right)[K]) {
    const int tid = threadIdx.x;
    if constexpr (NT == WAVE) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            left[k] = dpp_shift<DPP_WAVE_SHR1>(v[k][CH - 1]);
            right[k] = dpp_shift<DPP_WAVE_SHL1>(v[k][0]);
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        lds[(2 * k) * NT + tid] = v[k][0];
        lds[(2 * k + 1) * NT + tid] = v[k][CH - 1];
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < K; k++) {
        left[k] = tid > 0 ? lds[(2 * k + 1) * NT + tid - 1] : C(0);
        right[k] = tid < NT - 1 ? lds[(2 * k) * NT + tid + 1] : C(0);
    }
    lds_barrier();
}

template <typename C>
__device__ __forceinline__ C lap_half(C v, C vm, C vp, int64_t t, int64_t T, bool interior) {
    // half of row t of A0 (blockgraphsparse.cpp:60-89) in difference form:
    // (v_t - v_{t-1})[t>0] + (v_t - v_{t+1})[t<T-1]; the callers scale by 2*alpha (the factor
    // 2 is exact, so (2 alpha) * s equals alpha * (2 s) bit for bit, one multiply fewer)
    if (interior) return (v - vm) + (v - vp);
    C a = C(0);
    if (t > 0) a += v - vm;
    if (t < T - 1) a += v - vp;
    return a;
}

// One tile of the fused pass.  INT: every step of the tile, halos included, lies in [1, T-2]
// (the common case; no bounds or boundary-row tests).  Per-thread partials accumulate in the
// compute type C over the thread's owned steps and enter the FP64 slot sums once per tile.
template <typename S, typename C, int K, int CH, int NT>
__device__ __forceinline__ bool tile_is_interior(int64_t tile, int64_t T) {
    constexpr int64_t TILE = tile_steps<NT, CH>();
    constexpr int64_t H = (int64_t)halo_threads<CH>() * CH;
    return tile * TILE - H >= 1 && (tile + 1) * TILE + H <= T - 1;
}

// Per-instance constants of one pass, read once per partial slot into scalar registers (the
// descriptor and state loads are then never waited for inside a tile, where such a wait would
// also drain the asynchronous copies of the next tile).
template <typename S, typename C, int K>
struct TileCtx {
    int64_t T, ld;
    const S *src;          // gamma_{i-1} rows
    S *dst;                // gamma_i rows
    const S *xr;           // data row
    const S *br;           // explicit b rows or nullptr
    const uint8_t *heldout;
    C alpha, bscale, s_kkt, a_apply, b_apply, a_dir;
    C theta[K];
    bool first, need_f, want_kkt;
};

template <typename S, typename C, int K>
__device__ __forceinline__ void make_tile_ctx(const PassArgs &A, const InstDesc &D,
                                              const InstState &st, TileCtx<S, C, K> &c) {
    // pointers are formed from the kernel arguments plus scalar offsets, so that they stay
    // global-memory pointers in scalar registers
    const int par = uni(st.parity);
    const int64_t goff = uni(D.goff);
    c.T = uni(D.T);
    c.ld = uni(D.ld);
    c.src = reinterpret_cast<const S *>(par ? A.gamma1 : A.gamma0) + goff;
    c.dst = reinterpret_cast<S *>(par ? A.gamma0 : A.gamma1) + goff;
    c.xr = reinterpret_cast<const S *>(A.x) + uni(D.xoff);
    const int has_b = uni(D.b64 != nullptr ? 1 : 0);
    c.br = has_b ? reinterpret_cast<const S *>(A.b) + goff : nullptr;
    c.heldout = D.heldout;
    c.alpha = uni(C(D.alpha));
    c.bscale = uni(C(D.bscale));
    c.s_kkt = uni(C(D.s_kkt));
    c.a_apply = uni(C(st.alpha_apply));
    c.b_apply = uni(C(st.beta_apply));
    c.a_dir = uni(C(st.alpha_dir));
#pragma unroll
    for (int k = 0; k < K; k++) c.theta[k] = uni(C(D.theta[k]));
    c.first = uni(st.first) != 0;
    c.need_f = A.need_f || c.first;
    const int it = uni(st.it);
    c.want_kkt = (it % (A.kkt_every > 0 ? A.kkt_every : 1)) == 0;
}

// ---- staging of a tile in LDS (asynchronous global -> LDS copies) ----------------------------
// The stage holds the K gamma rows and the data row of one tile, each NT*CH contiguous elements
// (steps T0 - HT*CH .. T0 + TILE + HT*CH), in the order they lie in HBM.  global_load_lds_dwordx4
// writes the 16 bytes each lane loads at (LDS base + 16 * lane), so wave w's q-th copy of a row
// covers the row's 16-byte pieces (w*Q + q)*64 .. +63, Q = CH*sizeof(S)/16 per wave and row.
typedef __attribute__((address_space(3))) void lds_void_t;

template <typename S, int K, int CH, int NT>
struct Stage {
    static constexpr int ROW = NT * CH;             // elements per staged row
    static constexpr int Q = CH * (int)sizeof(S) / 16;  // 16-byte copies per lane and row
    static constexpr int ROWS = stored_rows<K>() + 1;  // the stored gamma rows + the data row
    static constexpr size_t BYTES = (size_t)ROWS * ROW * sizeof(S);
    static_assert(Q >= 1 && CH * sizeof(S) % 16 == 0, "a thread's chunk must be whole 16-byte pieces");
};

// Issues the copies of one interior tile into the stage (every lane of every wave takes part).
// The data row is copied also when the instance has an explicit b (then it is not read), so
// the number of copies in flight is the same on every path.
template <typename S, typename C, int K, int CH, int NT>
__device__ __forceinline__ void stage_issue(const TileCtx<S, C, K> &c, int tile, S *stage) {
    using SG = Stage<S, K, CH, NT>;
    constexpr int HT = halo_threads<CH>();
    constexpr int TILE = tile_steps<NT, CH>();
    const int64_t lo = (int64_t)tile * TILE - (int64_t)HT * CH;  // >= 1 for an interior tile
    constexpr int KS = stored_rows<K>();
    const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int r = 0; r < KS + 1; r++) {
        const char *g = reinterpret_cast<const char *>(r < KS ? c.src + (int64_t)r * c.ld + lo
                                                              : c.xr + lo);
        char *l = reinterpret_cast<char *>(stage + (size_t)r * SG::ROW);
#pragma unroll
        for (int q = 0; q < SG::Q; q++) {
            const int piece = (w * SG::Q + q) * WAVE;
            __builtin_amdgcn_global_load_lds(
                reinterpret_cast<const void *>(g + (size_t)(piece + lane) * 16),
                (lds_void_t *)(l + (size_t)piece * 16), 16, 0, 0);
        }
    }
}

// This thread's CH steps of staged row r.
template <typename S, typename C, int K, int CH, int NT>
__device__ __forceinline__ void stage_read(const S *stage, int r, C (&v)[CH]) {
    using VT = V16<S>;
    constexpr int VW = VT::n;
    const S *row = stage + (size_t)r * Stage<S, K, CH, NT>::ROW + threadIdx.x * CH;
#pragma unroll
    for (int q = 0; q < CH / VW; q++) {
        const typename VT::type vec = *reinterpret_cast<const typename VT::type *>(row + q * VW);
#pragma unroll
        for (int i = 0; i < VW; i++) v[q * VW + i] = C(VT::get(vec, i));
    }
}

// STAGED: the tile's gamma and data rows were copied into the LDS stage (interior tiles
// only).  next: a callable run once every thread has taken the stage's rows into registers
// (it issues the copies of the slot's next tile into the stage).
template <typename S, typename C, int K, int CH, int NT, bool INT, bool STAGED, typename Next>
__device__ __forceinline__ void pass_tile_1d(const TileCtx<S, C, K> &cx, int chunk, C *lds_x,
                                             double (&acc)[NPART], const S *stage, Next next) {
    static_assert(!STAGED || INT, "only interior tiles are staged");
    constexpr int HT = halo_threads<CH>();
    constexpr int TILE = tile_steps<NT, CH>();
    const int64_t T = cx.T;
    const int64_t ld = cx.ld;
    const int tid = threadIdx.x;
    const int64_t T0 = (int64_t)chunk * TILE;
    const int64_t t0 = T0 - (int64_t)HT * CH + (int64_t)tid * CH;
    const bool owner = tid >= HT && tid < NT - HT;  // owned steps are stored and reduced
    const S *src = cx.src;
    S *dst = cx.dst;
    const C alpha = cx.alpha;
    const C alpha2 = C(2) * alpha;
    const C bscale = cx.bscale;
    const C s_kkt = cx.s_kkt;
    const C a_apply = cx.a_apply;
    const C b_apply = cx.b_apply;
    const C a_dir = cx.a_dir;
    const bool first = cx.first;
    const bool need_f = cx.need_f;

    // data term q = -b of every step of the chunk, formed once for the two levels that use it:
    // bscale*(x - theta_k)^2 (zero when held out or outside [0,T)), or read from an explicit b
    C Q[K][CH];
    if (cx.br) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            load_chunk<S, C, CH, INT>(cx.br + k * ld, t0, T, ld, Q[k]);
#pragma unroll
            for (int c = 0; c < CH; c++) Q[k][c] = -Q[k][c];
        }
    } else {
        C X[CH];
        if (STAGED)
            stage_read<S, C, K, CH, NT>(stage, stored_rows<K>(), X);
        else
            load_chunk<S, C, CH, INT>(cx.xr, t0, T, ld, X);
        bool obs[CH];
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const int64_t t = t0 + c;
            obs[c] = INT || (t >= 0 && t < T);
        }
        if (cx.heldout) {
#pragma unroll
            for (int c = 0; c < CH; c++)
                if (obs[c]) obs[c] = cx.heldout[t0 + c] == 0;
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const C th = cx.theta[k];
#pragma unroll
            for (int c = 0; c < CH; c++) {
                const C dx = X[c] - th;
                Q[k][c] = obs[c] ? bscale * dx * dx : C(0);
            }
        }
    }
    // every step of the chunk is interior (both time neighbours exist): no boundary selects
    const bool interior = INT || (t0 >= 1 && t0 + CH <= T - 1);
    const bool want_kkt = cx.want_kkt;

    // level 1: gamma_i = gamma_{i-1} + beta_{i-1} (P(gamma_{i-1} - alpha_{i-1} g_{i-1}) - gamma_{i-1})
    constexpr int KS = stored_rows<K>();
    C G1[K][CH];
    {
        C G[K][CH];
#pragma unroll
        for (int k = 0; k < KS; k++) {
            if (STAGED)
                stage_read<S, C, K, CH, NT>(stage, k, G[k]);
            else
                load_chunk<S, C, CH, INT>(src + k * ld, t0, T, ld, G[k]);
        }
        if (KS < K) {
            if (first) {
                // gamma_0 may lie off the simplex: its last row is read (first passes are
                // never staged)
                load_chunk<S, C, CH, INT>(src + (K - 1) * ld, t0, T, ld, G[K - 1]);
            } else {
#pragma unroll
                for (int c = 0; c < CH; c++) {
                    C r = C(1);
#pragma unroll
                    for (int k = 0; k < K - 1; k++) r -= G[k][c];
                    const bool valid = INT || (t0 + c >= 0 && t0 + c < T);
                    G[K - 1][c] = valid ? (r > C(0) ? r : C(0)) : C(0);
                }
            }
        }
        C gl[K], gr[K];
        exchange_edges<C, K, CH, NT>(G, lds_x, gl, gr);
        // every stage read must be done before the next copies overwrite the stage: the
        // exchange's barriers see to that, and a one-wave workgroup waits for its own reads
        if constexpr (NT == WAVE && STAGED) asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        next();
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const int64_t t = t0 + c;
            C y[K];
            if (first) {
#pragma unroll
                for (int k = 0; k < K; k++) y[k] = G[k][c];
                project_simplex<K, C>(y);
#pragma unroll
                for (int k = 0; k < K; k++) G1[k][c] = y[k];
            } else {
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const C vm = c > 0 ? G[k][c - 1] : gl[k];
                    const C vp = c < CH - 1 ? G[k][c + 1] : gr[k];
                    const C g = alpha2 * lap_half(G[k][c], vm, vp, t, T, interior) + Q[k][c];
                    y[k] = G[k][c] - a_apply * g;
                }
                project_simplex<K, C>(y);
#pragma unroll
                for (int k = 0; k < K; k++) G1[k][c] = G[k][c] + b_apply * (y[k] - G[k][c]);
            }
        }
    }
    // level 2: g_i, d_i = P(gamma_i - alpha_i g_i) - gamma_i, and on the owned steps |d|^2, the
    // off-support term of <g,d>, f_i and the KKT residual of gamma_i
    C Dd[K][CH];
    C p_dd = C(0), p_off = C(0), p_f = C(0), p_kkt = C(0);
    {
        C gl[K], gr[K];
        exchange_edges<C, K, CH, NT>(G1, lds_x, gl, gr);
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const int64_t t = t0 + c;
            C g[K], y[K], z[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const C vm = c > 0 ? G1[k][c - 1] : gl[k];
                const C vp = c < CH - 1 ? G1[k][c + 1] : gr[k];
                g[k] = alpha2 * lap_half(G1[k][c], vm, vp, t, T, interior) + Q[k][c];
                y[k] = G1[k][c] - a_dir * g[k];
                z[k] = y[k];
            }
            const C tau = project_simplex<K, C>(z);
            // <g, d> = -(|d|^2 + sum_{k off the support} (tau - y_k) gamma_k) / alpha holds for the
            // projected-gradient step (it follows from g = (gamma - y)/alpha and sum_k d_k = 0);
            // every term is non-negative, so the computed value cannot lose its sign to rounding
            // the way the plain dot product does once d is tiny.
            C o = C(0);
#pragma unroll
            for (int k = 0; k < K; k++) {
                Dd[k][c] = z[k] - G1[k][c];
                // (tau - y_k)^+ gamma_k: the off-support terms (y_k <= tau), as one max
                o += cmax(tau - y[k], C(0)) * G1[k][c];
            }
            if (owner && (INT || t < T)) {
                p_off += o;
#pragma unroll
                for (int k = 0; k < K; k++) p_dd += Dd[k][c] * Dd[k][c];
                if (need_f) {
#pragma unroll
                    for (int k = 0; k < K; k++) p_f += (g[k] + Q[k][c]) * G1[k][c];
                }
                if (want_kkt) {
#pragma unroll
                    for (int k = 0; k < K; k++) z[k] = G1[k][c] - s_kkt * g[k];
                    project_simplex<K, C>(z);
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const C e = fabs(G1[k][c] - z[k]);
                        p_kkt = (e > p_kkt || e != e) ? e : p_kkt;
                    }
                }
            }
        }
    }
    // level 3: d'A d = alpha * sum over edges (t, t+1) of 2 (d_{t+1} - d_t)^2, store gamma_i
    C p_dad = C(0);
    {
        C dl[K], dr[K];
        exchange_edges<C, K, CH, NT>(Dd, lds_x, dl, dr);
        if (owner) {
#pragma unroll
            for (int c = 0; c < CH; c++) {
                const int64_t t = t0 + c;
                if (INT || interior || t < T - 1) {
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const C dn = c < CH - 1 ? Dd[k][c + 1] : dr[k];
                        const C e = dn - Dd[k][c];
                        p_dad += e * e;
                    }
                }
            }
            using VT = V16<S>;
            constexpr int VW = VT::n;
#pragma unroll
            for (int k = 0; k < KS; k++) {
#pragma unroll
                for (int q = 0; q < CH / VW; q++) {
                    const int64_t t = t0 + q * VW;
                    if (INT || t + VW <= T) {
                        typename VT::type vec;
#pragma unroll
                        for (int i = 0; i < VW; i++) VT::set(vec, i, S(G1[k][q * VW + i]));
                        // FP64 rows: nontemporal stores (95.8 -> 87.1 us per pass at
                        // T = 1e7, K = 3); the FP32 pass was slower with them (49.5 -> 53.8 us)
                        if constexpr (sizeof(S) == 8)
                            store_nt(reinterpret_cast<typename VT::type *>(dst + k * ld + t), vec);
                        else
                            *reinterpret_cast<typename VT::type *>(dst + k * ld + t) = vec;
                    } else {
#pragma unroll
                        for (int i = 0; i < VW; i++)
                            if (t + i < T) dst[k * ld + t + i] = S(G1[k][q * VW + i]);
                    }
                }
            }
        }
    }
    const double dd = (double)p_dd;
    acc[0] += dd;
    acc[1] += 2.0 * (double)alpha * (double)p_dad;
    acc[2] -= (dd + (double)p_off) / (double)a_dir;
    acc[3] += 0.5 * (double)p_f;
    if (want_kkt) {
        const double e = (double)p_kkt;
        acc[4] = (e > acc[4] || e != e) ? e : acc[4];
    } else {
        acc[4] = -1.0;  // not evaluated on this pass
    }
}

// Group partials of the fused step: DONE_SUB x NPART doubles after the counters.
__device__ __forceinline__ double *group_partials(int32_t *done) {
    return reinterpret_cast<double *>(done + DONE_WORDS);
}

// The step of one instance, run by the workgroup that completed its last group: the group
// partials are reduced in a fixed order and step_scalar advances the instance.
template <int NT>
__device__ __forceinline__ void fused_step(const InstDesc *desc, InstState *state, int32_t *done,
                                           const StepParams &sp, int32_t *dirty, int32_t *nactive,
                                           int inst, InstState *lds_state, double *lds_red) {
    const InstDesc &D = desc[inst];
    const int ns = uni(D.ntiles);
    double *gp = group_partials(done);
    step_instance<NT>(state + inst, ns < DONE_SUB ? ns : DONE_SUB, D.alpha * D.s_kkt, sp, dirty,
                      lds_state, lds_red,
                      [&](int i, int q) { return load_coherent(gp + (int64_t)i * NPART + q); },
                      [&](int i, int q, double v) {
                          return settle_published(gp + (int64_t)i * NPART + q, v);
                      });
    // the group partials are taken: empty again for the next pass
    const int ng = ns < DONE_SUB ? ns : DONE_SUB;
    for (int i = threadIdx.x; i < ng * NPART; i += NT) reset_published(gp + i);
    if (threadIdx.x == 0 && !lds_state->active) atomicSub(nactive, 1);
}

// Counts one completed partial slot of a single problem (its values already written with
// store_coherent) on the sub-counter of its group g = slot % DONE_SUB.  True
// in the workgroup that completed the group; the counter is back to 0 afterwards.  Thread 0.
__device__ __forceinline__ bool count_slot(int32_t *done, int g, int ns) {
    const int n_g = ns / DONE_SUB + (g < ns % DONE_SUB ? 1 : 0);
    int32_t *sub = done + g * DONE_STRIDE;
    if (atomicAdd(sub, 1) != n_g - 1) return false;
    __hip_atomic_store(sub, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// Run by the whole workgroup that completed group g: its slots g, g + DONE_SUB, ... are reduced
// (per thread in slot order, then across the workgroup) into group partial g, which is
// published and counted on the top counter.  True in thread 0 of the workgroup that completed
// the last group (the instance's step is then due); the top counter is back to 0 afterwards.
// Each slot is reduced once, by the workgroup completing its group, so the pass's last
// workgroup only reduces its own group and the DONE_SUB group partials.
template <int NT>
__device__ __forceinline__ bool reduce_group(int32_t *done, double *part, int g, int ns,
                                             double *lds_red) {
    double v[NPART] = {0.0, 0.0, 0.0, 0.0, -1.0};  // residual -1: not evaluated
    for (int j = g + (int)threadIdx.x * DONE_SUB; j < ns; j += NT * DONE_SUB) {
        double w[NPART];
#pragma unroll
        for (int q = 0; q < NPART; q++) w[q] = load_coherent(part + (int64_t)j * NPART + q);
#pragma unroll
        for (int q = 0; q < NPART; q++) w[q] = settle_published(part + (int64_t)j * NPART + q, w[q]);
#pragma unroll
        for (int q = 0; q < NPART - 1; q++) v[q] += w[q];
        const double m = w[NPART - 1];
        v[NPART - 1] = (m > v[NPART - 1] || m != m) ? m : v[NPART - 1];
    }
    block_reduce_part<NT>(v, lds_red);
    bool last = false;
    if (threadIdx.x == 0) {
        double *o = group_partials(done) + g * NPART;
#pragma unroll
        for (int q = 0; q < NPART; q++) store_coherent(o + q, v[q]);
        int32_t *top = done + DONE_SUB * DONE_STRIDE;
        const int n_top = ns < DONE_SUB ? ns : DONE_SUB;
        last = atomicAdd(top, 1) == n_top - 1;
        if (last) __hip_atomic_store(top, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // the group's slots are taken: empty again for the next pass
    for (int j = g + (int)threadIdx.x * DONE_SUB; j < ns; j += NT * DONE_SUB)
#pragma unroll
        for (int q = 0; q < NPART; q++) reset_published(part + (int64_t)j * NPART + q);
    return last;
}

// PF: software pipelining across the msub tiles of a slot.  While a tile is computed, the
// next interior tile's rows are copied into an LDS stage by the asynchronous global -> LDS path
// (no registers held for it); the wait for those copies is the only added synchronisation.
// Configurations whose stage would exceed PASC_STAGE_MAX_BYTES run without it.
constexpr size_t PASC_STAGE_MAX_BYTES = 40 * 1024;

template <typename S, typename C, int K, int CH, int NT, bool PF>
__global__ __launch_bounds__(NT) void spg_pass_1d(PassArgs A) {
    using SG = Stage<S, K, CH, NT>;
    constexpr bool USE_STAGE = PF && SG::BYTES <= PASC_STAGE_MAX_BYTES;
    __shared__ double lds_red[NPART * (NT / WAVE)];
    __shared__ C lds_x[2 * K * NT];
    __shared__ __attribute__((aligned(16))) S stage[USE_STAGE ? SG::ROWS * SG::ROW : 1];
    __shared__ uint64_t groups_done;  // fused step: bit g set if this workgroup completed group g
    __shared__ int32_t step_due;      // fused step: this workgroup completed the last group
    __shared__ InstState lds_state;  // fused step
    constexpr int TILE = tile_steps<NT, CH>();
    if (threadIdx.x == 0) {
        groups_done = 0;
        step_due = 0;
    }
    const int total = A.grp.off[A.grp.n];
    for (int j = blockIdx.x; j < total; j += gridDim.x) {
        int inst, slot;
        find_tile(A.grp, j, inst, slot);
        inst = uni(inst);
        slot = uni(slot);
        // an instance that stopped keeps its tiles until the next scan rebuild (batches) or for
        // good (single problems): its slots are skipped
        if (uni(A.state[inst].active) == 0) continue;
        TileCtx<S, C, K> cx;
        make_tile_ctx<S, C, K>(A, A.desc[inst], A.state[inst], cx);
        const int msub = uni(A.desc[inst].msub);
        const int32_t slot_out = uni(A.desc[inst].tile0) + slot;
        // one partial slot covers msub consecutive tiles of the instance
        double acc[NPART] = {0.0, 0.0, 0.0, 0.0, 0.0};
        const int first_tile = slot * msub;
        bool staged = false;  // the current tile's rows are in the stage
        // a first pass reads all K rows of gamma_0 and runs unstaged
        const bool can_stage = USE_STAGE && !cx.first;
        if (can_stage && tile_is_interior<S, C, K, CH, NT>(first_tile, cx.T)) {
            stage_issue<S, C, K, CH, NT>(cx, first_tile, stage);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            staged = true;
        }
        for (int m = 0; m < msub; m++) {
            const int tile = first_tile + m;
            if ((int64_t)tile * TILE >= cx.T) break;
            const int nt = tile + 1;
            const bool stage_next = can_stage && m + 1 < msub && (int64_t)nt * TILE < cx.T &&
                                    tile_is_interior<S, C, K, CH, NT>(nt, cx.T);
            auto issue = [&]() {
                if (stage_next) stage_issue<S, C, K, CH, NT>(cx, nt, stage);
            };
            if (USE_STAGE && staged)
                pass_tile_1d<S, C, K, CH, NT, true, true>(cx, tile, lds_x, acc, stage, issue);
            else if (tile_is_interior<S, C, K, CH, NT>(tile, cx.T))
                pass_tile_1d<S, C, K, CH, NT, true, false>(cx, tile, lds_x, acc, stage, issue);
            else
                pass_tile_1d<S, C, K, CH, NT, false, false>(cx, tile, lds_x, acc, stage, issue);
            if (stage_next) {
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
            }
            staged = stage_next;
        }
        block_reduce_part<NT>(acc, lds_red);
        if (threadIdx.x == 0) {
            double *o = A.part + (int64_t)slot_out * NPART;
            if (A.done) {
                // Single problems: publish the slot, then count it on its group's counter (the
                // group is reduced after the tile loop; store_coherent / settle_published).  An
                // agent-scope release fence would instead write back every dirty line of this
                // XCD's L2 (the gamma rows in flight), which costs a sixth of the pass.
                const int g = slot % DONE_SUB;
#pragma unroll
                for (int i = 0; i < NPART; i++) store_coherent(o + i, acc[i]);
                if (count_slot(A.done, g, uni(A.desc[inst].ntiles))) groups_done |= 1ull << g;
            } else {
#pragma unroll
                for (int i = 0; i < NPART; i++) o[i] = acc[i];
            }
        }
    }
    // After the tile loop, so that these registers do not add to the loop's: the groups this
    // workgroup completed are reduced, and the workgroup that completes the last group runs the
    // step.  Every reduction order is fixed by the slot tiling, so the result does not depend on
    // which workgroups these are.  The fused step serves single problems (instance 0).
    if (A.done) {
        __syncthreads();
        uint64_t gmask = (uint64_t)uni((int64_t)groups_done);
        if (gmask) {
            const int ns = uni(A.desc[0].ntiles);
            double *part0 = A.part + (int64_t)uni(A.desc[0].tile0) * NPART;
            while (gmask) {
                const int g = __builtin_ctzll(gmask);
                gmask &= gmask - 1;
                if (reduce_group<NT>(A.done, part0, g, ns, lds_red)) step_due = 1;
            }
            __syncthreads();
            if (step_due)
                fused_step<NT>(A.desc, A.state, A.done, A.sp, A.dirty, A.nactive, 0, &lds_state,
                               lds_red);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// General graph stencil (image grids with T = 1, spatio-temporal graphs with T > 1 and R > 1,
// CSR or implicit 4-neighbour grid; blockgraphsparse.cpp:48-109).  Unfused: the first kernel
// rebuilds gamma_i node-wise from gamma_{i-1} + beta*d_{i-1} (also at the neighbours), forms
// g_i and d_i, and stores gamma_i and d_i; the second forms |d|^2 and d'Ad over the edges.
// Data is stored SoA, rows c = 0..xdim-1 of T*R nodes.
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
    GraphGeo geo;
};

// XCD-aware tile order of the graph kernels.  gfx950 dispatches workgroup w to XCD w % 8, and
// each XCD has its own L2.  A grid node's neighbours one image row up or down lie a few tiles
// away, so with tile j on workgroup j those rows were fetched by other XCDs' L2s.  Tile j of the
// grid-stride loop is remapped so that the tiles with j % 8 == x (all on XCD x when the grid
// is a multiple of 8) form the x-th contiguous eighth of [0, total): a band of image rows per
// XCD.  A bijection of [0, total) for any total; the partial slots and their order of summation
// do not change.
constexpr int XCDS = 8;
__device__ __forceinline__ int xcd_tile(int j, int total) {
    const int q = total / XCDS, rem = total % XCDS;
    const int x = j % XCDS, i = j / XCDS;
    return x * q + (x < rem ? x : rem) + i;
}

// Neighbour walks of the graph kernels.  IMG: an image (T == 1, implicit 4-neighbour grid,
// fewer than 2^31 nodes) in 32-bit index math, visiting the neighbours in for_spatial's order,
// so that every sum is bitwise the same as through the general walk.
template <bool IMG, typename F>
__device__ __forceinline__ void grid_neighbours(const GraphGeo &G, int64_t n, F &&f) {
    if constexpr (IMG) {
        const uint32_t r = (uint32_t)n, W = (uint32_t)G.width, col = r % W;
        if (r >= W) f(n - W, 1.0);
        if (col > 0) f(n - 1, 1.0);
        if (col < W - 1) f(n + 1, 1.0);
        if (r + W < (uint32_t)G.R) f(n + W, 1.0);
    } else {
        for_neighbours(G, n, f);
    }
}

template <bool IMG, typename F>
__device__ __forceinline__ void grid_forward_edges(const GraphGeo &G, int64_t n, F &&f) {
    if constexpr (IMG) {
        const uint32_t r = (uint32_t)n, W = (uint32_t)G.width, col = r % W;
        if (col < W - 1) f(n + 1, 1.0);
        if (r + W < (uint32_t)G.R) f(n + W, 1.0);
    } else {
        for_forward_edges(G, n, f);
    }
}

// The graph kernels store K - 1 rows of gamma and of d as the 1-D pass stores K - 1 rows of
// gamma: after the first pass of a phase gamma_i = gamma_{i-1} + beta d_{i-1} lies on the simplex,
// so its last component is (1 - the others)^+, and the last component of d (the difference of
// two points of the simplex) is minus the sum of the others.  complete_last_row writes the last
// gamma row when a phase ends.
template <typename C, int K>
__device__ __forceinline__ C last_row(const C (&v)[K]) {
    C rest = C(1);
#pragma unroll
    for (int k = 0; k < K - 1; k++) rest -= v[k];
    return rest > C(0) ? rest : C(0);
}

template <typename C, int K>
__device__ __forceinline__ C last_dir(const C (&d)[K]) {
    C s = C(0);
#pragma unroll
    for (int k = 0; k < K - 1; k++) s -= d[k];
    return s;
}

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
        constexpr int KS = stored_rows<K>();
#pragma unroll
        for (int k = 0; k < KS; k++) v[k] = C(src[k * D.ld + r]) + b * C(dr[k * D.ld + r]);
        if (KS < K) v[K - 1] = last_row(v);
    }
}

// gamma_i at node r from per-slot values held in registers (as grid_gamma_at); FIRST: the
// pass starts from an unprojected gamma_0
template <typename S, typename C, int K, bool FIRST>
__device__ __forceinline__ void grid_gamma_u(const S *src, const S *dprev, C beta, int64_t ld,
                                             int64_t r, C (&v)[K]) {
    if constexpr (FIRST) {
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = C(src[k * ld + r]);
        project_simplex<K, C>(v);
    } else {
        constexpr int KS = stored_rows<K>();
#pragma unroll
        for (int k = 0; k < KS; k++) v[k] = C(src[k * ld + r]) + beta * C(dprev[k * ld + r]);
        if (KS < K) v[K - 1] = last_row(v);
    }
}

// The data-term fields of one instance in scalar registers (read once per slot), and the data
// term q_k = bscale * ||x_r - theta_k||^2 of node r from them (0 at held-out nodes; -b for an
// explicit linear term).
struct GridData {
    const uint8_t *heldout;
    const double *b64;
    const float *b32;
    const double *theta;
    int64_t xoff, ld;
    int32_t xdim;
    double bscale;
};

__device__ __forceinline__ GridData grid_data(const InstDesc &D) {
    return GridData{uni(D.heldout), uni(D.b64), uni(D.b32), D.theta,
                    uni(D.xoff), uni(D.ld), uni(D.xdim), uni(D.bscale)};
}

template <typename C, int K>
__device__ __forceinline__ void grid_dataterm_u(const GridData &G, const void *xv, int64_t r,
                                                bool is_f32, C (&q)[K]) {
    if (G.b64) {  // explicit linear term: q = -b
#pragma unroll
        for (int k = 0; k < K; k++)
            q[k] = is_f32 ? -C(G.b32[k * G.ld + r]) : -C(G.b64[k * G.ld + r]);
        return;
    }
    const bool obs = !(G.heldout && G.heldout[r]);
#pragma unroll
    for (int k = 0; k < K; k++) q[k] = C(0);
    if (!obs) return;
    for (int c = 0; c < G.xdim; c++) {
        const C xc = is_f32 ? C(reinterpret_cast<const float *>(xv)[G.xoff + c * G.ld + r])
                            : C(reinterpret_cast<const double *>(xv)[G.xoff + c * G.ld + r]);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const C dx = xc - C(G.theta[k * G.xdim + c]);
            q[k] += dx * dx;
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) q[k] *= C(G.bscale);
}

template <typename S, typename C, int K, int NT, bool IMG>
__global__ __launch_bounds__(NT) void spg_pass_grid_dir(GridArgs A) {
    __shared__ double lds[NPART * (NT / WAVE)];
    const int total = A.grp.off[A.grp.n];
    const bool is_f32 = sizeof(S) == 4;
    for (int j = blockIdx.x; j < total; j += gridDim.x) {
        int inst, slot;
        find_tile(A.grp, xcd_tile(j, total), inst, slot);
        // a stopped single problem keeps its tiles (no scan runs): they are skipped
        if (uni(A.state[inst].active) == 0) continue;
        const InstDesc &D = A.desc[inst];
        const InstState &st = A.state[inst];
        // per-slot fields in scalar registers (read once, not after every store)
        const int64_t N = uni(D.T);  // nodes T*R
        const int64_t ld = uni(D.ld), goff = uni(D.goff);
        const int msub = uni(D.msub);
        const int par = uni(st.parity);
        const bool first = uni(st.first) != 0;
        const C beta = C(uni(st.beta_apply));
        const C alpha = C(uni(D.alpha)), a_dir = C(uni(st.alpha_dir)), s = C(uni(D.s_kkt));
        const GridData gd = grid_data(D);
        // the KKT residual on every pass: these kernels are bound by their neighbour loads, the
        // extra projection costs nothing measurable, and the solve stops at the first pass that
        // reaches the target (the residual is not monotone along SPG)
        double acc[NPART] = {0.0, 0.0, 0.0, 0.0, 0.0};
      for (int m = 0; m < msub; m++) {
        const int64_t r = (int64_t)(slot * msub + m) * NT + threadIdx.x;
        if ((int64_t)(slot * msub + m) * NT >= N) break;
        const S *src = reinterpret_cast<const S *>(par ? A.gamma1 : A.gamma0) + goff;
        S *dst = reinterpret_cast<S *>(par ? A.gamma0 : A.gamma1) + goff;
        S *dr = reinterpret_cast<S *>(par ? A.dir0 : A.dir1) + goff;
        if (r < N) {
            C g0[K];
            C lap[K];
#pragma unroll
            for (int k = 0; k < K; k++) lap[k] = C(0);
            if constexpr (IMG) {
                // The walk's order (up, left, right, down) without branches: a missing
                // neighbour reads the node itself and its term is dropped by a select, so the
                // sums are bitwise the general walk's and all loads are issued before any use.
                const uint32_t ru = (uint32_t)r, W = (uint32_t)A.geo.width, col = ru % W;
                const bool has[4] = {ru >= W, col > 0, col < W - 1, ru + W < (uint32_t)A.geo.R};
                const int64_t nb[4] = {has[0] ? r - (int64_t)W : r, has[1] ? r - 1 : r,
                                       has[2] ? r + 1 : r, has[3] ? r + (int64_t)W : r};
                const S *dprev = reinterpret_cast<const S *>(par ? A.dir1 : A.dir0) + goff;
                // one branch-free block per case, so that the five nodes' loads issue together
                auto gather = [&](auto first_case) {
                    constexpr bool FIRST = decltype(first_case)::value;
                    C gn[4][K];
                    grid_gamma_u<S, C, K, FIRST>(src, dprev, beta, ld, r, g0);
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        grid_gamma_u<S, C, K, FIRST>(src, dprev, beta, ld, nb[q], gn[q]);
#pragma unroll
                    for (int q = 0; q < 4; q++)
#pragma unroll
                        for (int k = 0; k < K; k++)
                            lap[k] = has[q] ? lap[k] + C(1.0) * (g0[k] - gn[q][k]) : lap[k];
                };
                if (first)
                    gather(std::true_type{});
                else
                    gather(std::false_type{});
            } else {
                grid_gamma_at<S, C, K>(A, D, st, src, r, g0);
                grid_neighbours<IMG>(A.geo, r, [&](int64_t n2, double w) {
                    C gn[K];
                    grid_gamma_at<S, C, K>(A, D, st, src, n2, gn);
#pragma unroll
                    for (int k = 0; k < K; k++) lap[k] += C(w) * (g0[k] - gn[k]);
                });
            }
            C q[K];
            grid_dataterm_u<C, K>(gd, A.x, r, is_f32, q);
            C y[K], p[K], z[K], g[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                g[k] = alpha * lap[k] + q[k];
                y[k] = g0[k] - a_dir * g[k];
                p[k] = y[k];
                z[k] = g0[k] - s * g[k];
            }
            const C tau = project_simplex<K, C>(p);
            project_simplex<K, C>(z);
            double ddt = 0.0, off = 0.0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const C d = p[k] - g0[k];
                if (k < stored_rows<K>()) {
                    dst[k * ld + r] = S(g0[k]);
                    dr[k * ld + r] = S(d);
                }
                ddt += (double)d * (double)d;
                if (y[k] <= tau) off += (double)(tau - y[k]) * (double)g0[k];
                acc[3] += 0.5 * ((double)g[k] + (double)q[k]) * (double)g0[k];
                const double e = fabs((double)g0[k] - (double)z[k]);
                acc[4] = (e > acc[4] || e != e) ? e : acc[4];
            }
            // <g, d> in the sign-safe form of the projected-gradient step (see pass_tile_1d)
            acc[2] -= (ddt + off) / (double)a_dir;
        }
      }
        block_reduce_part<NT>(acc, lds);
        if (threadIdx.x == 0) {
            double *o = A.part + (int64_t)(D.tile0 + slot) * NPART;
            o[2] = acc[2];
            o[3] = acc[3];
            o[4] = acc[4];
        }
    }
}

// d_i at node n: K - 1 stored rows, the last rebuilt (FP64)
template <typename S, int K>
__device__ __forceinline__ void load_dir(const S *dr, int64_t ld, int64_t n, double (&d)[K]) {
#pragma unroll
    for (int k = 0; k < stored_rows<K>(); k++) d[k] = (double)dr[k * ld + n];
    if (stored_rows<K>() < K) d[K - 1] = last_dir<double, K>(d);
}

template <typename S, typename C, int K, int NT, bool IMG>
__global__ __launch_bounds__(NT) void spg_pass_grid_ad(GridArgs A) {
    __shared__ double lds[NPART * (NT / WAVE)];
    const int total = A.grp.off[A.grp.n];
    for (int j = blockIdx.x; j < total; j += gridDim.x) {
        int inst, slot;
        find_tile(A.grp, xcd_tile(j, total), inst, slot);
        if (uni(A.state[inst].active) == 0) continue;
        const InstDesc &D = A.desc[inst];
        const int64_t N = uni(D.T);
        const int64_t ld = uni(D.ld);
        const int msub = uni(D.msub);
        const double alpha = uni(D.alpha);
        // d_i was written by spg_pass_grid_dir into the buffer opposite to the parity
        const S *dr = reinterpret_cast<const S *>(uni(A.state[inst].parity) ? A.dir0 : A.dir1) +
                      uni(D.goff);
        double acc[NPART] = {0.0, 0.0, 0.0, 0.0, 0.0};
      for (int m = 0; m < msub; m++) {
        const int64_t r = (int64_t)(slot * msub + m) * NT + threadIdx.x;
        if ((int64_t)(slot * msub + m) * NT >= N) break;
        if (r < N) {
            if constexpr (IMG) {
                // the forward edges (right, down) without branches, as in the direction kernel
                const uint32_t ru = (uint32_t)r, W = (uint32_t)A.geo.width, col = ru % W;
                const bool hr = col < W - 1, hd = ru + W < (uint32_t)A.geo.R;
                const int64_t nr = hr ? r + 1 : r, nd = hd ? r + (int64_t)W : r;
                const double w = 1.0;
                double d0[K], d1[K], d2[K];
                load_dir<S, K>(dr, ld, r, d0);
                load_dir<S, K>(dr, ld, nr, d1);
                load_dir<S, K>(dr, ld, nd, d2);
                // the general walk's order: every component of one edge, then the next edge
#pragma unroll
                for (int k = 0; k < K; k++) acc[0] += d0[k] * d0[k];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double e1 = d0[k] - d1[k];
                    acc[1] = hr ? acc[1] + alpha * w * e1 * e1 : acc[1];
                }
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double e2 = d0[k] - d2[k];
                    acc[1] = hd ? acc[1] + alpha * w * e2 * e2 : acc[1];
                }
            } else {
                double d0[K];
                load_dir<S, K>(dr, ld, r, d0);
#pragma unroll
                for (int k = 0; k < K; k++) acc[0] += d0[k] * d0[k];
                // d'A d = alpha * sum over the graph's edges of w (d_a - d_b)^2, each edge once
                grid_forward_edges<IMG>(A.geo, r, [&](int64_t n2, double w) {
                    double d1[K];
                    load_dir<S, K>(dr, ld, n2, d1);
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const double e = d0[k] - d1[k];
                        acc[1] += alpha * w * e * e;
                    }
                });
            }
        }
      }
        block_reduce_part<NT>(acc, lds);
        if (threadIdx.x == 0) {
            double *o = A.part + (int64_t)(D.tile0 + slot) * NPART;
            o[0] = acc[0];
            o[1] = acc[1];
        }
    }
}

}  // namespace pasc

// Device-side building blocks shared by the HIP translation units of libpasc_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pasc_internal.h"

namespace pasc {

// Workgroup barrier for LDS traffic only: waits for this wave's LDS operations and then for the
// other waves, but not for its outstanding global loads and stores (the fence that
// __syncthreads() adds would drain those too, including an asynchronous global -> LDS copy of
// the next tile that is meant to stay in flight across the barrier).
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

constexpr int WAVE = 64;

template <typename S> struct V16;
template <> struct V16<float> {
    using type = float4;
    static constexpr int n = 4;
    __device__ static inline float get(const type &v, int i) {
        return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
    }
    __device__ static inline void set(type &v, int i, float a) {
        if (i == 0) v.x = a; else if (i == 1) v.y = a; else if (i == 2) v.z = a; else v.w = a;
    }
};
template <> struct V16<double> {
    using type = double2;
    static constexpr int n = 2;
    __device__ static inline double get(const type &v, int i) { return i == 0 ? v.x : v.y; }
    __device__ static inline void set(type &v, int i, double a) {
        if (i == 0) v.x = a; else v.y = a;
    }
};

// Nontemporal 16-byte stores (global_store_dwordx4 with the nt bit).  An iterate written by
// a pass is read again only by the next pass, after the whole vector has streamed through.
// Measured on one MI355X: a probe stream of three read and two write vectors of 80 MiB moved
// 5.4 TB/s with plain stores and 8.0 TB/s with nontemporal stores; in the FP64 pass they took
// the pass from 95.8 to 87.1 us at T = 1e7, K = 3 (the FP32 pass was slower with them).
typedef float nt_float4 __attribute__((ext_vector_type(4)));
typedef double nt_double2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_nt(float4 *p, const float4 &v) {
    const nt_float4 n = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(n, reinterpret_cast<nt_float4 *>(p));
}
__device__ __forceinline__ void store_nt(double2 *p, const double2 &v) {
    const nt_double2 n = {v.x, v.y};
    __builtin_nontemporal_store(n, reinterpret_cast<nt_double2 *>(p));
}

// max via the hardware min/max instructions (one VALU op instead of a compare and a select)
__device__ inline float cmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ inline double cmax(double a, double b) { return __builtin_fmax(a, b); }

template <typename C, int K>
__device__ __forceinline__ C recip_small(int n) {
    // 1/n for n in [1, K] without a division (select chain over compile-time constants)
    C r = C(1);
#pragma unroll
    for (int m = 2; m <= K; m++) r = (n == m) ? C(1) / C(m) : r;
    return r;
}

// Threshold of the Euclidean projection onto the unit simplex {y : sum y = 1, y >= 0}, for
// K <= 4: tau* = max over all non-empty subsets S of (sum_S y - 1)/|S|.  (With u sorted
// descending, h(j) = (u_1 + .. + u_j - 1)/j is unimodal and peaks at the support size, so its
// maximum is the Chen & Ye / Duchi threshold; any other subset of size j has a smaller sum than
// the top-j prefix.)  Branch-free: no sort and no data-dependent loop.
template <int K, typename C>
__device__ __forceinline__ C simplex_tau_small(const C (&y)[K]) {
    static_assert(K >= 1 && K <= 4, "closed form for K <= 4");
    const C third = C(1) / C(3);
    if constexpr (K == 1) {
        return y[0] - C(1);
    } else if constexpr (K == 2) {
        const C t = cmax(y[0], y[1]) - C(1);
        return cmax(t, C(0.5) * (y[0] + y[1] - C(1)));
    } else if constexpr (K == 3) {
        C t = cmax(cmax(y[0], y[1]), y[2]) - C(1);
        const C s01 = y[0] + y[1], s02 = y[0] + y[2], s12 = y[1] + y[2];
        t = cmax(t, C(0.5) * (cmax(cmax(s01, s02), s12) - C(1)));
        return cmax(t, third * (s01 + y[2] - C(1)));
    } else {
        C t = cmax(cmax(y[0], y[1]), cmax(y[2], y[3])) - C(1);
        const C s01 = y[0] + y[1], s02 = y[0] + y[2], s03 = y[0] + y[3];
        const C s12 = y[1] + y[2], s13 = y[1] + y[3], s23 = y[2] + y[3];
        const C p2 = cmax(cmax(cmax(s01, s02), cmax(s03, s12)), cmax(s13, s23));
        t = cmax(t, C(0.5) * (p2 - C(1)));
        const C s012 = s01 + y[2], s013 = s01 + y[3], s023 = s02 + y[3], s123 = s12 + y[3];
        const C p3 = cmax(cmax(s012, s013), cmax(s023, s123));
        t = cmax(t, third * (p3 - C(1)));
        return cmax(t, C(0.25) * (s012 + y[3] - C(1)));
    }
}

// Euclidean projection of one K-vector (in registers) onto the unit simplex; replaces
// SimplexFeasibleSet_Local::project_sub (simplex_local.h:176-249) with a sort-free threshold:
//   * K <= 4: the closed form above;
//   * K > 4: tau is bracketed in [max(y) - 1, max(y)) (phi(tau) = sum max(y - tau, 0) - 1 changes
//     sign there), NB bisection steps shrink the bracket, and the exact threshold is the fixed
//     point tau = (sum_{y > tau} y - 1) / |{y > tau}| iterated from the lower bracket end (the
//     support only shrinks, so at most K sweeps), the same threshold as Chen & Ye's search.
// Returns tau; y is overwritten with max(y - tau, 0).
template <int K, typename C>
__device__ __forceinline__ C project_simplex(C (&y)[K]) {
    if constexpr (K <= 4) {
        const C tau = simplex_tau_small<K, C>(y);
#pragma unroll
        for (int k = 0; k < K; k++) y[k] = cmax(y[k] - tau, C(0));
        return tau;
    } else {
        C mx = y[0];
#pragma unroll
        for (int k = 1; k < K; k++) mx = cmax(mx, y[k]);
        C tau = mx - C(1);
        constexpr int NB = 4;
        C hi = mx;
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const C mid = C(0.5) * (tau + hi);
            C phi = C(-1);
#pragma unroll
            for (int k = 0; k < K; k++) phi += cmax(y[k] - mid, C(0));
            if (phi >= C(0)) tau = mid; else hi = mid;
        }
        int cnt_prev = -1;
#pragma unroll
        for (int it = 0; it < K; it++) {
            C sum = C(0);
            int cnt = 0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const bool in = y[k] > tau;
                sum += in ? y[k] : C(0);
                cnt += in ? 1 : 0;
            }
            if (cnt == cnt_prev || cnt == 0) break;
            cnt_prev = cnt;
            tau = (sum - C(1)) * recip_small<C, K>(cnt);
        }
#pragma unroll
        for (int k = 0; k < K; k++) y[k] = cmax(y[k] - tau, C(0));
        return tau;
    }
}

// Wavefront reductions (64 lanes).
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, WAVE);
        v = (w > v || w != w) ? w : v;
    }
    return v;
}
__device__ inline int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// Block reduction of NPART partials (sum for the first 4, max for the last).  Thread 0 returns
// the result in v.  lds must hold NPART * (NT/64) doubles.
template <int NT>
__device__ inline void block_reduce_part(double (&v)[NPART], double *lds) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
#pragma unroll
    for (int i = 0; i < NPART - 1; i++) v[i] = wave_sum(v[i]);
    v[NPART - 1] = wave_max(v[NPART - 1]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NPART; i++) lds[w * NPART + i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < NT / WAVE; q++) {
#pragma unroll
            for (int i = 0; i < NPART - 1; i++) v[i] += lds[q * NPART + i];
            const double m = lds[q * NPART + NPART - 1];
            if (m > v[NPART - 1] || m != m) v[NPART - 1] = m;
        }
    }
    __syncthreads();
}

// Tile enumeration over the active instances of a K group (see engine scan kernel).
struct GroupView {
    const int32_t *ids;
    const int32_t *off;
    const int32_t *hint;
    int32_t n;
};

__device__ inline void find_tile(const GroupView &g, int j, int &inst, int &chunk) {
    int a = g.hint[j >> 6];
    while (g.off[a + 1] <= j) a++;
    inst = g.ids[a];
    chunk = j - g.off[a];
}

// Neighbour walks of the spatio-temporal graph (GraphGeo).  A0 is the weighted Laplacian with
// edge weights 2 for (t,r)-(t+-1,r) and 1 for (t,r)-(t,s) and (t,r)-(t+-1,s), s a spatial
// neighbour of r: its diagonal is Wsum = deg (T == 1), 2 deg + 2 (t at an end), 3 deg + 4
// (interior t), as blockgraphsparse.cpp:57-73 assembles it.
template <typename F>
__device__ __forceinline__ void for_spatial(const GraphGeo &G, int64_t r, F &&f) {
    if (G.ptr) {
        for (int64_t e = G.ptr[r]; e < G.ptr[r + 1]; e++) f((int64_t)G.idx[e]);
    } else if (G.width > 0) {
        const int64_t W = G.width, row = r / W, col = r - row * W;
        if (row > 0) f(r - W);
        if (col > 0) f(r - 1);
        if (col < W - 1) f(r + 1);
        if (r + W < G.R) f(r + W);
    }
}

// every neighbour n2 of node n with its edge weight w
template <typename F>
__device__ __forceinline__ void for_neighbours(const GraphGeo &G, int64_t n, F &&f) {
    const int64_t t = n / G.R, r = n - t * G.R;
    if (t > 0) f(n - G.R, 2.0);
    if (t < G.T - 1) f(n + G.R, 2.0);
    for_spatial(G, r, [&](int64_t s) {
        f(t * G.R + s, 1.0);
        if (t > 0) f((t - 1) * G.R + s, 1.0);
        if (t < G.T - 1) f((t + 1) * G.R + s, 1.0);
    });
}

// the edges of node n to nodes of larger index: every undirected edge once
template <typename F>
__device__ __forceinline__ void for_forward_edges(const GraphGeo &G, int64_t n, F &&f) {
    const int64_t t = n / G.R, r = n - t * G.R;
    if (t < G.T - 1) f(n + G.R, 2.0);
    for_spatial(G, r, [&](int64_t s) {
        if (s > r) f(t * G.R + s, 1.0);
        if (t < G.T - 1) f((t + 1) * G.R + s, 1.0);
    });
}

// Wave-uniform copies in scalar registers (the value of the first active lane).  Values that
// every lane already holds equal become scalars, so later uses wait on nothing.
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni(int64_t v) {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)(u >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double uni(double v) {
    return __builtin_bit_cast(double, uni(__builtin_bit_cast(int64_t, v)));
}
__device__ __forceinline__ float uni(float v) {
    return __builtin_bit_cast(float, uni(__builtin_bit_cast(int32_t, v)));
}
template <typename P>
__device__ __forceinline__ P *uni(P *v) {
    return reinterpret_cast<P *>(uni((int64_t)reinterpret_cast<uintptr_t>(v)));
}

// Hand-over of a few values between workgroups within one launch (the partial sums of the
// fused step), without an agent-scope fence (on gfx950 a release fence writes back every dirty
// line of the XCD's L2, here the gamma rows in flight).  Each value location carries its own
// state: it holds PART_EMPTY (a NaN whose payload no arithmetic produces) until its writer
// publishes the value with an agent-scope atomic store; the reader, which learned from a
// counter that the writer is done, reads the location with agent-scope atomic loads until it
// no longer holds PART_EMPTY (settle_published), and puts PART_EMPTY back for the next launch
// (reset_published).
// Correctness relies only on the coherence of each single location, not on an order between
// two locations (the value and the counter), so the writer counts without waiting for its
// stores.  The reader's wait is bounded: a value that does not appear reads as a quiet NaN,
// which the step reports as PASC_ERR_NUMERIC.
constexpr uint64_t PART_EMPTY = 0x7ff4504153430000ull;  // NaN, payload "PASC"
constexpr int PUBLISH_SPIN_LIMIT = 1 << 16;

__device__ __forceinline__ void store_coherent(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_coherent(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// v: a load_coherent of p issued earlier (the loads of a slot go out together); returns the
// published value
__device__ __forceinline__ double settle_published(const double *p, double v) {
    for (int n = 0; __builtin_bit_cast(uint64_t, v) == PART_EMPTY; n++) {
        if (n >= PUBLISH_SPIN_LIMIT) return __builtin_nan("");
        __builtin_amdgcn_s_sleep(1);
        v = load_coherent(p);
    }
    return v;
}
// marks a location empty again once its value was taken; a plain store, issued after the
// reader's own publication so that no wait covers it (the launch boundary makes it visible to
// the next launch's writer and reader)
__device__ __forceinline__ void reset_published(double *p) {
    *p = __builtin_bit_cast(double, PART_EMPTY);
}

}  // namespace pasc

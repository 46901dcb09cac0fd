// Device-side building blocks shared by the level kernels (tt_kernels.hip)
// and the persistent batched kernel (tt_batch.hip): one (node, pattern) step
// of each pass, so both launch shapes run identical arithmetic.
#pragma once
#include "tt_common.h"

#include <float.h>

namespace ttd {

__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

template <int N>
__device__ __forceinline__ void load_row(const float *__restrict__ src, double (&r)[N]) {
    if constexpr (N % 4 == 0) {
        const float4 *s4 = reinterpret_cast<const float4 *>(src);
#pragma unroll
        for (int q = 0; q < N / 4; ++q) {
            float4 x = s4[q];
            r[4 * q + 0] = x.x;
            r[4 * q + 1] = x.y;
            r[4 * q + 2] = x.z;
            r[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) r[i] = src[i];
    }
}

template <int N>
__device__ __forceinline__ void store_row(float *__restrict__ dst, const float (&r)[N]) {
    if constexpr (N % 4 == 0) {
        float4 *d4 = reinterpret_cast<float4 *>(dst);
#pragma unroll
        for (int q = 0; q < N / 4; ++q) d4[q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) dst[i] = r[i];
    }
}

template <int N>
__device__ __forceinline__ void store_bytes(uint8_t *__restrict__ dst, const int (&c)[N]) {
    if constexpr (N == 4) {
        uint32_t w = (uint32_t)c[0] | ((uint32_t)c[1] << 8) | ((uint32_t)c[2] << 16) | ((uint32_t)c[3] << 24);
        *reinterpret_cast<uint32_t *>(dst) = w;
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) dst[i] = (uint8_t)c[i];
    }
}

// Child message m_j = sum_i r_i P[i][j] (GTR.propagate_profile, gtr.py:992-995).
template <int N>
__device__ __forceinline__ void child_message(const double (&r)[N], const double *__restrict__ M, double (&m)[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) m[j] = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int j = 0; j < N; ++j) m[j] = fma(r[i], M[i * N + j], m[j]);
    }
}

// Subtree row of a child: stored row of an internal child, profile-map row
// of a tip's character (seq2prof, seq_utils.py:207-229).
template <int N>
__device__ __forceinline__ void child_row(const TreeDev &T, int cref, int p, double (&r)[N]) {
    if (cref < 0) {
        const int code = T.tip_codes[(size_t)(-cref - 1) * T.P + p];
        const float *lr = T.lut + code * N;
#pragma unroll
        for (int i = 0; i < N; ++i) r[i] = lr[i];
    } else {
        load_row<N>(T.sub + ((size_t)cref * T.P + p) * N, r);
    }
}

// Postorder step for internal row v and pattern p (treeanc.py:857-878): the
// product over children of (child row @ P(t)), renormalised so the largest
// entry is 1.  Returns the log normaliser (reference: normalize_profile
// offset, seq_utils.py:296-307).
template <int N>
__device__ __forceinline__ double post_node(const TreeDev &T, int v, int p) {
    double prod[N];
#pragma unroll
    for (int j = 0; j < N; ++j) prod[j] = 1.0;
    double lscale = 0.0;
    const int e0 = T.child_ptr[v], e1 = T.child_ptr[v + 1];
    for (int e = e0; e < e1; ++e) {
        const int cref = T.child_ref[e];
        const double *__restrict__ M = T.P64 + (size_t)T.child_node[e] * (N * N);
        double r[N], m[N];
        child_row<N>(T, cref, p, r);
        child_message<N>(r, M, m);
#pragma unroll
        for (int j = 0; j < N; ++j) prod[j] *= m[j];
        if (e1 - e0 > 2) {  // polytomies: keep the running product away from underflow
            double mx = prod[0];
#pragma unroll
            for (int j = 1; j < N; ++j) mx = fmax(mx, prod[j]);
            if (mx < 0x1p-500 && mx > 0.0) {
#pragma unroll
                for (int j = 0; j < N; ++j) prod[j] *= 0x1p+500;
                lscale -= 500.0 * 0.6931471805599453;
            }
        }
    }
    double mx = prod[0];
    int jm = 0;
#pragma unroll
    for (int j = 1; j < N; ++j)
        if (prod[j] > mx) {
            mx = prod[j];
            jm = j;
        }
    const double inv = 1.0 / mx;
    float out[N];
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = (j == jm) ? 1.0f : (float)(prod[j] * inv);
    store_row<N>(T.sub + ((size_t)v * T.P + p) * N, out);
    return log(mx) + lscale;
}

// Root (treeanc.py:817-828) for pattern p: profile = normalise(pi * subtree),
// logL = log(sum pi * subtree) + ``offsets`` (the sum of all node log
// normalisers of this pattern).  Writes sequence_LH, the root profile and
// state; returns mult[p] * logL.
template <int N>
__device__ __forceinline__ double root_step(const TreeDev &T, int p, double offsets) {
    double r[N], x[N];
    load_row<N>(T.sub + ((size_t)T.root_row * T.P + p) * N, r);
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        x[j] = T.pi[j] * r[j];
        s += x[j];
    }
    const double lp = log(s) + offsets;
    T.seqlh[p] = lp;
    const double inv = 1.0 / s;
    float out[N];
    int jm = 0;
    double mx = x[0];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        out[j] = (float)(x[j] * inv);
        if (x[j] > mx) {
            mx = x[j];
            jm = j;
        }
    }
    store_row<N>(T.prof + ((size_t)T.root * T.P + p) * N, out);
    T.states[(size_t)T.root * T.P + p] = (uint8_t)jm;
    return T.mult[p] * lp;
}

// Preorder step for node c (treeanc.py:889-930): outgroup message by the
// reference's division trick, normalise(log max(TINY, parent profile) -
// log(child message)), then evolve it down the branch and multiply with the
// subtree row.
template <int N>
__device__ __forceinline__ void pre_node(const TreeDev &T, int c, int p) {
    const int par = T.parent[c];
    const int ir = T.int_row[c];
    const double *__restrict__ M = T.P64 + (size_t)c * (N * N);
    double pp[N], sub[N], m[N], q[N], msg[N];
    {
        const float *src = T.prof + ((size_t)par * T.P + p) * N;
        if constexpr (N % 4 == 0) {
            const float4 *s4 = reinterpret_cast<const float4 *>(src);
#pragma unroll
            for (int u = 0; u < N / 4; ++u) {
                float4 x = s4[u];
                pp[4 * u] = x.x;
                pp[4 * u + 1] = x.y;
                pp[4 * u + 2] = x.z;
                pp[4 * u + 3] = x.w;
            }
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) pp[i] = src[i];
        }
    }
    child_row<N>(T, ir >= 0 ? ir : -(T.tip_row[c] + 1), p, sub);
    child_message<N>(sub, M, m);
    double qs = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        q[j] = fmax(TT_TINY, pp[j]) / m[j];
        qs += q[j];
    }
    const double qi = 1.0 / qs;
#pragma unroll
    for (int j = 0; j < N; ++j) q[j] *= qi;
    // GTR.evolve (gtr.py:1023-1025): msg_i = sum_j outgroup_j P[i][j]
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double a = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) a = fma(q[j], M[i * N + j], a);
        msg[i] = a * sub[i];
    }
    double xs = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) xs += msg[i];
    const double xi = 1.0 / xs;
    float out[N];
    int jm = 0;
    double mx = msg[0];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        out[i] = (float)(msg[i] * xi);
        if (msg[i] > mx) {
            mx = msg[i];
            jm = i;
        }
    }
    store_row<N>(T.prof + ((size_t)c * T.P + p) * N, out);
    T.states[(size_t)c * T.P + p] = (uint8_t)jm;
}

// Deterministic block reduction of one double per thread (wave shuffles,
// then LDS across the block's waves).  Result valid in thread 0.
__device__ __forceinline__ double block_sum(double x) {
    __shared__ double s_w[TT_BLOCK / TT_WAVE];
#pragma unroll
    for (int off = TT_WAVE / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, TT_WAVE);
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    if (lane == 0) s_w[w] = x;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + TT_WAVE - 1) / TT_WAVE;
        for (int q = 0; q < nw; ++q) r += s_w[q];
    }
    return r;
}

}  // namespace ttd

// Device-side building blocks shared by the level kernels (tt_kernels.hip)
// and the persistent batched kernel (tt_batch.hip): one (node, pattern) step
// of each pass, so both launch shapes run identical arithmetic.
//
// Memory-level parallelism: a wave handles two nodes per step; for binary
// nodes every child load of both nodes is issued before any arithmetic
// (fetch / finish split), so one wave keeps four row loads in flight.
#pragma once
#include "tt_common.h"

#include <float.h>

#include <type_traits>

namespace ttd {

__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

template <int N, class Ptr>
__device__ __forceinline__ void load_rowf(Ptr src, float (&r)[N]) {
    if constexpr (N % 4 == 0) {
        const float4 TT_G *s4 = reinterpret_cast<const float4 TT_G *>(src);
#pragma unroll
        for (int q = 0; q < N / 4; ++q) {
            const float4 x = s4[q];
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

template <int N, class Ptr>
__device__ __forceinline__ void load_row(Ptr src, double (&r)[N]) {
    float f[N];
    load_rowf<N>(src, f);
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] = f[i];
}

template <int N, class Ptr>
__device__ __forceinline__ void store_row(Ptr dst, const float (&r)[N]) {
    if constexpr (N % 4 == 0) {
        float4 TT_G *d4 = reinterpret_cast<float4 TT_G *>(dst);
#pragma unroll
        for (int q = 0; q < N / 4; ++q) d4[q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) dst[i] = r[i];
    }
}

template <int N, class Ptr>
__device__ __forceinline__ void store_bytes(Ptr dst, const int (&c)[N]) {
    if constexpr (N == 4) {
        uint32_t w = (uint32_t)c[0] | ((uint32_t)c[1] << 8) | ((uint32_t)c[2] << 16) | ((uint32_t)c[3] << 24);
        *reinterpret_cast<uint32_t TT_G *>(dst) = w;
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) dst[i] = (uint8_t)c[i];
    }
}

// Joint argmax rows (Cx), layout in tt_common.h (cx_packed).
template <int N>
__device__ __forceinline__ void cx_store(const TreeDev &T, int node, int p, const int (&c)[N]) {
    if constexpr (cx_packed(N)) {
        uint32_t b = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) b |= (uint32_t)c[j] << (2 * j);
        T.Cx[(size_t)node * T.P + p] = (uint8_t)b;
    } else {
        store_bytes<N>(T.Cx + ((size_t)node * T.P + p) * N, c);
    }
}

// State of node under parent state ps from its argmax row.
__device__ __forceinline__ int cx_state(const TreeDev &T, int node, int p, int ps) {
    if (cx_packed(T.n)) return (T.Cx[(size_t)node * T.P + p] >> (2 * ps)) & 3;
    return T.Cx[((size_t)node * T.P + p) * T.n + ps];
}

// Child message m_j = sum_i r_i P[i][j] (GTR.propagate_profile, gtr.py:992-995).
template <int N, class Ptr>
__device__ __forceinline__ void child_message(const double (&r)[N], Ptr M, double (&m)[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) m[j] = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int j = 0; j < N; ++j) m[j] = fma(r[i], M[i * N + j], m[j]);
    }
}

// ---- postorder: FP32 arithmetic with deficit messages ---------------------
//
// For child c and state j the message is m_cj = r_cj + d_cj with
// d_cj = sum_i r_ci (P - I)_ij, where P - I is built in FP64 and rounded once
// (D32).  The deficit d is small on short branches and keeps its relative
// precision in FP32; r + d is then formed exactly in FP64, so the per-node
// normaliser of a nearly constant pattern is accurate to ~1e-7 of its own
// (small) distance from 1 instead of 6e-8 absolute.  Normalisers multiply
// into an FP64 running product (LogAcc).

template <int N, class MP>
__device__ __forceinline__ void deficit(const float (&r)[N], MP D, float (&d)[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) d[j] = 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int j = 0; j < N; ++j) d[j] = fmaf(r[i], D[i * N + j], d[j]);
    }
}

// Running product of per-node normalisers, value = m * 2^e, FP64.  A node
// contributes prod_c (r_c + d_c) at its selected state, each factor formed
// exactly in FP64 from the FP32 row entry r and deficit d; one FP64 log per
// pattern at the end replaces a log per (node, child, pattern).
struct LogAcc {
    double m = 1.0;
    int e = 0;
    __device__ __forceinline__ void mul(double x) {
        m *= x;
        if (m < 0x1p-500) {
            m *= 0x1p+500;
            e -= 500;
        } else if (m > 0x1p+500) {
            m *= 0x1p-500;
            e += 500;
        }
    }
    __device__ __forceinline__ double log_value() const { return log(m) + (double)e * 0.6931471805599453094; }
};

__device__ __forceinline__ double msg64(float r, float d) { return (double)r + (double)d; }

// log of an FP32 message for the per-state sums of the polytomy path (those
// only shape the stored row; the normaliser itself goes through LogAcc)
__device__ __forceinline__ float log_msg(float r, float d) { return r == 1.0f ? log1pf(d) : logf(r + d); }

template <int N>
__device__ __forceinline__ int argmax_first(const float (&x)[N]) {
    int jm = 0;
    float mx = x[0];
#pragma unroll
    for (int j = 1; j < N; ++j)
        if (x[j] > mx) {
            mx = x[j];
            jm = j;
        }
    return jm;
}

// Store row v normalised by its entry jm (value xm), which is stored as
// exactly 1.
// lds_out (nullable): the same row also goes to this LDS slot.
template <int N>
__device__ __forceinline__ void store_post(const TreeDev &T, int v, int p, const float (&x)[N], int jm, float xm,
                                           float *lds_out = nullptr) {
    // correctly rounded quotient: these entries feed the parents' messages,
    // and a biased approximate reciprocal would bias variable patterns' logL
    float out[N];
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = (j == jm) ? 1.0f : x[j] / xm;
    store_row<N>(T.sub + ((size_t)v * T.P + p) * N, out);
    if (lds_out) {
#pragma unroll
        for (int j = 0; j < N; ++j) lds_out[j] = out[j];
    }
}

// Subtree row of child edge e: stored row of an internal child or the
// profile-map row of a tip's character (seq2prof, seq_utils.py:207-229).
template <int N, class LP>
__device__ __forceinline__ void child_row(const TreeDev &T, int cref, int p, LP lut, float (&r)[N]) {
    if (cref < 0) {
        const int code = T.tip_codes[(size_t)(-cref - 1) * T.P + p];
#pragma unroll
        for (int i = 0; i < N; ++i) r[i] = lut[code * N + i];
    } else {
        load_rowf<N>(T.sub + ((size_t)cref * T.P + p) * N, r);
    }
}

// Postorder step for internal row v with any number of children
// (treeanc.py:857-878), in log space per state as the reference does.
// post_node with the child rows read through rowf(cref, r) (tips and
// internal children), uni() applied to the tree-structure values (wave
// uniform in the level kernels, per lane elsewhere) and the stored row also
// copied to lds_out when given.
template <int N, class LP, class RowF, class Uni>
__device__ __forceinline__ void post_node_f(const TreeDev &T, int v, int p, LP lut, LogAcc &acc, RowF rowf, Uni uni,
                                            float *lds_out = nullptr) {
    float L[N];
#pragma unroll
    for (int j = 0; j < N; ++j) L[j] = 0.0f;
    const int e0 = uni(T.child_ptr[v]), e1 = uni(T.child_ptr[v + 1]);
    for (int e = e0; e < e1; ++e) {
        const int cref = uni(T.child_ref[e]);
        const int cnode = uni(T.child_node[e]);
        float r[N], d[N];
        rowf(cref, r);
        deficit<N>(r, T.D32 + (size_t)cnode * (N * N), d);
#pragma unroll
        for (int j = 0; j < N; ++j) L[j] += log_msg(r[j], d[j]);
    }
    // first maximum, value carried along (no dynamic register indexing)
    int jm = 0;
    float Lm = L[0];
#pragma unroll
    for (int j = 1; j < N; ++j)
        if (L[j] > Lm) {
            Lm = L[j];
            jm = j;
        }
    float out[N];
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = (j == jm) ? 1.0f : expf(L[j] - Lm);
    // normaliser at the selected state, children revisited (cache hits)
    double lnode = 0.0;
    for (int e = e0; e < e1; ++e) {
        const int cref = uni(T.child_ref[e]);
        const int cnode = uni(T.child_node[e]);
        float r[N], d[N];
        rowf(cref, r);
        deficit<N>(r, T.D32 + (size_t)cnode * (N * N), d);
        float rj = r[0], dj = d[0];
#pragma unroll
        for (int j = 1; j < N; ++j)
            if (j == jm) {
                rj = r[j];
                dj = d[j];
            }
        acc.mul(msg64(rj, dj));
        if (T.offs) lnode += log(msg64(rj, dj));
    }
    if (T.offs) T.offs[(size_t)v * T.P + p] = lnode;
    store_row<N>(T.sub + ((size_t)v * T.P + p) * N, out);
    if (lds_out) {
#pragma unroll
        for (int j = 0; j < N; ++j) lds_out[j] = out[j];
    }
}

template <int N, class LP>
__device__ __forceinline__ void post_node(const TreeDev &T, int v, int p, LP lut, LogAcc &acc) {
    post_node_f<N>(T, v, p, lut, acc, [&](int cref, float (&r)[N]) { child_row<N>(T, cref, p, lut, r); },
                   [](int e) { return wave_uniform(e); });
}

// Binary fast path, split in a fetch phase (global loads only) and a finish
// phase (arithmetic), so the loads of several nodes overlap.
template <int N>
struct BinRaw {
    float row[2][N];
    int code[2];
};

template <int N>
__device__ __forceinline__ void bin_fetch_refs(const TreeDev &T, int cref0, int cref1, int p, BinRaw<N> &R) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int cref = c ? cref1 : cref0;
        if (cref < 0) {
            R.code[c] = T.tip_codes[(size_t)(-cref - 1) * T.P + p];
        } else {
            R.code[c] = 0;
            load_rowf<N>(T.sub + ((size_t)cref * T.P + p) * N, R.row[c]);
        }
    }
}

template <int N>
__device__ __forceinline__ void bin_fetch(const TreeDev &T, int e0, int p, BinRaw<N> &R) {
    bin_fetch_refs<N>(T, wave_uniform(T.child_ref[e0]), wave_uniform(T.child_ref[e0 + 1]), p, R);
}

// Binary postorder step with the children's P - I given as D0 and D1
// (matrices in memory or in registers).
template <int N, class LP, class MP0, class MP1>
__device__ __forceinline__ void bin_finish_m(const TreeDev &T, int v, int cref0, int cref1, MP0 D0, MP1 D1, int p,
                                             const BinRaw<N> &R, LP lut, LogAcc &acc, float *lds_out = nullptr) {
    float r[2][N], d[2][N], x[N];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int cref = c ? cref1 : cref0;
        if (cref < 0) {
#pragma unroll
            for (int i = 0; i < N; ++i) r[c][i] = lut[R.code[c] * N + i];
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) r[c][i] = R.row[c][i];
        }
        if (c)
            deficit<N>(r[c], D1, d[c]);
        else
            deficit<N>(r[c], D0, d[c]);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) x[j] = (r[0][j] + d[0][j]) * (r[1][j] + d[1][j]);
    // first maximum; the children's (row, deficit) at that state ride along
    int jm = 0;
    float xm = x[0], r0 = r[0][0], d0 = d[0][0], r1 = r[1][0], d1 = d[1][0];
#pragma unroll
    for (int j = 1; j < N; ++j)
        if (x[j] > xm) {
            xm = x[j];
            jm = j;
            r0 = r[0][j];
            d0 = d[0][j];
            r1 = r[1][j];
            d1 = d[1][j];
        }
    const double f = msg64(r0, d0) * msg64(r1, d1);
    acc.mul(f);
    if (T.offs) T.offs[(size_t)v * T.P + p] = log(f);  // update mode: keep the node's own normaliser
    store_post<N>(T, v, p, x, jm, xm, lds_out);
}

template <int N, class LP>
__device__ __forceinline__ void bin_finish_refs(const TreeDev &T, int v, int cref0, int cref1, int cnode0, int cnode1,
                                                int p, const BinRaw<N> &R, LP lut, LogAcc &acc,
                                                float *lds_out = nullptr) {
    bin_finish_m<N>(T, v, cref0, cref1, T.D32 + (size_t)cnode0 * (N * N), T.D32 + (size_t)cnode1 * (N * N), p, R, lut,
                    acc, lds_out);
}

template <int N, class LP>
__device__ __forceinline__ void bin_finish(const TreeDev &T, int v, int e0, int p, const BinRaw<N> &R, LP lut,
                                           LogAcc &acc) {
    bin_finish_refs<N>(T, v, wave_uniform(T.child_ref[e0]), wave_uniform(T.child_ref[e0 + 1]),
                       wave_uniform(T.child_node[e0]), wave_uniform(T.child_node[e0 + 1]), p, R, lut, acc);
}

// A postorder descriptor (kPostDesc ints) read with one scalar load.
struct PostDesc {
    int v, e0, nchild, cref0, cref1, cnode0, cnode1, vn;  // vn: node id of row v
};
__device__ __forceinline__ PostDesc load_post_desc(const int32_t TT_C *d) {
    return PostDesc{d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7]};
}

template <int N>
__device__ __forceinline__ void bin_fetch_d(const TreeDev &T, const PostDesc &d, int p, BinRaw<N> &R) {
    bin_fetch_refs<N>(T, d.cref0, d.cref1, p, R);
}

template <int N, class LP>
__device__ __forceinline__ void bin_finish_d(const TreeDev &T, const PostDesc &d, int p, const BinRaw<N> &R, LP lut,
                                             LogAcc &acc) {
    bin_finish_refs<N>(T, d.v, d.cref0, d.cref1, d.cnode0, d.cnode1, p, R, lut, acc);
}

// One postorder step: the binary path for two children, else post_node.
template <int N, class LP>
__device__ __forceinline__ void post_one(const TreeDev &T, int v, int e0, bool binary, int p, LP lut, LogAcc &acc) {
    if (binary) {
        BinRaw<N> R;
        bin_fetch<N>(T, e0, p, R);
        bin_finish<N>(T, v, e0, p, R, lut, acc);
    } else {
        post_node<N>(T, v, p, lut, acc);
    }
}

// Two postorder steps (rows v0 and, if has1, v1) for lane pattern p.
template <int N, class LP>
__device__ __forceinline__ void post_pair(const TreeDev &T, int v0, int v1, bool has1, int p, LP lut,
                                          LogAcc &acc) {
    const int a0 = wave_uniform(T.child_ptr[v0]), a1 = wave_uniform(T.child_ptr[v0 + 1]);
    const int b0 = wave_uniform(T.child_ptr[v1]), b1 = wave_uniform(T.child_ptr[v1 + 1]);
    if (a1 - a0 == 2 && b1 - b0 == 2) {
        BinRaw<N> RA, RB;  // both fetched unconditionally so they stay in registers
        bin_fetch<N>(T, a0, p, RA);
        bin_fetch<N>(T, b0, p, RB);
        bin_finish<N>(T, v0, a0, p, RA, lut, acc);
        if (has1) bin_finish<N>(T, v1, b0, p, RB, lut, acc);
    } else {
        // a node's arithmetic depends only on its own arity, never on the
        // node it is paired with (results identical for every schedule)
        post_one<N>(T, v0, a0, a1 - a0 == 2, p, lut, acc);
        if (has1) post_one<N>(T, v1, b0, b1 - b0 == 2, p, lut, acc);
    }
}

// post_pair from descriptors (no dependent scalar loads of the tree structure).
template <int N, class LP>
__device__ __forceinline__ void post_one_d(const TreeDev &T, const PostDesc &d, int p, LP lut, LogAcc &acc) {
    if (d.nchild == 2) {
        BinRaw<N> R;
        bin_fetch_d<N>(T, d, p, R);
        bin_finish_d<N>(T, d, p, R, lut, acc);
    } else {
        post_node<N>(T, d.v, p, lut, acc);
    }
}

template <int N, class LP>
__device__ __forceinline__ void post_pair_d(const TreeDev &T, const PostDesc &d0, const PostDesc &d1, bool has1, int p,
                                            LP lut, LogAcc &acc) {
    if (d0.nchild == 2 && d1.nchild == 2) {
        BinRaw<N> RA, RB;  // both fetched unconditionally so they stay in registers
        bin_fetch_d<N>(T, d0, p, RA);
        bin_fetch_d<N>(T, d1, p, RB);
        bin_finish_d<N>(T, d0, p, RA, lut, acc);
        if (has1) bin_finish_d<N>(T, d1, p, RB, lut, acc);
    } else {
        post_one_d<N>(T, d0, p, lut, acc);
        if (has1) post_one_d<N>(T, d1, p, lut, acc);
    }
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

// Preorder step for node c (treeanc.py:889-930), FP32: outgroup message by
// the reference's division trick, normalise(log max(TINY, parent profile) -
// log(child message)), evolved down the branch (GTR.evolve,
// gtr.py:1023-1025) and multiplied with the subtree row.  The outgroup's own
// normalisation is a per-pattern constant that cancels in the final
// normalisation, so it is not computed.
template <int N>
struct PreRaw {
    float pp[N];
    float sub[N];
    int code;
};

struct PreDesc {
    int c, par, ir, tr;
};
__device__ __forceinline__ PreDesc load_pre_desc(const int32_t TT_C *d) { return PreDesc{d[0], d[1], d[2], d[3]}; }

template <int N>
__device__ __forceinline__ void pre_fetch_d(const TreeDev &T, const PreDesc &q, int p, PreRaw<N> &R) {
    load_rowf<N>(T.prof + ((size_t)q.par * T.P + p) * N, R.pp);
    if (q.ir >= 0) {
        R.code = 0;
        load_rowf<N>(T.sub + ((size_t)q.ir * T.P + p) * N, R.sub);
    } else {
        R.code = T.tip_codes[(size_t)q.tr * T.P + p];
    }
}

template <int N>
__device__ __forceinline__ void pre_fetch(const TreeDev &T, int c, int p, PreRaw<N> &R) {
    pre_fetch_d<N>(T, PreDesc{c, wave_uniform(T.parent[c]), wave_uniform(T.int_row[c]), wave_uniform(T.tip_row[c])},
                   p, R);
}

// Preorder step with the branch's P(t) given as M (memory or registers).
template <int N, class LP, class MP>
__device__ __forceinline__ void pre_finish_m(const TreeDev &T, int c, int ir, int p, const PreRaw<N> &R, LP lut, MP M,
                                             float *lds_out = nullptr) {
    float sub[N], m[N], q[N], msg[N];
    if (ir >= 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) sub[i] = R.sub[i];
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) sub[i] = lut[R.code * N + i];
    }
#pragma unroll
    for (int j = 0; j < N; ++j) m[j] = 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int j = 0; j < N; ++j) m[j] = fmaf(sub[i], M[i * N + j], m[j]);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) q[j] = fmaxf((float)TT_TINY, R.pp[j]) * __builtin_amdgcn_rcpf(m[j]);
    float xs = 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float a = 0.0f;
#pragma unroll
        for (int j = 0; j < N; ++j) a = fmaf(q[j], M[i * N + j], a);
        msg[i] = a * sub[i];
        xs += msg[i];
    }
    const float xi = __builtin_amdgcn_rcpf(xs);
    float out[N];
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = msg[i] * xi;
    const int jm = argmax_first<N>(msg);
    store_row<N>(T.prof + ((size_t)c * T.P + p) * N, out);
    T.states[(size_t)c * T.P + p] = (uint8_t)jm;
    if (lds_out) {
#pragma unroll
        for (int i = 0; i < N; ++i) lds_out[i] = out[i];
    }
}

template <int N, class LP>
__device__ __forceinline__ void pre_finish_ir(const TreeDev &T, int c, int ir, int p, const PreRaw<N> &R, LP lut,
                                              float *lds_out = nullptr) {
    pre_finish_m<N>(T, c, ir, p, R, lut, T.P32 + (size_t)c * (N * N), lds_out);
}

template <int N, class LP>
__device__ __forceinline__ void pre_finish(const TreeDev &T, int c, int p, const PreRaw<N> &R, LP lut) {
    pre_finish_ir<N>(T, c, wave_uniform(T.int_row[c]), p, R, lut);
}

template <int N, class LP>
__device__ __forceinline__ void pre_pair_d(const TreeDev &T, const PreDesc &q0, const PreDesc &q1, bool has1, int p,
                                           LP lut) {
    PreRaw<N> RA, RB;
    pre_fetch_d<N>(T, q0, p, RA);
    pre_fetch_d<N>(T, q1, p, RB);
    pre_finish_ir<N>(T, q0.c, q0.ir, p, RA, lut);
    if (has1) pre_finish_ir<N>(T, q1.c, q1.ir, p, RB, lut);
}

template <int N, class LP>
__device__ __forceinline__ void pre_node(const TreeDev &T, int c, int p, LP lut) {
    PreRaw<N> R;
    pre_fetch<N>(T, c, p, R);
    pre_finish<N>(T, c, p, R, lut);
}

template <int N, class LP>
__device__ __forceinline__ void pre_pair(const TreeDev &T, int c0, int c1, bool has1, int p, LP lut) {
    PreRaw<N> RA, RB;
    pre_fetch<N>(T, c0, p, RA);
    pre_fetch<N>(T, c1, p, RB);
    pre_finish<N>(T, c0, p, RA, lut);
    if (has1) pre_finish<N>(T, c1, p, RB, lut);
}

// ---- joint (max-product) reconstruction -------------------------------------
//
// The reference works on logs (treeanc.py:960-1037): a message is
// Lx_j = max_i log max(TINY, P_ij) + (sum of the children's log messages)_i.
// Max and product commute with the log, so the same decisions are taken in
// product space: messages are stored in FP32, divided by their largest entry
// (which is therefore exactly 1), and every product and comparison is formed
// from the tables max(TINY, P) and max(profile, TINY) (FP32 from 4 states
// on, FP64 below, see JR).  A stored entry is >= TINY (a row whose largest
// entry is 1, times max(TINY, P)), so the product of two children's entries
// times max(TINY, P) is >= 1e-36, inside the normal FP32 range; from a
// polytomy's third child on, the running product is rescaled by a power of
// two after each child (an entry that still underflows is more than 1e-24
// below the entry at the message's maximum and never decides an argmax).  The messages only decide the states: the joint
// log-likelihood is the FP64 log-space score of the decoded states
// (joint_tip_term / joint_int_term), summed after the backtrack, so no FP32
// rounding reaches it.

// Arithmetic of the joint messages: FP32 from 4 states on, where the n^2
// products and comparisons per child are most of the pass's instructions
// (FP64 runs at half the FP32 rate, and its compares and selects take two
// registers per value); FP64 for binary and ternary alphabets.  The messages
// only decide the argmax states (the log-likelihood is the FP64 score of
// those states), and FP32 keeps every decision whose FP64 margin exceeds
// ~1e-6.
template <int N>
using JR = std::conditional_t<(N > 3), float, double>;

template <int N>
__device__ __forceinline__ const JR<N> TT_C *joint_P(const TreeDev &T) {
    if constexpr (N > 3)
        return T.Pj32;
    else
        return T.Pj64;
}

template <int N>
__device__ __forceinline__ const JR<N> TT_C *joint_lut(const TreeDev &T) {
    if constexpr (N > 3)
        return T.lutj32;
    else
        return T.lutj;
}

// Tip child (treeanc.py:970-978, 997-1011): msg_j *= max_i max(TINY, P_ij)
// max(profile_i, TINY); Cx[j] = the first argmax.
template <int N, class LJ>
__device__ __forceinline__ void tip_joint_code(const TreeDev &T, LJ lutj, int code, int cnode, int p,
                                               JR<N> (&msg)[N], int store_cx) {
    using R = JR<N>;
    const auto lj = lutj + code * N;
    const auto PJ = joint_P<N>(T) + (size_t)cnode * (N * N);
    const auto cxo = T.Cx + ((size_t)cnode * T.P + p) * N;  // byte rows (N > 8)
    R tl[N];
#pragma unroll
    for (int i = 0; i < N; ++i) tl[i] = lj[i];
    int cx[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        R best = PJ[j] * tl[0];
        int bi = 0;
#pragma unroll
        for (int i = 1; i < N; ++i) {
            const R val = PJ[i * N + j] * tl[i];
            if (val > best) {
                best = val;
                bi = i;
            }
        }
        msg[j] *= best;
        if constexpr (N > 8) {
            if (store_cx) cxo[j] = (uint8_t)bi;  // large alphabets: no packed argmax row in registers
        } else {
            cx[j] = bi;
        }
    }
    if constexpr (N <= 8) {
        if (store_cx) cx_store<N>(T, cnode, p, cx);
    }
}

template <int N>
__device__ __forceinline__ void tip_joint(const TreeDev &T, int cref, int cnode, int p, JR<N> (&msg)[N],
                                          int store_cx) {
    if constexpr (N > 8) {
        if (T.jtab) {  // the tip branch's table row of the pattern's code (the same numbers as below)
            const int tr = -cref - 1;
            const size_t row = ((size_t)tr * T.n_cc + T.tip_cc[(size_t)tr * T.P + p]) * N;
            float m[N];
            load_rowf<N>(T.jtab + row, m);
#pragma unroll
            for (int j = 0; j < N; ++j) msg[j] *= m[j];
            if (store_cx) {
                uint8_t TT_G *cxo = T.Cx + ((size_t)cnode * T.P + p) * N;
#pragma unroll
                for (int j = 0; j < N; ++j) cxo[j] = T.jcx[row + j];
            }
            return;
        }
    }
    tip_joint_code<N>(T, joint_lut<N>(T), T.tip_codes[(size_t)(-cref - 1) * T.P + p], cnode, p, msg, store_cx);
}

// msg = product over the children of row v of their messages (FP64); a
// polytomy's running product is rescaled by 2^-e (largest entry into
// [0.5, 1)) after each child from the third on (ex_sum, when given, adds up
// the exponents e).
template <int N>
__device__ __forceinline__ void joint_children(const TreeDev &T, int v, int p, JR<N> (&msg)[N], int store_tip_cx,
                                               int *ex_sum = nullptr) {
    using R = JR<N>;
#pragma unroll
    for (int i = 0; i < N; ++i) msg[i] = R(1);
    const int e0 = T.child_ptr[v], e1 = T.child_ptr[v + 1];
    for (int e = e0; e < e1; ++e) {
        const int cref = T.child_ref[e];
        if (cref < 0) {
            tip_joint<N>(T, cref, T.child_node[e], p, msg, store_tip_cx);
        } else {
            float r[N];
            load_rowf<N>(T.Lx + ((size_t)cref * T.P + p) * N, r);
#pragma unroll
            for (int i = 0; i < N; ++i) msg[i] *= (R)r[i];
        }
        if (e - e0 >= 2) {
            R mx = msg[0];
#pragma unroll
            for (int i = 1; i < N; ++i) mx = fmax(mx, msg[i]);
            int ex;
            (void)frexp(mx, &ex);
            if (ex_sum) *ex_sum += ex;
#pragma unroll
            for (int i = 0; i < N; ++i) msg[i] = ldexp(msg[i], -ex);
        }
    }
}

// Joint postorder step for internal non-root row v with node id vn
// (treeanc.py:980-1011): Lx_j = max_i max(TINY, P_ij) msg_i, Cx[j] = the
// first argmax; the row is stored divided by its largest entry.
template <int N>
__device__ __forceinline__ void joint_store(const TreeDev &T, int v, int vn, int p, const JR<N> (&msg)[N]) {
    using R = JR<N>;
    const auto PJ = joint_P<N>(T) + (size_t)vn * (N * N);
    const auto dst = T.Lx + ((size_t)v * T.P + p) * N;
    const auto cxo = T.Cx + ((size_t)vn * T.P + p) * N;  // byte rows (N > 8)
    // large alphabets write the argmax bytes one by one
    R best[N];
    int cx[N];
    R off = R(0);
    int jm = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        R b = PJ[j] * msg[0];
        int bi = 0;
#pragma unroll
        for (int i = 1; i < N; ++i) {
            const R val = PJ[i * N + j] * msg[i];
            if (val > b) {
                b = val;
                bi = i;
            }
        }
        best[j] = b;
        if (b > off) {
            off = b;
            jm = j;
        }
        if constexpr (N > 8) {
            cxo[j] = (uint8_t)bi;
        } else {
            cx[j] = bi;
        }
    }
    const R inv = R(1) / off;
    float out[N];
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = (j == jm) ? 1.0f : (float)(best[j] * inv);
    store_row<N>(dst, out);
    if constexpr (N <= 8) cx_store<N>(T, vn, p, cx);
}

// joint_store for alphabets of more than 8 states with max(TINY, P) of the
// branch read from LDS, transposed (sp[j N + i] = max(TINY, P_ij)): every
// read is one address for the whole wave, and the 400 entries never have to
// live in scalar registers.  Same arithmetic, in the same order.
template <int N>
__device__ __forceinline__ void joint_store_lds(const TreeDev &T, int v, int vn, int p, const float (&msg)[N],
                                                const float *sp) {
    const auto dst = T.Lx + ((size_t)v * T.P + p) * N;
    const auto cxo = T.Cx + ((size_t)vn * T.P + p) * N;
    float best[N];
    float off = 0.0f;
    int jm = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const float *col = sp + j * N;
        float b = col[0] * msg[0];
        int bi = 0;
#pragma unroll
        for (int i = 1; i < N; ++i) {
            const float val = col[i] * msg[i];
            if (val > b) {
                b = val;
                bi = i;
            }
        }
        best[j] = b;
        if (b > off) {
            off = b;
            jm = j;
        }
        cxo[j] = (uint8_t)bi;
    }
    const float inv = 1.0f / off;
    float out[N];
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = (j == jm) ? 1.0f : best[j] * inv;
    store_row<N>(dst, out);
}

template <int N>
__device__ __forceinline__ void post_joint_node(const TreeDev &T, int v, int vn, int p, int store_tip_cx) {
    JR<N> msg[N];
    joint_children<N>(T, v, p, msg, store_tip_cx);
    joint_store<N>(T, v, vn, p, msg);
}

// Binary joint steps split into fetch (the two children's rows or tip codes)
// and finish, so that a wave issues the loads of two nodes before any
// arithmetic (the batched kernel); same arithmetic as post_joint_node.
template <int N>
struct JointRaw {
    int code[2];
    float row[2][N];
};

template <int N>
__device__ __forceinline__ void joint_fetch_d(const TreeDev &T, const PostDesc &d, int p, JointRaw<N> &R) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int cref = c ? d.cref1 : d.cref0;
        if (cref < 0) {
            R.code[c] = T.tip_codes[(size_t)(-cref - 1) * T.P + p];
        } else {
            R.code[c] = 0;
            load_rowf<N>(T.Lx + ((size_t)cref * T.P + p) * N, R.row[c]);
        }
    }
}

template <int N, class LJ>
__device__ __forceinline__ void joint_finish_d(const TreeDev &T, const PostDesc &d, int p, const JointRaw<N> &Raw,
                                               LJ lutj, int store_tip_cx) {
    using R = JR<N>;
    R msg[N];
#pragma unroll
    for (int i = 0; i < N; ++i) msg[i] = R(1);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        if ((c ? d.cref1 : d.cref0) < 0) {
            tip_joint_code<N>(T, lutj, Raw.code[c], c ? d.cnode1 : d.cnode0, p, msg, store_tip_cx);
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) msg[i] *= (R)Raw.row[c][i];
        }
    }
    joint_store<N>(T, d.v, d.vn, p, msg);
}

// Joint root state for pattern p (treeanc.py:1015-1037): the first argmax of
// pi times the product of the children's messages.  The root's row of Lx
// (unused by the messages: the root sends none) receives that product
// divided by its largest entry, exp(normalized_profile) of treeanc.py:1019,
// from which a caller samples the root state.
template <int N>
__device__ __forceinline__ int root_joint_state(const TreeDev &T, int p, int store_tip_cx) {
    using R = JR<N>;
    R msg[N];
    joint_children<N>(T, T.root_row, p, msg, store_tip_cx);
    int jm = 0;
    R mx = R(-1);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        msg[i] *= (R)T.pi[i];
        if (msg[i] > mx) {
            mx = msg[i];
            jm = i;
        }
    }
    const R inv = R(1) / mx;
    float out[N];
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = (i == jm) ? 1.0f : (float)(msg[i] * inv);
    store_row<N>(T.Lx + ((size_t)T.root_row * T.P + p) * N, out);
    T.states[(size_t)T.root * T.P + p] = (uint8_t)jm;
    return jm;
}

// Log-space score terms of the decoded states (the reference's joint
// log-likelihood, treeanc.py:970-1037, evaluated on its own argmax path):
// an internal child c in state sc under parent state sp contributes
// log max(TINY, P_c)[sc][sp]; a tip contributes its message
// max_i log max(TINY, P_c)[i][sp] + log max(profile_i, TINY), which is the
// same number whether or not the tip's state is reconstructed.
template <int N>
__device__ __forceinline__ double joint_int_term(const TreeDev &T, int c, int sc, int sp) {
    return T.logP64[(size_t)c * (N * N) + sc * N + sp];
}

template <int N, class LL>
__device__ __forceinline__ double joint_tip_term(const TreeDev &T, LL loglut, int c, int code, int sp) {
    const auto LP = T.logP64 + (size_t)c * (N * N) + sp;
    const auto ll = loglut + code * N;
    double b = -DBL_MAX;
#pragma unroll
    for (int i = 0; i < N; ++i) b = fmax(b, LP[i * N] + ll[i]);
    return b;
}

// joint_tip_term for small alphabets with wave-uniform c: the branch's
// log P block is read through the scalar unit and each lane selects its
// column sp (no per-lane gathers); the same sums and maxima, so the same value.
template <int N, class LL>
__device__ __forceinline__ double joint_tip_term_sel(const TreeDev &T, LL loglut, int c, int code, int sp) {
    const auto LP = T.logP64 + (size_t)c * (N * N);
    const auto ll = loglut + code * N;
    double b = -DBL_MAX;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double x = LP[i * N];
#pragma unroll
        for (int j = 1; j < N; ++j) x = (sp == j) ? LP[i * N + j] : x;
        b = fmax(b, x + ll[i]);
    }
    return b;
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

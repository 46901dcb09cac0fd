// Site-specific GTR models (reference treetime/gtr_site_specific.py): every
// pattern p has its own eigensystem, P_p(t) = V_p diag(exp(t rho_p)) V_p^-1
// with rho_p = mu_p * eigenvalues_p, and its own root prior pi_p.
//
// A lane owns one pattern, so its model is per-lane data; nothing about the
// matrix is wave-uniform any more.  Messages are formed in the eigenbasis,
// as deficits from the identity (V V^-1 = I):
//   postorder (propagate_profile, gtr_site_specific.py:376-406)
//     m_j = r_j + sum_k (sum_i r_i V_ik) delta_k Vinv_kj
//   preorder (evolve, gtr_site_specific.py:408-437)
//     e_i = q_i + sum_k V_ik delta_k (sum_j Vinv_kj q_j)
// with delta_k = exp(t rho_k) - 1, 2n^2 + n multiply-adds per (branch,
// pattern) instead of n^3.  With the reference's approximate=True
// interpolator (gtr_site_specific.py:331-371) P(t) is linear between the two
// grid points around t, and so is delta: delta = d(t_lo) + w (d(t_hi) -
// d(t_lo)).  The per-branch (t_lo, t_hi, w) come from the host (s_br).
//
// The row layout, the max-normalised FP32 rows, the FP64 normaliser product
// and the partial-row sums are those of the standard level kernels; the
// joint pass builds each lane's P(t) in FP64 and takes its maxima in product
// space.
#include "tt_device.h"

namespace {
using namespace ttd;

template <int N>
struct SiteModel {
    float V[N][N], Vi[N][N], rho[N];
};

template <int N>
__device__ __forceinline__ void load_site(const TreeDev &T, int p, SiteModel<N> &m) {
    const auto v = T.s_V32 + (size_t)p * (N * N);
    const auto vi = T.s_Vi32 + (size_t)p * (N * N);
    const auto r = T.s_rho32 + (size_t)p * N;
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            m.V[i][k] = v[i * N + k];
            m.Vi[i][k] = vi[i * N + k];
        }
        m.rho[i] = r[i];
    }
}

// Branch parameters of node c: (t_lo, t_hi, w); w == 0 means P(t_lo) itself.
struct SiteBranch {
    float ta, tb, w;
};
__device__ __forceinline__ SiteBranch site_branch(const TreeDev &T, int c) {
    const auto b = T.s_br + 3 * (size_t)c;
    return SiteBranch{(float)b[0], (float)b[1], (float)b[2]};
}

template <int N>
__device__ __forceinline__ void site_delta(const SiteModel<N> &m, const SiteBranch &b, float (&d)[N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) d[k] = expm1f(b.ta * m.rho[k]);
    if (b.w != 0.0f) {
#pragma unroll
        for (int k = 0; k < N; ++k) d[k] = fmaf(b.w, expm1f(b.tb * m.rho[k]) - d[k], d[k]);
    }
}

// m = r + Delta (returned as Delta)
template <int N>
__device__ __forceinline__ void site_deficit(const SiteModel<N> &m, const float (&d)[N], const float (&r)[N],
                                             float (&D)[N]) {
    float y[N];
#pragma unroll
    for (int k = 0; k < N; ++k) y[k] = 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int k = 0; k < N; ++k) y[k] = fmaf(r[i], m.V[i][k], y[k]);
    }
#pragma unroll
    for (int k = 0; k < N; ++k) y[k] *= d[k];
#pragma unroll
    for (int j = 0; j < N; ++j) D[j] = 0.0f;
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
        for (int j = 0; j < N; ++j) D[j] = fmaf(y[k], m.Vi[k][j], D[j]);
    }
}

// propagate_profile with return_log clamps at TINY (gtr_site_specific.py:404-406)
__device__ __forceinline__ float site_msg32(float r, float D) { return fmaxf((float)TT_TINY, r + D); }
__device__ __forceinline__ double site_msg64(float r, float D) { return fmax(TT_TINY, (double)r + (double)D); }

template <int N>
__device__ __forceinline__ void site_child(const TreeDev &T, const SiteModel<N> &m, int cref, int cnode, int p,
                                           float (&r)[N], float (&D)[N]) {
    child_row<N>(T, cref, p, T.lut, r);
    float d[N];
    site_delta<N>(m, site_branch(T, cnode), d);
    site_deficit<N>(m, d, r, D);
}

// Postorder step for internal row v (treeanc.py:857-878).  Up to two
// children keep their messages in registers; larger polytomies take the
// log-sum path of the standard kernels and revisit the children for the
// normaliser.
template <int N>
__device__ __forceinline__ void site_post_node(const TreeDev &T, const SiteModel<N> &m, const PostDesc &dd, int p,
                                               LogAcc &acc) {
    float x[N];
    int jm = 0;
    if (dd.nchild <= 2) {
        float r0[N], D0[N], r1[N], D1[N];
#pragma unroll
        for (int j = 0; j < N; ++j) r1[j] = D1[j] = 0.0f;
        site_child<N>(T, m, dd.cref0, dd.cnode0, p, r0, D0);
        if (dd.nchild == 2) site_child<N>(T, m, dd.cref1, dd.cnode1, p, r1, D1);
#pragma unroll
        for (int j = 0; j < N; ++j) {
            x[j] = site_msg32(r0[j], D0[j]);
            if (dd.nchild == 2) x[j] *= site_msg32(r1[j], D1[j]);
        }
        jm = argmax_first<N>(x);
        float a0 = r0[0], b0 = D0[0], a1 = r1[0], b1 = D1[0];
#pragma unroll
        for (int j = 1; j < N; ++j)
            if (j == jm) {
                a0 = r0[j];
                b0 = D0[j];
                a1 = r1[j];
                b1 = D1[j];
            }
        double z = site_msg64(a0, b0);
        if (dd.nchild == 2) z *= site_msg64(a1, b1);
        acc.mul(z);
        float xm = x[0];
#pragma unroll
        for (int j = 1; j < N; ++j)
            if (j == jm) xm = x[j];
        store_post<N>(T, dd.v, p, x, jm, xm);
        return;
    }
    float Ls[N];
#pragma unroll
    for (int j = 0; j < N; ++j) Ls[j] = 0.0f;
    for (int c = 0; c < dd.nchild; ++c) {
        const int cref = wave_uniform(T.child_ref[dd.e0 + c]);
        const int cnode = wave_uniform(T.child_node[dd.e0 + c]);
        float r[N], D[N];
        site_child<N>(T, m, cref, cnode, p, r, D);
#pragma unroll
        for (int j = 0; j < N; ++j) Ls[j] += logf(site_msg32(r[j], D[j]));
    }
    jm = argmax_first<N>(Ls);
    float Lm = Ls[0];
#pragma unroll
    for (int j = 1; j < N; ++j)
        if (j == jm) Lm = Ls[j];
#pragma unroll
    for (int j = 0; j < N; ++j) x[j] = (j == jm) ? 1.0f : expf(Ls[j] - Lm);
    for (int c = 0; c < dd.nchild; ++c) {
        const int cref = wave_uniform(T.child_ref[dd.e0 + c]);
        const int cnode = wave_uniform(T.child_node[dd.e0 + c]);
        float r[N], D[N];
        site_child<N>(T, m, cref, cnode, p, r, D);
        float a = r[0], b = D[0];
#pragma unroll
        for (int j = 1; j < N; ++j)
            if (j == jm) {
                a = r[j];
                b = D[j];
            }
        acc.mul(site_msg64(a, b));
    }
    store_row<N>(T.sub + ((size_t)dd.v * T.P + p) * N, x);
}

template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_post_site(TreeDev T, LevelDev L, int n_ptiles) {
    __shared__ double s_acc[TT_BLOCK / TT_WAVE][TT_WAVE];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int pt = blockIdx.x % n_ptiles, cg = blockIdx.x / n_ptiles;
    const int p = pt * TT_WAVE + lane;
    const bool act = p < T.P;
    const int chunk = cg * (TT_BLOCK / TT_WAVE) + w;
    LogAcc acc;
    if (chunk < L.n_chunks && act) {
        SiteModel<N> m;
        load_site<N>(T, p, m);
        const int kb = chunk * L.chunk;
        const int ke = min(kb + L.chunk, L.n_nodes);
        for (int k = wave_uniform(kb); k < ke; ++k)
            site_post_node<N>(T, m, load_post_desc(L.desc + kPostDesc * k), p, acc);
    }
    s_acc[w][lane] = act ? acc.log_value() : 0.0;
    __syncthreads();
    if (w == 0 && act) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < TT_BLOCK / TT_WAVE; ++q) s += s_acc[q][lane];
        T.part[(size_t)(L.row_off + cg) * T.P + p] = s;
    }
}

// Root with the pattern's own prior (treeanc.py:817-828 with
// gtr.Pi of shape (n, L)).
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_root_site(TreeDev T, int S) {
    const int p = blockIdx.x * TT_BLOCK + threadIdx.x;
    double contrib = 0.0;
    if (p < T.P) {
        double off = 0.0;
        for (int k = 0; k < S; ++k) off += T.part2[(size_t)k * T.P + p];
        double r[N], x[N];
        load_row<N>(T.sub + ((size_t)T.root_row * T.P + p) * N, r);
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            x[j] = T.s_pi[(size_t)p * N + j] * r[j];
            s += x[j];
        }
        const double lp = log(s) + off;
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
        contrib = T.mult[p] * lp;
    }
    const double b = block_sum(contrib);
    if (threadIdx.x == 0) T.bpart[blockIdx.x] = b;
}

// Preorder step for node q.c (treeanc.py:889-930): outgroup by the division
// trick, evolved down the branch in the eigenbasis, times the subtree row.
template <int N>
__device__ __forceinline__ void site_pre_node(const TreeDev &T, const SiteModel<N> &m, const PreDesc &q, int p) {
    float pp[N], sub[N];
    load_rowf<N>(T.prof + ((size_t)q.par * T.P + p) * N, pp);
    if (q.ir >= 0) {
        load_rowf<N>(T.sub + ((size_t)q.ir * T.P + p) * N, sub);
    } else {
        const int code = T.tip_codes[(size_t)q.tr * T.P + p];
#pragma unroll
        for (int i = 0; i < N; ++i) sub[i] = T.lut[code * N + i];
    }
    float d[N], D[N];
    site_delta<N>(m, site_branch(T, q.c), d);
    site_deficit<N>(m, d, sub, D);
    float qv[N];
#pragma unroll
    for (int j = 0; j < N; ++j) qv[j] = fmaxf((float)TT_TINY, pp[j]) / site_msg32(sub[j], D[j]);
    // e = q + V (delta * (Vinv q))
    float y[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        float a = 0.0f;
#pragma unroll
        for (int j = 0; j < N; ++j) a = fmaf(m.Vi[k][j], qv[j], a);
        y[k] = a * d[k];
    }
    float msg[N];
    float xs = 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float a = qv[i];
#pragma unroll
        for (int k = 0; k < N; ++k) a = fmaf(m.V[i][k], y[k], a);
        msg[i] = a * sub[i];
        xs += msg[i];
    }
    const float xi = 1.0f / xs;
    float out[N];
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = msg[i] * xi;
    store_row<N>(T.prof + ((size_t)q.c * T.P + p) * N, out);
    T.states[(size_t)q.c * T.P + p] = (uint8_t)argmax_first<N>(msg);
}

template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_pre_site(TreeDev T, LevelDev L, int n_ptiles) {
    const int lane = threadIdx.x & (TT_WAVE - 1);
    const int task = blockIdx.x * (TT_BLOCK / TT_WAVE) + threadIdx.x / TT_WAVE;
    const int n_quads = (L.n_nodes + 3) / 4;
    if (task >= n_quads * n_ptiles) return;
    const int k0 = 4 * (task / n_ptiles), pt = task % n_ptiles;
    const int p = pt * TT_WAVE + lane;
    if (p >= T.P) return;
    SiteModel<N> m;
    load_site<N>(T, p, m);
    const int ke = min(k0 + 4, L.n_nodes);
    for (int k = wave_uniform(k0); k < ke; ++k) site_pre_node<N>(T, m, load_pre_desc(L.desc + kPreDesc * k), p);
}

// ---- joint (FP64) -------------------------------------------------------------

// max(TINY, P_p(t)) of branch c for this lane's pattern, FP64 (the
// reference's log_transitions before the log, treeanc.py:964-967, with P of
// shape (n, n, L)).
template <int N>
__device__ __forceinline__ void site_Pc(const TreeDev &T, int c, int p, double (&P)[N][N]) {
    const auto b = T.s_br + 3 * (size_t)c;
    const double ta = b[0], tb = b[1], w = b[2];
    const auto rho = T.s_rho64 + (size_t)p * N;
    const auto V = T.s_V64 + (size_t)p * (N * N);
    const auto Vi = T.s_Vi64 + (size_t)p * (N * N);
    double d[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        d[k] = expm1(ta * rho[k]);
        if (w != 0.0) d[k] += w * (expm1(tb * rho[k]) - d[k]);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double vd[N];
#pragma unroll
        for (int k = 0; k < N; ++k) vd[k] = V[i * N + k] * d[k];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double a = (i == j) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k) a = fma(vd[k], Vi[k * N + j], a);
            P[i][j] = fmax(TT_TINY, a);
        }
    }
}

// Lx[j] = max_i log max(TINY, P_ij) + in_i, first argmax in cx, formed in
// product space: with E_i = exp(in_i - m) (m the shift passed in), Lx[j] =
// m + log(max_i P_ij E_i) -- one log per state instead of one per (i, j).
template <int N>
__device__ __forceinline__ void site_joint_max(const double (&P)[N][N], const double (&E)[N], double m,
                                               double (&best)[N], int (&cx)[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double b = P[0][j] * E[0];
        int bi = 0;
#pragma unroll
        for (int i = 1; i < N; ++i) {
            const double val = P[i][j] * E[i];
            if (val > b) {
                b = val;
                bi = i;
            }
        }
        best[j] = m + log(b);
        cx[j] = bi;
    }
}

template <int N>
__device__ __forceinline__ void site_joint_children(const TreeDev &T, int e0, int e1, int p, double (&msg)[N],
                                                    int store_tip_cx) {
    for (int e = e0; e < e1; ++e) {
        const int cref = wave_uniform(T.child_ref[e]);
        if (cref < 0) {
            // tip: exp(log max(profile, TINY)) = max(profile, TINY), no shift
            const int cnode = wave_uniform(T.child_node[e]);
            const int code = T.tip_codes[(size_t)(-cref - 1) * T.P + p];
            double E[N], P[N][N], best[N];
            int cx[N];
#pragma unroll
            for (int i = 0; i < N; ++i) E[i] = fmax(TT_TINY, (double)T.lut[code * N + i]);
            site_Pc<N>(T, cnode, p, P);
            site_joint_max<N>(P, E, 0.0, best, cx);
#pragma unroll
            for (int j = 0; j < N; ++j) msg[j] += best[j];
            if (store_tip_cx) cx_store<N>(T, cnode, p, cx);
        } else {
            const auto src = T.Lx64 + ((size_t)cref * T.P + p) * N;
#pragma unroll
            for (int i = 0; i < N; ++i) msg[i] += src[i];
        }
    }
}

template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_post_joint_site(TreeDev T, LevelDev L, int n_ptiles, int store_tip_cx) {
    __shared__ double s_acc[TT_BLOCK / TT_WAVE][TT_WAVE];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int pt = blockIdx.x % n_ptiles, cg = blockIdx.x / n_ptiles;
    const int p = pt * TT_WAVE + lane;
    const bool act = p < T.P;
    const int chunk = cg * (TT_BLOCK / TT_WAVE) + w;
    double acc = 0.0;
    if (chunk < L.n_chunks && act) {
        const int kb = chunk * L.chunk;
        const int ke = min(kb + L.chunk, L.n_nodes);
        for (int k = wave_uniform(kb); k < ke; ++k) {
            const int v = wave_uniform(L.nodes[k]);
            const int vn = wave_uniform(T.int_node[v]);
            double msg[N];
#pragma unroll
            for (int i = 0; i < N; ++i) msg[i] = 0.0;
            site_joint_children<N>(T, wave_uniform(T.child_ptr[v]), wave_uniform(T.child_ptr[v + 1]), p, msg,
                                   store_tip_cx);
            double P[N][N], E[N], best[N];
            int cx[N];
            double m = msg[0];
#pragma unroll
            for (int i = 1; i < N; ++i) m = fmax(m, msg[i]);
#pragma unroll
            for (int i = 0; i < N; ++i) E[i] = exp(msg[i] - m);
            site_Pc<N>(T, vn, p, P);
            site_joint_max<N>(P, E, m, best, cx);
            double off = best[0];
#pragma unroll
            for (int j = 1; j < N; ++j) off = fmax(off, best[j]);
            const auto dst = T.Lx64 + ((size_t)v * T.P + p) * N;
#pragma unroll
            for (int j = 0; j < N; ++j) dst[j] = best[j] - off;
            cx_store<N>(T, vn, p, cx);
            acc += off;
        }
    }
    s_acc[w][lane] = acc;
    __syncthreads();
    if (w == 0 && act) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < TT_BLOCK / TT_WAVE; ++q) s += s_acc[q][lane];
        T.part[(size_t)(L.row_off + cg) * T.P + p] = s;
    }
}

// Joint root (treeanc.py:1015-1037) with the pattern's own log pi.
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_root_joint_site(TreeDev T, int S, int store_tip_cx) {
    const int p = blockIdx.x * TT_BLOCK + threadIdx.x;
    double contrib = 0.0;
    if (p < T.P) {
        double msg[N];
#pragma unroll
        for (int i = 0; i < N; ++i) msg[i] = 0.0;
        site_joint_children<N>(T, T.child_ptr[T.root_row], T.child_ptr[T.root_row + 1], p, msg, store_tip_cx);
        int jm = 0;
        double mx = -DBL_MAX;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            msg[i] += log(T.s_pi[(size_t)p * N + i]);
            if (msg[i] > mx) {
                mx = msg[i];
                jm = i;
            }
        }
        // the root's row of Lx64 (the root sends no message) keeps its
        // log message, root.joint_Lx (treeanc.py:1017-1018) less the node
        // offsets, for root-state sampling
        const auto dst = T.Lx64 + ((size_t)T.root_row * T.P + p) * N;
#pragma unroll
        for (int i = 0; i < N; ++i) dst[i] = msg[i];
        double lp = mx;
        for (int k = 0; k < S; ++k) lp += T.part2[(size_t)k * T.P + p];
        T.jseqlh[p] = lp;
        T.states[(size_t)T.root * T.P + p] = (uint8_t)jm;
        contrib = T.mult[p] * lp;
    }
    const double b = block_sum(contrib);
    if (threadIdx.x == 0) T.bpart[blockIdx.x] = b;
}

// Debug rows of a joint pass under a site-specific model, as k_joint_rows
// (tt_kernels.hip) for the shared model: internal c (the root too): rel =
// c's log message less its maximum, own = that maximum (the root: of
// log pi_p + the children's log messages); tip c: rel = its log message,
// own = 0, its argmax row stored in Cx.
// A wave handles 64 patterns of one node (the child walk reads the tree
// structure as wave-uniform values).
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_joint_rows_site(TreeDev T, double *rel, double *own) {
    const int ntiles = (T.P + TT_WAVE - 1) / TT_WAVE;
    const long wave = ((long)blockIdx.x * TT_BLOCK + threadIdx.x) / TT_WAVE;
    const int c = (int)(wave / ntiles), p = (int)(wave % ntiles) * TT_WAVE + (int)(threadIdx.x & (TT_WAVE - 1));
    if (c >= T.n_nodes || p >= T.P) return;
    const size_t gid = (size_t)c * T.P + p;
    double *o = rel + gid * N;
    const int ir = T.int_row[c];
    double best[N];
    int cx[N];
    if (ir < 0) {
        const int code = T.tip_codes[(size_t)T.tip_row[c] * T.P + p];
        double E[N], P[N][N];
#pragma unroll
        for (int i = 0; i < N; ++i) E[i] = fmax(TT_TINY, (double)T.lut[code * N + i]);
        site_Pc<N>(T, c, p, P);
        site_joint_max<N>(P, E, 0.0, best, cx);
#pragma unroll
        for (int j = 0; j < N; ++j) o[j] = best[j];
        cx_store<N>(T, c, p, cx);
        own[gid] = 0.0;
        return;
    }
    double msg[N];
#pragma unroll
    for (int i = 0; i < N; ++i) msg[i] = 0.0;
    site_joint_children<N>(T, T.child_ptr[ir], T.child_ptr[ir + 1], p, msg, 0);
    double m = -DBL_MAX;
    if (T.parent[c] < 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            best[i] = msg[i] + log(T.s_pi[(size_t)p * N + i]);
            m = fmax(m, best[i]);
        }
    } else {
        double P[N][N], E[N];
#pragma unroll
        for (int i = 0; i < N; ++i) m = fmax(m, msg[i]);
#pragma unroll
        for (int i = 0; i < N; ++i) E[i] = exp(msg[i] - m);
        site_Pc<N>(T, c, p, P);
        site_joint_max<N>(P, E, m, best, cx);
        m = best[0];
#pragma unroll
        for (int j = 1; j < N; ++j) m = fmax(m, best[j]);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) o[j] = best[j] - m;
    own[gid] = m;
}

// Joint log-likelihood of given root states (written into the root's state
// row by the caller), site-specific model: sequence_LH = root.joint_Lx at the
// chosen state (treeanc.py:1036), the root's stored log message at that
// state plus the node offsets of the last joint pass.
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_root_given_site(TreeDev T, int S) {
    const int p = blockIdx.x * TT_BLOCK + threadIdx.x;
    double contrib = 0.0;
    if (p < T.P) {
        const int s = T.states[(size_t)T.root * T.P + p];
        double lp = T.Lx64[((size_t)T.root_row * T.P + p) * N + s];
        for (int k = 0; k < S; ++k) lp += T.part2[(size_t)k * T.P + p];
        T.jseqlh[p] = lp;
        contrib = T.mult[p] * lp;
    }
    const double b = block_sum(contrib);
    if (threadIdx.x == 0) T.bpart[blockIdx.x] = b;
}

// GTR_site_specific.propagate_profile / evolve (gtr_site_specific.py:376-437)
// with the matrices Q[a][i][j] = P_a(t)[i][j] of every site a:
//   transpose = 0: out[a][j] = sum_i in[a][i] Q[a][i][j], then
//                  log(max(TINY, x)) or max(0, x)
//   transpose = 1: out[a][j] = sum_i in[a][i] Q[a][j][i], then log(x) or x
__global__ void k_propagate_site(const double *in, int P, int n, const double *Q, int transpose, int ret_log,
                                 double *out) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (size_t)P * n) return;
    const size_t a = gid / n;
    const int j = (int)(gid % n);
    const double *q = Q + a * n * n;
    double acc = 0.0;
    for (int i = 0; i < n; ++i) acc += in[a * n + i] * (transpose ? q[j * n + i] : q[i * n + j]);
    if (transpose)
        out[gid] = ret_log ? log(acc) : acc;
    else
        out[gid] = ret_log ? log(fmax(TT_TINY, acc)) : fmax(0.0, acc);
}

inline int n_ptiles(int P) { return (P + TT_WAVE - 1) / TT_WAVE; }
inline unsigned grid_for_tasks(long tasks) { return (unsigned)((tasks + 3) / 4); }

#define TT_SITE_DISPATCH(n, CALL)                       \
    switch (n) {                                        \
        case 2: { constexpr int N_ = 2; CALL; } break;  \
        case 3: { constexpr int N_ = 3; CALL; } break;  \
        case 4: { constexpr int N_ = 4; CALL; } break;  \
        case 5: { constexpr int N_ = 5; CALL; } break;  \
        default: return hipErrorInvalidValue;           \
    }

}  // namespace

namespace ttk {

hipError_t post_site(const TreeDev &T, const LevelDev &L, hipStream_t s) {
    if (L.n_nodes <= 0) return hipSuccess;
    const int npt = n_ptiles(T.P);
    const int groups = (L.n_chunks + 3) / 4;
    TT_SITE_DISPATCH(T.n, hipLaunchKernelGGL(k_post_site<N_>, dim3(groups * npt), dim3(TT_BLOCK), 0, s, T, L, npt));
    return hipGetLastError();
}

hipError_t root_site(const TreeDev &T, int S, hipStream_t s) {
    TT_SITE_DISPATCH(T.n, hipLaunchKernelGGL(k_root_site<N_>, dim3(root_blocks(T.P)), dim3(TT_BLOCK), 0, s, T, S));
    return hipGetLastError();
}

hipError_t pre_site(const TreeDev &T, const LevelDev &L, hipStream_t s) {
    if (L.n_nodes <= 0) return hipSuccess;
    const int npt = n_ptiles(T.P);
    TT_SITE_DISPATCH(T.n, hipLaunchKernelGGL(k_pre_site<N_>, dim3(grid_for_tasks((long)((L.n_nodes + 3) / 4) * npt)),
                                             dim3(TT_BLOCK), 0, s, T, L, npt));
    return hipGetLastError();
}

hipError_t post_joint_site(const TreeDev &T, const LevelDev &L, int store_tip_cx, hipStream_t s) {
    if (L.n_nodes <= 0) return hipSuccess;
    const int npt = n_ptiles(T.P);
    const int groups = (L.n_chunks + 3) / 4;
    TT_SITE_DISPATCH(T.n, hipLaunchKernelGGL(k_post_joint_site<N_>, dim3(groups * npt), dim3(TT_BLOCK), 0, s, T, L,
                                             npt, store_tip_cx));
    return hipGetLastError();
}

hipError_t propagate_site(const double *in, int P, int n, const double *Q, int transpose, int ret_log, double *out,
                          hipStream_t s) {
    const size_t total = (size_t)P * n;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_propagate_site, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, P, n, Q,
                       transpose, ret_log, out);
    return hipGetLastError();
}

hipError_t root_joint_site(const TreeDev &T, int S, int store_tip_cx, hipStream_t s) {
    TT_SITE_DISPATCH(T.n, hipLaunchKernelGGL(k_root_joint_site<N_>, dim3(root_blocks(T.P)), dim3(TT_BLOCK), 0, s, T,
                                             S, store_tip_cx));
    return hipGetLastError();
}

hipError_t joint_rows_site(const TreeDev &T, double *rel, double *own, hipStream_t s) {
    const size_t threads = (size_t)T.n_nodes * n_ptiles(T.P) * TT_WAVE;
    if (!threads) return hipSuccess;
    TT_SITE_DISPATCH(T.n, hipLaunchKernelGGL(k_joint_rows_site<N_>,
                                             dim3((unsigned)((threads + TT_BLOCK - 1) / TT_BLOCK)), dim3(TT_BLOCK), 0,
                                             s, T, rel, own));
    return hipGetLastError();
}

hipError_t root_given_site(const TreeDev &T, int S, hipStream_t s) {
    TT_SITE_DISPATCH(T.n, hipLaunchKernelGGL(k_root_given_site<N_>, dim3(root_blocks(T.P)), dim3(TT_BLOCK), 0, s, T,
                                             S));
    return hipGetLastError();
}

}  // namespace ttk

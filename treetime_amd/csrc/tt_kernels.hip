// HIP kernels of the pruning engine for gfx950 (CDNA4, wave64).
//
// Layout in HBM (pattern-major inside each node):
//   sub  [n_int][P][n]   float  subtree partials, normalised so the largest
//                               entry is exactly 1.0 (reference:
//                               node.marginal_subtree_LH, treeanc.py:877)
//   prof [n_nodes][P][n] float  marginal profiles, sum = 1 (treeanc.py:909-917)
//   Lx   [n_int][P][n]   float  joint (max-product) messages, largest entry exactly 1
//   Cx   [n_nodes][P] (n <= 4, 2 bits per parent state) or [n_nodes][P][n]
//        uint8  joint argmax table (treeanc.py:996-1011)
//   part [R][P]          double per-chunk-group sums of log normalisers
//
// Precision scheme: rows are stored in FP32, arithmetic is FP64.  Because a
// stored row's dominant entry is exactly 1, the only FP32
// rounding that reaches the likelihood is on entries weighted by their
// (small) posterior, and the per-node log normalisers are accumulated in
// FP64.  This keeps the total log-likelihood within ~1e-9 relative of the
// FP64 reference, where naive FP32 storage drifts by ~1e-6.
//
// Work decomposition of the level kernels: a wave owns 64 consecutive
// patterns of one node (lane = pattern), so everything that depends on the
// node (child list, P(t) of the branch) is wave-uniform and is read through
// the scalar unit.
#include "tt_device.h"

#include <cstdlib>

namespace {
using namespace ttd;

template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_post_marginal(TreeDev T, LevelDev L, int n_ptiles) {
    __shared__ double s_acc[TT_BLOCK / TT_WAVE][TT_WAVE];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int pt = blockIdx.x % n_ptiles, cg = blockIdx.x / n_ptiles;
    const int p = pt * TT_WAVE + lane;
    const bool act = p < T.P;
    const int chunk = cg * (TT_BLOCK / TT_WAVE) + w;
    LogAcc acc;
    if (chunk < L.n_chunks) {
        const int kb = chunk * L.chunk;
        const int ke = min(kb + L.chunk, L.n_nodes);
        for (int k = wave_uniform(kb); k < ke; k += 2) {
            const bool has1 = k + 1 < ke;
            const PostDesc d0 = load_post_desc(L.desc + kPostDesc * k);
            const PostDesc d1 = has1 ? load_post_desc(L.desc + kPostDesc * (k + 1)) : d0;
            if (act) post_pair_d<N>(T, d0, d1, has1, p, T.lut, acc);
        }
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

// Sums row slices of the partial-normaliser matrix: part2[s][p] = sum of
// part[r][p] over r in slice s.
__global__ __launch_bounds__(TT_BLOCK) void k_rowsum(TreeDev T, int R, int S, int G, int n_ptiles) {
    const int lane = threadIdx.x & (TT_WAVE - 1);
    const int task = blockIdx.x * (TT_BLOCK / TT_WAVE) + threadIdx.x / TT_WAVE;
    const int pt = task % n_ptiles, s = task / n_ptiles;
    const int p = pt * TT_WAVE + lane;
    if (s >= S || p >= T.P) return;
    const int r0 = s * G, r1 = min(R, r0 + G);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int r = r0;
    for (; r + 3 < r1; r += 4) {
        a0 += T.part[(size_t)r * T.P + p];
        a1 += T.part[(size_t)(r + 1) * T.P + p];
        a2 += T.part[(size_t)(r + 2) * T.P + p];
        a3 += T.part[(size_t)(r + 3) * T.P + p];
    }
    for (; r < r1; ++r) a0 += T.part[(size_t)r * T.P + p];
    T.part2[(size_t)s * T.P + p] = (a0 + a1) + (a2 + a3);
}

// Root (treeanc.py:817-828) for pattern p: profile = normalise(pi * subtree),
// per-pattern logL = root log-norm + all node log normalisers (passed in as
// ``offsets``); returns mult[p] * logL.
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_root_marginal(TreeDev T, int S) {
    const int p = blockIdx.x * TT_BLOCK + threadIdx.x;
    double contrib = 0.0;
    if (p < T.P) {
        double off = 0.0;
        for (int k = 0; k < S; ++k) off += T.part2[(size_t)k * T.P + p];
        contrib = root_step<N>(T, p, off);
    }
    const double b = block_sum(contrib);
    if (threadIdx.x == 0) T.bpart[blockIdx.x] = b;
}

__global__ __launch_bounds__(TT_BLOCK) void k_final_sum(TreeDev T, int nblocks) {
    double x = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += TT_BLOCK) x += T.bpart[i];
    const double b = block_sum(x);
    if (threadIdx.x == 0) T.total[0] = b;
}

// Preorder level: a wave handles two nodes of the level for 64 patterns.
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_pre_marginal(TreeDev T, LevelDev L, int n_ptiles) {
    const int lane = threadIdx.x & (TT_WAVE - 1);
    const int task = blockIdx.x * (TT_BLOCK / TT_WAVE) + threadIdx.x / TT_WAVE;
    const int n_pairs = (L.n_nodes + 1) / 2;
    if (task >= n_pairs * n_ptiles) return;
    const int k = 2 * (task / n_ptiles), pt = task % n_ptiles;
    const int p = pt * TT_WAVE + lane;
    if (p >= T.P) return;
    const bool has1 = k + 1 < L.n_nodes;
    const int ku = wave_uniform(k);
    const PreDesc q0 = load_pre_desc(L.desc + kPreDesc * ku);
    const PreDesc q1 = has1 ? load_pre_desc(L.desc + kPreDesc * (ku + 1)) : q0;
    pre_pair_d<N>(T, q0, q1, has1, p, T.lut);
}

// Runs of narrow levels (few nodes each, typically near the root): one launch
// walks all of them.  A block of 4 waves takes one 64-pattern tile through
// every level of the run -- the level's nodes split over the waves, two per
// step -- with a block barrier between levels, so a narrow level costs one
// or two node steps instead of a kernel launch.  Same per-(node, pattern)
// arithmetic as the level kernels.
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_post_run(TreeDev T, RunDev R) {
    constexpr int kW = TT_BLOCK / TT_WAVE;
    __shared__ double s_acc[kW][TT_WAVE];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int p = blockIdx.x * TT_WAVE + lane;
    const bool act = p < T.P;
    LogAcc acc;
    for (int l = R.l0; l < R.l1; ++l) {
        const int kb = R.off[l], ke = R.off[l + 1];
        for (int k = wave_uniform(kb + w); k < ke; k += 2 * kW) {
            const bool has1 = k + kW < ke;
            const PostDesc d0 = load_post_desc(R.desc + kPostDesc * k);
            const PostDesc d1 = has1 ? load_post_desc(R.desc + kPostDesc * (k + kW)) : d0;
            if (act) post_pair_d<N>(T, d0, d1, has1, p, T.lut, acc);
        }
        __syncthreads();
    }
    s_acc[w][lane] = act ? acc.log_value() : 0.0;
    __syncthreads();
    if (w == 0 && act) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < kW; ++q) s += s_acc[q][lane];
        T.part[(size_t)R.row_first * T.P + p] = s;
        for (int r = R.row_first + 1; r < R.row_end; ++r) T.part[(size_t)r * T.P + p] = 0.0;
    }
}

template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_pre_run(TreeDev T, RunDev R) {
    constexpr int kW = TT_BLOCK / TT_WAVE;
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int p = blockIdx.x * TT_WAVE + lane;
    const bool act = p < T.P;
    for (int l = R.l0; l < R.l1; ++l) {
        const int kb = R.off[l];
        const int ke = R.tips ? R.off[l + 1] : kb + R.nint[l];
        for (int k = wave_uniform(kb + w); k < ke; k += 2 * kW) {
            const bool has1 = k + kW < ke;
            const PreDesc q0 = load_pre_desc(R.desc + kPreDesc * k);
            const PreDesc q1 = has1 ? load_pre_desc(R.desc + kPreDesc * (k + kW)) : q0;
            if (act) pre_pair_d<N>(T, q0, q1, has1, p, T.lut);
        }
        __syncthreads();
    }
}

template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_post_joint(TreeDev T, LevelDev L, int n_ptiles, int store_tip_cx) {
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int pt = blockIdx.x % n_ptiles, cg = blockIdx.x / n_ptiles;
    const int p = pt * TT_WAVE + lane;
    const int chunk = cg * (TT_BLOCK / TT_WAVE) + w;
    if (chunk >= L.n_chunks || p >= T.P) return;
    const int kb = chunk * L.chunk;
    const int ke = min(kb + L.chunk, L.n_nodes);
    for (int k = kb; k < ke; ++k) {
        const int v = wave_uniform(L.nodes[k]);
        post_joint_node<N>(T, v, wave_uniform(T.int_node[v]), p, store_tip_cx);
    }
}

// Joint postorder level for alphabets of more than 8 states: as
// k_post_joint, with each node's max(TINY, P) staged transposed in the
// wave's own LDS slice (joint_store_lds).
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_post_joint_wide(TreeDev T, LevelDev L, int n_ptiles,
                                                              int store_tip_cx) {
    __shared__ float s_p[TT_BLOCK / TT_WAVE][N * N];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int pt = blockIdx.x % n_ptiles, cg = blockIdx.x / n_ptiles;
    const int p = pt * TT_WAVE + lane;
    const int chunk = cg * (TT_BLOCK / TT_WAVE) + w;
    if (chunk >= L.n_chunks) return;  // whole wave
    const int kb = chunk * L.chunk;
    const int ke = min(kb + L.chunk, L.n_nodes);
    float *sp = s_p[w];
    for (int k = kb; k < ke; ++k) {
        const int v = wave_uniform(L.nodes[k]);
        const int vn = wave_uniform(T.int_node[v]);
        const float TT_C *PJ = T.Pj32 + (size_t)vn * (N * N);
        __builtin_amdgcn_wave_barrier();  // the previous node's reads of the slice come first
        for (int q = lane; q < N * N; q += TT_WAVE) sp[(q % N) * N + q / N] = PJ[q];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (p < T.P) {
            float msg[N];
            joint_children<N>(T, v, p, msg, store_tip_cx);
            joint_store_lds<N>(T, v, vn, p, msg, sp);
        }
    }
}

// Runs of narrow joint levels in one launch (as k_post_run): a block of 4
// waves takes one 64-pattern tile through every level of the run, the
// level's nodes dealt to the waves, a block barrier between levels.
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_post_joint_run(TreeDev T, RunDev R, int store_tip_cx) {
    constexpr int kW = TT_BLOCK / TT_WAVE;
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int p = blockIdx.x * TT_WAVE + lane;
    for (int l = R.l0; l < R.l1; ++l) {
        const int kb = R.off[l], ke = R.off[l + 1];
        for (int k = wave_uniform(kb + w); k < ke; k += kW) {
            const int v = wave_uniform(R.nodes[k]);
            if (p < T.P) post_joint_node<N>(T, v, wave_uniform(T.int_node[v]), p, store_tip_cx);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(TT_BLOCK) void k_backtrack_run(TreeDev T, RunDev R) {
    constexpr int kW = TT_BLOCK / TT_WAVE;
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int p = blockIdx.x * TT_WAVE + lane;
    for (int l = R.l0; l < R.l1; ++l) {
        const int kb = R.off[l];
        const int ke = R.tips ? R.off[l + 1] : kb + R.nint[l];
        for (int k = wave_uniform(kb + w); k < ke; k += kW) {
            const int c = wave_uniform(R.nodes[k]);
            if (p < T.P) {
                const int ps = T.states[(size_t)T.parent[c] * T.P + p];
                T.states[(size_t)c * T.P + p] = cx_state(T, c, p, ps);
            }
        }
        __syncthreads();
    }
}

// Joint root state (treeanc.py:1015-1037), one thread per pattern.
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_root_joint(TreeDev T, int store_tip_cx) {
    const int p = blockIdx.x * TT_BLOCK + threadIdx.x;
    if (p < T.P) (void)root_joint_state<N>(T, p, store_tip_cx);
}

// Score of the decoded joint states, part 1: the wave of (slice sl, pattern
// tile) sums the terms of node ids [sl G, sl G + G) (every non-root branch
// once) into part row sl.
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_joint_score(TreeDev T, int S, int G, int n_ptiles) {
    const int lane = threadIdx.x & (TT_WAVE - 1);
    const int task = blockIdx.x * (TT_BLOCK / TT_WAVE) + threadIdx.x / TT_WAVE;
    const int pt = task % n_ptiles, sl = task / n_ptiles;
    const int p = pt * TT_WAVE + lane;
    if (sl >= S || p >= T.P) return;
    const int v0 = sl * G, v1 = min(T.n_nodes, v0 + G);
    double a = 0.0;
    for (int v = wave_uniform(v0); v < v1; ++v) {
        const int par = wave_uniform(T.parent[v]);
        if (par < 0) continue;
        const int sp = T.states[(size_t)par * T.P + p];
        const int tr = wave_uniform(T.tip_row[v]);
        if (tr >= 0)
            a += joint_tip_term<N>(T, T.loglut, v, T.tip_codes[(size_t)tr * T.P + p], sp);
        else
            a += joint_int_term<N>(T, v, T.states[(size_t)v * T.P + p], sp);
    }
    T.part[(size_t)sl * T.P + p] = a;
}

// Part 2, one thread per pattern: jseqlh = log pi[root state] + the S2
// row-slice sums; block partials of sum(mult * jseqlh).
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_joint_final(TreeDev T, int S2) {
    const int p = blockIdx.x * TT_BLOCK + threadIdx.x;
    double contrib = 0.0;
    if (p < T.P) {
        double lp = log(T.pi[T.states[(size_t)T.root * T.P + p]]);
        for (int k = 0; k < S2; ++k) lp += T.part2[(size_t)k * T.P + p];
        T.jseqlh[p] = lp;
        contrib = T.mult[p] * lp;
    }
    const double b = block_sum(contrib);
    if (threadIdx.x == 0) T.bpart[blockIdx.x] = b;
}

// Joint backtrack (treeanc.py:1048-1063): state = Cx[state of parent].
__global__ __launch_bounds__(TT_BLOCK) void k_backtrack(TreeDev T, LevelDev L, int n_ptiles) {
    const int lane = threadIdx.x & (TT_WAVE - 1);
    const int task = blockIdx.x * (TT_BLOCK / TT_WAVE) + threadIdx.x / TT_WAVE;
    if (task >= L.n_nodes * n_ptiles) return;
    const int k = task / n_ptiles, pt = task % n_ptiles;
    const int p = pt * TT_WAVE + lane;
    if (p >= T.P) return;
    const int c = wave_uniform(L.nodes[k]);
    const int par = T.parent[c];
    const int ps = T.states[(size_t)par * T.P + p];
    T.states[(size_t)c * T.P + p] = cx_state(T, c, p, ps);
}

// Branch re-evaluation along one root path (update mode).  Starting at the
// changed branch's parent, every internal node up to the root is recomputed
// from its children (the changed one was just rewritten by this thread), and
// sequence_LH moves by the change of those nodes' log normalisers and of the
// root term.  Equals a full postorder + root (treeanc.py:840-878, 814-829)
// for a tree whose other rows are current.
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_path_update(TreeDev T, int node) {
    const int p = blockIdx.x * TT_BLOCK + threadIdx.x;
    double contrib = 0.0;
    if (p < T.P) {
        auto root_term = [&]() {
            float r[N];
            load_rowf<N>(T.sub + ((size_t)T.root_row * T.P + p) * N, r);
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < N; ++j) s += T.pi[j] * (double)r[j];
            return log(s);
        };
        const double root_old = root_term();
        double dl = 0.0;
        LogAcc unused;
        for (int a = wave_uniform(T.parent[node]); a >= 0; a = wave_uniform(T.parent[a])) {
            const int v = wave_uniform(T.int_row[a]);
            const double old = T.offs[(size_t)v * T.P + p];
            post_pair<N>(T, v, v, false, p, T.lut, unused);
            dl += T.offs[(size_t)v * T.P + p] - old;
        }
        const double lp = T.seqlh[p] + dl + (root_term() - root_old);
        T.seqlh[p] = lp;
        contrib = T.mult[p] * lp;
    }
    const double b = block_sum(contrib);
    if (threadIdx.x == 0) T.bpart[blockIdx.x] = b;
}

__global__ void k_fill_states(TreeDev T, const int32_t *nodes, int n_nodes, uint8_t value) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n_nodes * T.P) return;
    const int k = (int)(i / T.P), p = (int)(i % T.P);
    T.states[(size_t)nodes[k] * T.P + p] = value;
}

// P(t) for a list of branches (gtr.py:1027-1067): v diag(exp(mu t lambda))
// v_inv, clipped at 0, plus max(TINY, P) and its log for the joint pass
// (treeanc.py:970).
__global__ void k_expqt(int n, const double *t, const int32_t *branches, int k, const double *eig,
                        const double *v, const double *vinv, double mu, double *P64, double *Pj64, double *logP64,
                        float *Pj32, float *P32, float *D32) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nn = n * n;
    if (gid >= k * nn) return;
    const int b = gid / nn, ij = gid % nn, i = ij / n, j = ij % n;
    const double tb = t[b];
    double acc = 0.0;
    for (int q = 0; q < n; ++q) acc += v[i * n + q] * (exp(mu * tb * eig[q]) * vinv[q * n + j]);
    const double P = fmax(0.0, acc);
    const size_t o = (size_t)branches[b] * nn + ij;
    P64[o] = P;
    Pj64[o] = fmax(TT_TINY, P);
    logP64[o] = log(fmax(TT_TINY, P));
    Pj32[o] = (float)fmax(TT_TINY, P);
    P32[o] = (float)P;
    D32[o] = (float)(P - (i == j ? 1.0 : 0.0));  // deficit form keeps the diagonal's small part exact
}

// out = in @ Q (transpose=0, GTR.propagate_profile) or in @ Q^T (transpose=1,
// GTR.evolve); optional log.
__global__ void k_propagate(const double *in, int P, int n, const double *Q, int transpose, int ret_log,
                            double *out) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (size_t)P * n) return;
    const size_t a = gid / n;
    const int j = (int)(gid % n);
    double acc = 0.0;
    for (int i = 0; i < n; ++i) acc += in[a * n + i] * (transpose ? Q[j * n + i] : Q[i * n + j]);
    out[gid] = ret_log ? log(acc) : acc;
}

// GTR.prob_t_profiles (gtr.py:950-961) for K matrices; one block partial
// per (k, block).
__global__ __launch_bounds__(TT_BLOCK) void k_prob_t_profiles(const double *pp, const double *pc,
                                                              const double *mult, int P, int n, const double *Q,
                                                              int gap, double *block_out, int nblk) {
    const int a = blockIdx.x * TT_BLOCK + threadIdx.x;
    const int kq = blockIdx.y;
    const double *Qk = Q + (size_t)kq * n * n;
    double contrib = 0.0;
    if (a < P) {
        double res = 0.0;
        for (int i = 0; i < n; ++i) {
            double inner = 0.0;
            for (int j = 0; j < n; ++j) inner += Qk[i * n + j] * pp[(size_t)a * n + j];
            res += pc[(size_t)a * n + i] * inner;
        }
        double lg = mult[a] * log(res + 1e-24);
        if (gap >= 0) lg *= (1.0 - pp[(size_t)a * n + gap]) * (1.0 - pc[(size_t)a * n + gap]);
        contrib = lg;
    }
    const double b = block_sum(contrib);
    if (threadIdx.x == 0) block_out[(size_t)kq * nblk + blockIdx.x] = b;
}

// Rows of a joint pass for the debug attributes (treeanc.py:1066-1077 keeps
// joint_Lx and joint_Cx with debug=True), recomputed from the stored
// messages with the joint kernels' arithmetic, one thread per (node c,
// pattern p):
//   internal c (the root too): rel[c][p][j] = log of c's message (the root:
//     of pi times the children's product) less its log maximum, and
//     own[c][p] = that log maximum plus ln 2 per polytomy rescaling step, so
//     that the absolute log message is rel plus the sum of own over the
//     internal nodes of c's subtree;
//   tip c: rel = its absolute log message max_i log max(TINY, P_ij) +
//     log max(profile_i, TINY), own = 0, and its argmax row goes to Cx (the
//     joint pass stores tips' rows only when it reconstructs tips).
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_joint_rows(TreeDev T, double *rel, double *own) {
    using R = JR<N>;
    const size_t gid = (size_t)blockIdx.x * TT_BLOCK + threadIdx.x;
    if (gid >= (size_t)T.n_nodes * T.P) return;
    const int c = (int)(gid / T.P), p = (int)(gid % T.P);
    double *o = rel + gid * N;
    const int ir = T.int_row[c];
    if (ir < 0) {
        const int code = T.tip_codes[(size_t)T.tip_row[c] * T.P + p];
        for (int j = 0; j < N; ++j) o[j] = joint_tip_term<N>(T, T.loglut, c, code, j);
        R msg[N];
#pragma unroll
        for (int i = 0; i < N; ++i) msg[i] = R(1);
        tip_joint_code<N>(T, joint_lut<N>(T), code, c, p, msg, 1);
        own[gid] = 0.0;
        return;
    }
    R msg[N], x[N];
    int ex = 0;
    joint_children<N>(T, ir, p, msg, 0, &ex);
    R mx = R(-1);
    if (T.parent[c] < 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            x[i] = msg[i] * (R)T.pi[i];
            mx = x[i] > mx ? x[i] : mx;
        }
    } else {
        const auto PJ = joint_P<N>(T) + (size_t)c * (N * N);
        for (int j = 0; j < N; ++j) {
            R b = PJ[j] * msg[0];
            for (int i = 1; i < N; ++i) {
                const R val = PJ[i * N + j] * msg[i];
                if (val > b) b = val;
            }
            x[j] = b;
            mx = b > mx ? b : mx;
        }
    }
    const double lm = log((double)mx);
    for (int j = 0; j < N; ++j) o[j] = log((double)x[j]) - lm;
    own[gid] = lm + ex * 0.6931471805599453094;
}

inline int n_ptiles(int P) { return (P + TT_WAVE - 1) / TT_WAVE; }
inline unsigned grid_for_tasks(long tasks) { return (unsigned)((tasks + 3) / 4); }

}  // namespace

#define TT_DISPATCH(n, CALL)               \
    switch (n) {                           \
        case 2: { constexpr int N_ = 2; CALL; } break;   \
        case 3: { constexpr int N_ = 3; CALL; } break;   \
        case 4: { constexpr int N_ = 4; CALL; } break;   \
        case 5: { constexpr int N_ = 5; CALL; } break;   \
        case 20: { constexpr int N_ = 20; CALL; } break; \
        case 21: { constexpr int N_ = 21; CALL; } break; \
        case 22: { constexpr int N_ = 22; CALL; } break; \
        default: return hipErrorInvalidValue;            \
    }

namespace {

// Tip tables of the joint pass (n > 8): for tip row t, compact code c and
// parent state j, the best child state's message max_i max(TINY, P_ij)
// max(profile_i, TINY) and the first argmax, with the arithmetic of
// tip_joint_code (so the table holds the same numbers per pattern).
template <int N>
__global__ void k_joint_tiptab(TreeDev T, const int32_t *tip_nodes) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int mN = T.n_cc * N;
    if (gid >= (long)T.n_tips * mN) return;
    const int t = (int)(gid / mN), r = (int)(gid % mN), c = r / N, j = r % N;
    const float TT_C *PJ = T.Pj32 + (size_t)tip_nodes[t] * (N * N);
    const float TT_C *lj = T.lutj_c + c * N;
    float best = PJ[j] * lj[0];
    int bi = 0;
    for (int i = 1; i < N; ++i) {
        const float val = PJ[i * N + j] * lj[i];
        if (val > best) {
            best = val;
            bi = i;
        }
    }
    T.jtab[(size_t)t * mN + r] = best;
    T.jcx[(size_t)t * mN + r] = (uint8_t)bi;
}

}  // namespace

namespace ttk {

hipError_t joint_tiptab(const TreeDev &T, const int32_t *tip_nodes, hipStream_t s) {
    const long total = (long)T.n_tips * T.n_cc * T.n;
    if (!T.jtab || total <= 0) return hipSuccess;
    const dim3 grid((unsigned)((total + 255) / 256)), block(256);
    switch (T.n) {
        case 20: hipLaunchKernelGGL(k_joint_tiptab<20>, grid, block, 0, s, T, tip_nodes); break;
        case 21: hipLaunchKernelGGL(k_joint_tiptab<21>, grid, block, 0, s, T, tip_nodes); break;
        case 22: hipLaunchKernelGGL(k_joint_tiptab<22>, grid, block, 0, s, T, tip_nodes); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int root_blocks(int P) { return (P + TT_BLOCK - 1) / TT_BLOCK; }

// The 20-state passes use the MFMA kernels unless TT_NO_MFMA=1 (then the
// VALU template instance runs; both are parity-tested).
bool mfma_enabled() {
    static bool v = [] {
        const char *e = getenv("TT_NO_MFMA");
        return !(e && e[0] == '1');
    }();
    return v;
}

hipError_t expqt(int n, const double *t, const int32_t *branches, int k, const double *eig, const double *v,
                 const double *vinv, double mu, double *P64, double *Pj64, double *logP64, float *Pj32, float *P32,
                 float *D32, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    const int total = k * n * n;
    hipLaunchKernelGGL(k_expqt, dim3((total + 255) / 256), dim3(256), 0, s, n, t, branches, k, eig, v, vinv, mu,
                       P64, Pj64, logP64, Pj32, P32, D32);
    return hipGetLastError();
}

hipError_t post_marginal(const TreeDev &T, const LevelDev &L, hipStream_t s) {
    if (T.n == 20 && mfma_enabled()) return post_marginal20(T, L, s);
    const int npt = n_ptiles(T.P);
    const int groups = (L.n_chunks + 3) / 4;
    TT_DISPATCH(T.n, hipLaunchKernelGGL(k_post_marginal<N_>, dim3(groups * npt), dim3(TT_BLOCK), 0, s, T, L, npt));
    return hipGetLastError();
}

hipError_t rowsum(const TreeDev &T, int R, int S, hipStream_t s) {
    if (S <= 0) return hipSuccess;
    const int npt = n_ptiles(T.P);
    const int G = (R + S - 1) / S;
    hipLaunchKernelGGL(k_rowsum, dim3(grid_for_tasks((long)S * npt)), dim3(TT_BLOCK), 0, s, T, R, S, G, npt);
    return hipGetLastError();
}

hipError_t root_marginal(const TreeDev &T, int S, hipStream_t s) {
    TT_DISPATCH(T.n, hipLaunchKernelGGL(k_root_marginal<N_>, dim3(root_blocks(T.P)), dim3(TT_BLOCK), 0, s, T, S));
    return hipGetLastError();
}

hipError_t final_sum(const TreeDev &T, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(TT_BLOCK), 0, s, T, nblocks);
    return hipGetLastError();
}

hipError_t post_run(const TreeDev &T, const RunDev &R, hipStream_t s) {
    if (T.n == 20 && mfma_enabled()) return post_run20(T, R, s);
    TT_DISPATCH(T.n, hipLaunchKernelGGL(k_post_run<N_>, dim3(n_ptiles(T.P)), dim3(TT_BLOCK), 0, s, T, R));
    return hipGetLastError();
}

hipError_t pre_run(const TreeDev &T, const RunDev &R, hipStream_t s) {
    if (T.n == 20 && mfma_enabled()) return pre_run20(T, R, s);
    TT_DISPATCH(T.n, hipLaunchKernelGGL(k_pre_run<N_>, dim3(n_ptiles(T.P)), dim3(TT_BLOCK), 0, s, T, R));
    return hipGetLastError();
}

hipError_t pre_marginal(const TreeDev &T, const LevelDev &L, hipStream_t s) {
    if (L.n_nodes <= 0) return hipSuccess;
    if (T.n == 20 && mfma_enabled()) return pre_marginal20(T, L, s);
    const int npt = n_ptiles(T.P);
    TT_DISPATCH(T.n, hipLaunchKernelGGL(k_pre_marginal<N_>, dim3(grid_for_tasks((long)((L.n_nodes + 1) / 2) * npt)),
                                        dim3(TT_BLOCK), 0, s, T, L, npt));
    return hipGetLastError();
}

hipError_t joint_rows(const TreeDev &T, double *rel, double *own, hipStream_t s) {
    const size_t total = (size_t)T.n_nodes * T.P;
    if (!total) return hipSuccess;
    TT_DISPATCH(T.n, hipLaunchKernelGGL(k_joint_rows<N_>, dim3((unsigned)((total + TT_BLOCK - 1) / TT_BLOCK)),
                                        dim3(TT_BLOCK), 0, s, T, rel, own));
    return hipGetLastError();
}

hipError_t root_joint(const TreeDev &T, int store_tip_cx, hipStream_t s) {
    TT_DISPATCH(T.n, hipLaunchKernelGGL(k_root_joint<N_>, dim3(root_blocks(T.P)), dim3(TT_BLOCK), 0, s, T,
                                        store_tip_cx));
    return hipGetLastError();
}

hipError_t joint_score(const TreeDev &T, int S, int G, hipStream_t s) {
    if (S <= 0) return hipSuccess;
    const int npt = n_ptiles(T.P);
    TT_DISPATCH(T.n, hipLaunchKernelGGL(k_joint_score<N_>, dim3(grid_for_tasks((long)S * npt)), dim3(TT_BLOCK), 0,
                                        s, T, S, G, npt));
    return hipGetLastError();
}

hipError_t joint_final(const TreeDev &T, int S2, hipStream_t s) {
    TT_DISPATCH(T.n, hipLaunchKernelGGL(k_joint_final<N_>, dim3(root_blocks(T.P)), dim3(TT_BLOCK), 0, s, T, S2));
    return hipGetLastError();
}

hipError_t post_joint_run(const TreeDev &T, const RunDev &R, int store_tip_cx, hipStream_t s) {
    if (T.n > 8) return hipErrorInvalidValue;
    TT_DISPATCH(T.n, hipLaunchKernelGGL(k_post_joint_run<N_>, dim3(n_ptiles(T.P)), dim3(TT_BLOCK), 0, s, T, R,
                                        store_tip_cx));
    return hipGetLastError();
}

hipError_t backtrack_run(const TreeDev &T, const RunDev &R, hipStream_t s) {
    hipLaunchKernelGGL(k_backtrack_run, dim3(n_ptiles(T.P)), dim3(TT_BLOCK), 0, s, T, R);
    return hipGetLastError();
}

hipError_t backtrack(const TreeDev &T, const LevelDev &L, hipStream_t s) {
    if (L.n_nodes <= 0) return hipSuccess;
    const int npt = n_ptiles(T.P);
    hipLaunchKernelGGL(k_backtrack, dim3(grid_for_tasks((long)L.n_nodes * npt)), dim3(TT_BLOCK), 0, s, T, L, npt);
    return hipGetLastError();
}

hipError_t path_update(const TreeDev &T, int node, hipStream_t s) {
    TT_DISPATCH(T.n, hipLaunchKernelGGL(k_path_update<N_>, dim3(root_blocks(T.P)), dim3(TT_BLOCK), 0, s, T, node));
    return hipGetLastError();
}

hipError_t fill_states(const TreeDev &T, const int32_t *nodes, int n_nodes, uint8_t value, hipStream_t s) {
    if (n_nodes <= 0) return hipSuccess;
    const size_t total = (size_t)n_nodes * T.P;
    hipLaunchKernelGGL(k_fill_states, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, T, nodes, n_nodes,
                       value);
    return hipGetLastError();
}

hipError_t propagate(const double *in, int P, int n, const double *Q, int transpose, int ret_log, double *out,
                     hipStream_t s) {
    const size_t total = (size_t)P * n;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_propagate, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, P, n, Q, transpose,
                       ret_log, out);
    return hipGetLastError();
}

hipError_t prob_t_profiles(const double *pp, const double *pc, const double *mult, int P, int n, const double *Q,
                           int K, int gap, double *block_out, int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_prob_t_profiles, dim3(nblk, K), dim3(TT_BLOCK), 0, s, pp, pc, mult, P, n, Q, gap,
                       block_out, nblk);
    return hipGetLastError();
}

}  // namespace ttk

hipError_t ttk::post_joint(const TreeDev &T, const LevelDev &L, int store_tip_cx, hipStream_t s) {
    if (L.n_nodes <= 0) return hipSuccess;
    const int npt = n_ptiles(T.P);
    const int groups = (L.n_chunks + 3) / 4;
    const dim3 grid(groups * npt), block(TT_BLOCK);
    switch (T.n) {  // wide alphabets: the branch matrix from LDS
        case 20: hipLaunchKernelGGL(k_post_joint_wide<20>, grid, block, 0, s, T, L, npt, store_tip_cx); break;
        case 21: hipLaunchKernelGGL(k_post_joint_wide<21>, grid, block, 0, s, T, L, npt, store_tip_cx); break;
        case 22: hipLaunchKernelGGL(k_post_joint_wide<22>, grid, block, 0, s, T, L, npt, store_tip_cx); break;
        default:
            TT_DISPATCH(T.n, hipLaunchKernelGGL(k_post_joint<N_>, grid, block, 0, s, T, L, npt, store_tip_cx));
    }
    return hipGetLastError();
}

// 20-state (amino-acid) level kernels on FP32 MFMA (v_mfma_f32_32x32x2f32).
//
// A node's messages for 32 patterns are one MFMA product, computed
// transposed: states x patterns = M^T (32 x 20, padded) times R^T (20 x 32
// patterns).  In the 32x32 accumulator lane l holds column (pattern) l % 32
// and the 16 rows 8 * (r / 4) + 4 * h + r % 4 (slot r, h = l / 32).  The
// output rows are assigned to states so that slot r < 10 of half h holds
// state 10 * h + r and slots 10..15 are padding rows (zero A rows): the two
// lanes l and l ^ 32 own one pattern's 20 states, ten each, as one
// contiguous 40-byte half of the pattern's 80-byte row.
//
// Every vector lives in that slot layout, the inputs included: step s of a
// contraction takes input state 10 * h + s from lane half h, which is the
// value the lane already holds in slot s.  A contraction is therefore 10
// MFMA steps with no data exchanged between lanes, and per-pattern
// reductions (argmax, sums) are a within-lane pass over 10 slots plus one
// exchange with the partner lane.  Operand layouts of the instruction: lane
// l supplies A[row l % 32][k = l / 32] and B[k = l / 32][col l % 32].
//
// Launch shapes:
// * wide levels (k_post20 / k_pre20): a wave takes one node and a strided
//   set of its 32-pattern tiles; the branch matrix (the A operand, 10
//   registers per contraction) is loaded once per wave, and the next tile's
//   rows are fetched while the current tile's MFMA chain runs;
// * runs of narrow levels (k_post_run20 / k_pre_run20): a block of 8 waves
//   takes one 32-pattern tile through every level of the run, one node per
//   wave per level, a block barrier between levels.
// Tip rows come from the character profile table, staged in LDS when it
// has at most 64 codes.
//
// Arithmetic follows tt_device.h: deficit messages m = r + d with
// d = r (P - I), the normaliser factor formed in FP64 (LogAcc), the stored
// row scaled by the correctly rounded reciprocal of its maximum; preorder by
// the division trick with the reciprocal estimate of the VALU kernels.
// Polytomies sum per-state log messages over their children (the
// reference's own log-space form, treeanc.py:857-878).
#include "tt_device.h"

#include <algorithm>

namespace {
using namespace ttd;

typedef float v16f __attribute__((ext_vector_type(16)));
constexpr int NS = 20;
constexpr int kHalf = 10;      // states per lane half = MFMA steps per contraction
constexpr int kTilesMax = 8;   // 32-pattern tiles per wave in the wide postorder kernel
constexpr int kLutLds = 64;    // character codes whose profile rows are staged in LDS
constexpr int kRunWaves = 8;   // waves per block of the narrow-run kernels

// LDS pointers carry their address space, so that a tip's table row and an
// internal child's registers are never merged into one generic load.
#define TT_L __attribute__((address_space(3)))

// State of accumulator row i (-1: padding row).
__device__ __forceinline__ int row_state(int i) {
    const int hh = (i >> 2) & 1, r = ((i >> 3) << 2) | (i & 3);
    return r < kHalf ? kHalf * hh + r : -1;
}

// This lane's half (10 states at 10 h) of an 80-byte row.
template <class Ptr>
__device__ __forceinline__ void load_half(Ptr row, int h, float (&o)[kHalf]) {
    const float2 TT_G *s = reinterpret_cast<const float2 TT_G *>(row + kHalf * h);
#pragma unroll
    for (int q = 0; q < kHalf / 2; ++q) {
        const float2 v = s[q];
        o[2 * q] = v.x;
        o[2 * q + 1] = v.y;
    }
}

template <class Ptr>
__device__ __forceinline__ void store_half(Ptr row, int h, const float (&o)[kHalf]) {
    float2 TT_G *d = reinterpret_cast<float2 TT_G *>(row + kHalf * h);
#pragma unroll
    for (int q = 0; q < kHalf / 2; ++q) d[q] = make_float2(o[2 * q], o[2 * q + 1]);
}

// A operand of the 10 steps for accumulator row i (state si, -1 padding).
struct AOp {
    float a[kHalf];
};

// out_i = sum_k x_k M[k][i]: A[i][k] = M[k * NS + i].
__device__ __forceinline__ void load_a_in(const float TT_C *M, int si, int h, AOp &A) {
    const int col = si >= 0 ? si : 0;
#pragma unroll
    for (int s = 0; s < kHalf; ++s) {
        const float v = M[(kHalf * h + s) * NS + col];
        A.a[s] = si >= 0 ? v : 0.0f;
    }
}

// out_i = sum_k M[i][k] x_k: A[i][k] = M[i * NS + k].
__device__ __forceinline__ void load_a_out(const float TT_C *M, int si, int h, AOp &A) {
    const int row = si >= 0 ? si : 0;
#pragma unroll
    for (int s = 0; s < kHalf; ++s) {
        const float v = M[row * NS + kHalf * h + s];
        A.a[s] = si >= 0 ? v : 0.0f;
    }
}

__device__ __forceinline__ v16f contract(const AOp &A, const float (&x)[kHalf]) {
    v16f acc = {0.0f};
#pragma unroll
    for (int s = 0; s < kHalf; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(A.a[s], x[s], acc, 0, 0, 0);
    return acc;
}

// Raw child data of one tile: the lane's half of an internal child's row,
// or a tip's character code (bits of slot 0).  Loads are unconditional at a
// clamped pattern index, so their number does not depend on the tile.
template <bool TIP>
__device__ __forceinline__ void fetch_kind(const TreeDev &T, int cref, int pc, int h, float (&raw)[kHalf]) {
    if constexpr (TIP) {
        raw[0] = __int_as_float((int)T.tip_codes[(size_t)(-cref - 1) * T.P + pc]);
#pragma unroll
        for (int s = 1; s < kHalf; ++s) raw[s] = 0.0f;
    } else {
        load_half(T.sub + ((size_t)cref * T.P + pc) * NS, h, raw);
    }
}

template <bool TIP, class LP>
__device__ __forceinline__ void expand_kind(const float (&raw)[kHalf], LP lut, int h, float (&x)[kHalf]) {
    if constexpr (TIP) {
        const int base = __float_as_int(raw[0]) * NS + kHalf * h;
#pragma unroll
        for (int s = 0; s < kHalf; ++s) x[s] = lut[base + s];
    } else {
#pragma unroll
        for (int s = 0; s < kHalf; ++s) x[s] = raw[s];
    }
}

__device__ __forceinline__ void fetch_child(const TreeDev &T, int cref, int pc, int h, float (&raw)[kHalf]) {
    if (cref < 0)
        fetch_kind<true>(T, cref, pc, h, raw);
    else
        fetch_kind<false>(T, cref, pc, h, raw);
}

template <class LP>
__device__ __forceinline__ void expand_child(int cref, const float (&raw)[kHalf], LP lut, int h, float (&x)[kHalf]) {
    if (cref < 0) {
        const int base = __float_as_int(raw[0]) * NS + kHalf * h;
#pragma unroll
        for (int s = 0; s < kHalf; ++s) x[s] = lut[base + s];
    } else {
#pragma unroll
        for (int s = 0; s < kHalf; ++s) x[s] = raw[s];
    }
}

__device__ __forceinline__ double pair_pick(double mine_value, bool mine) {
    const double other = __shfl_xor(mine_value, 32, TT_WAVE);
    return mine ? mine_value : other;
}

// First maximum over the pattern's 20 states held by lanes l and l ^ 32.
__device__ __forceinline__ int pair_argmax(const float (&x)[kHalf], int h, float &best, bool &mine) {
    best = x[0];
    int bi = kHalf * h;
#pragma unroll
    for (int s = 1; s < kHalf; ++s)
        if (x[s] > best) {
            best = x[s];
            bi = kHalf * h + s;
        }
    const float ob = __shfl_xor(best, 32, TT_WAVE);
    const int obi = __shfl_xor(bi, 32, TT_WAVE);
    mine = !(ob > best || (ob == best && obi < bi));
    if (!mine) best = ob;
    return mine ? bi : obi;
}

// Postorder step of binary row v for one tile (children already fetched);
// returns the FP64 normaliser factor of the pattern.
// FULL: every lane's pattern exists (stores without a lane mask, so the
// number of memory operations in flight does not depend on the tile);
// OFFS: the per-node log normaliser is also stored (update mode).
template <bool TIP0, bool TIP1, bool FULL, bool OFFS, class LP>
__device__ __forceinline__ double post_bin20(const TreeDev &T, int v, int p, bool act, int h, const AOp &A0,
                                             const AOp &A1, const float (&raw0)[kHalf], const float (&raw1)[kHalf],
                                             LP lut) {
    float r0[kHalf], r1[kHalf];
    expand_kind<TIP0>(raw0, lut, h, r0);
    expand_kind<TIP1>(raw1, lut, h, r1);
    const v16f d0 = contract(A0, r0);
    const v16f d1 = contract(A1, r1);
    float x[kHalf];
#pragma unroll
    for (int s = 0; s < kHalf; ++s) x[s] = (r0[s] + d0[s]) * (r1[s] + d1[s]);
    // first maximum of this lane's states; the children's (r, d) ride along
    float best = x[0], a0 = r0[0], b0 = d0[0], a1 = r1[0], b1 = d1[0];
    int bi = kHalf * h;
#pragma unroll
    for (int s = 1; s < kHalf; ++s)
        if (x[s] > best) {
            best = x[s];
            bi = kHalf * h + s;
            a0 = r0[s];
            b0 = d0[s];
            a1 = r1[s];
            b1 = d1[s];
        }
    const float ob = __shfl_xor(best, 32, TT_WAVE);
    const int obi = __shfl_xor(bi, 32, TT_WAVE);
    const bool mine = !(ob > best || (ob == best && obi < bi));
    const int jm = mine ? bi : obi;
    const float xm = mine ? best : ob;
    const double f = pair_pick(msg64(a0, b0) * msg64(a1, b1), mine);
    if constexpr (OFFS) {
        if (act && h == 0) T.offs[(size_t)v * T.P + p] = log(f);
    }
    // entries times the correctly rounded reciprocal of the maximum (within
    // 1.5 ulp of the quotient, unbiased; the maximum itself is exactly 1)
    const float inv = 1.0f / xm;
    float o[kHalf];
#pragma unroll
    for (int s = 0; s < kHalf; ++s) o[s] = (kHalf * h + s == jm) ? 1.0f : x[s] * inv;
    if (FULL || act) store_half(T.sub + ((size_t)v * T.P + p) * NS, h, o);
    return f;
}

// Postorder step of a row with any number of children, one tile.
template <class LP>
__device__ __forceinline__ double post_poly20(const TreeDev &T, const PostDesc &dd, int p, bool act, int h, int si,
                                              LP lut) {
    const int pc = p < T.P ? p : T.P - 1;
    const int e0 = dd.e0, e1 = dd.e0 + dd.nchild;
    float L[kHalf];
#pragma unroll
    for (int s = 0; s < kHalf; ++s) L[s] = 0.0f;
    for (int e = e0; e < e1; ++e) {
        const int cref = wave_uniform(T.child_ref[e]);
        AOp A;
        load_a_in(T.D32 + (size_t)wave_uniform(T.child_node[e]) * (NS * NS), si, h, A);
        float raw[kHalf], r[kHalf];
        fetch_child(T, cref, pc, h, raw);
        expand_child(cref, raw, lut, h, r);
        const v16f d = contract(A, r);
#pragma unroll
        for (int s = 0; s < kHalf; ++s) L[s] += log_msg(r[s], d[s]);
    }
    float best;
    bool mine;
    const int jm = pair_argmax(L, h, best, mine);
    // normaliser at the selected state, children revisited (cache hits)
    double f = 1.0, lnode = 0.0;
    for (int e = e0; e < e1; ++e) {
        const int cref = wave_uniform(T.child_ref[e]);
        AOp A;
        load_a_in(T.D32 + (size_t)wave_uniform(T.child_node[e]) * (NS * NS), si, h, A);
        float raw[kHalf], r[kHalf];
        fetch_child(T, cref, pc, h, raw);
        expand_child(cref, raw, lut, h, r);
        const v16f d = contract(A, r);
        float rj = r[0], dj = d[0];
#pragma unroll
        for (int s = 1; s < kHalf; ++s)
            if (kHalf * h + s == jm) {
                rj = r[s];
                dj = d[s];
            }
        const double m = pair_pick(msg64(rj, dj), jm >= kHalf * h && jm < kHalf * h + kHalf);
        f *= m;
        if (T.offs) lnode += log(m);
    }
    if (T.offs && act && h == 0) T.offs[(size_t)dd.v * T.P + p] = lnode;
    float o[kHalf];
#pragma unroll
    for (int s = 0; s < kHalf; ++s) o[s] = (kHalf * h + s == jm) ? 1.0f : expf(L[s] - best);
    if (act) store_half(T.sub + ((size_t)dd.v * T.P + p) * NS, h, o);
    return f;
}

// One binary postorder node over tiles g, g + G, ... (nt of them); the
// normaliser factor of tile j goes to sf[j][pattern lane].  Two register
// sets alternate over the full tiles: the next tile's rows are in flight
// while the current tile's MFMA chain runs, and no copy between the sets
// waits for them.  A partial last tile (lanes past the last pattern) runs
// after the loop with masked stores.
template <bool TIP0, bool TIP1, class LP>
__device__ __forceinline__ void post_bin_tiles20(const TreeDev &T, const PostDesc &dd, int npt, int g, int G, int nt,
                                                 int h, int i, int si, LP lut, double (*sf)[32]) {
    AOp A0, A1;
    load_a_in(T.D32 + (size_t)dd.cnode0 * (NS * NS), si, h, A0);
    load_a_in(T.D32 + (size_t)dd.cnode1 * (NS * NS), si, h, A1);
    const bool partial = (T.P & 31) != 0 && g + (nt - 1) * G == npt - 1;
    const int nf = nt - (partial ? 1 : 0);
    auto tile_p = [&](int j) { return (g + min(j, nt - 1) * G) * 32 + i; };
    float a0[kHalf], a1[kHalf], b0[kHalf], b1[kHalf];
    {
        const int pc = min(tile_p(0), T.P - 1);
        fetch_kind<TIP0>(T, dd.cref0, pc, h, a0);
        fetch_kind<TIP1>(T, dd.cref1, pc, h, a1);
    }
    for (int j = 0; j < nf; j += 2) {
        {
            const int pc = min(tile_p(j + 1), T.P - 1);
            fetch_kind<TIP0>(T, dd.cref0, pc, h, b0);
            fetch_kind<TIP1>(T, dd.cref1, pc, h, b1);
        }
        sf[j][i] = post_bin20<TIP0, TIP1, true, false>(T, dd.v, tile_p(j), true, h, A0, A1, a0, a1, lut);
        if (j + 1 >= nf) {
#pragma unroll
            for (int s = 0; s < kHalf; ++s) {  // the partial tile's rows, if any, were fetched into b
                a0[s] = b0[s];
                a1[s] = b1[s];
            }
            break;
        }
        {
            const int pc = min(tile_p(j + 2), T.P - 1);
            fetch_kind<TIP0>(T, dd.cref0, pc, h, a0);
            fetch_kind<TIP1>(T, dd.cref1, pc, h, a1);
        }
        sf[j + 1][i] = post_bin20<TIP0, TIP1, true, false>(T, dd.v, tile_p(j + 1), true, h, A0, A1, b0, b1, lut);
    }
    if (partial) {
        const int p = tile_p(nt - 1);
        sf[nt - 1][i] = post_bin20<TIP0, TIP1, false, false>(T, dd.v, p, p < T.P, h, A0, A1, a0, a1, lut);
    }
}

// Update mode (per-node normalisers stored): one tile after the other.
template <bool TIP0, bool TIP1, class LP>
__device__ __forceinline__ void post_bin_tiles20_offs(const TreeDev &T, const PostDesc &dd, int g, int G, int nt,
                                                      int h, int i, int si, LP lut, double (*sf)[32]) {
    AOp A0, A1;
    load_a_in(T.D32 + (size_t)dd.cnode0 * (NS * NS), si, h, A0);
    load_a_in(T.D32 + (size_t)dd.cnode1 * (NS * NS), si, h, A1);
    for (int j = 0; j < nt; ++j) {
        const int p = (g + j * G) * 32 + i, pc = min(p, T.P - 1);
        float c0[kHalf], c1[kHalf];
        fetch_kind<TIP0>(T, dd.cref0, pc, h, c0);
        fetch_kind<TIP1>(T, dd.cref1, pc, h, c1);
        sf[j][i] = post_bin20<TIP0, TIP1, false, true>(T, dd.v, p, p < T.P, h, A0, A1, c0, c1, lut);
    }
}

template <class LP>
__device__ __forceinline__ void post_tiles20(const TreeDev &T, const PostDesc &dd, int npt, int g, int G, int nt,
                                             int h, int i, int si, LP lut, double (*sf)[32]) {
    if (dd.nchild != 2) {
        for (int j = 0; j < nt; ++j) {
            const int p = (g + j * G) * 32 + i;
            sf[j][i] = post_poly20(T, dd, p, p < T.P, h, si, lut);
        }
        return;
    }
    auto go = [&](auto tip0, auto tip1) {
        constexpr bool t0 = decltype(tip0)::value, t1 = decltype(tip1)::value;
        if (T.offs)
            post_bin_tiles20_offs<t0, t1>(T, dd, g, G, nt, h, i, si, lut, sf);
        else
            post_bin_tiles20<t0, t1>(T, dd, npt, g, G, nt, h, i, si, lut, sf);
    };
    using Y = std::true_type;
    using N = std::false_type;
    if (dd.cref0 < 0) {
        if (dd.cref1 < 0)
            go(Y{}, Y{});
        else
            go(Y{}, N{});
    } else {
        if (dd.cref1 < 0)
            go(N{}, Y{});
        else
            go(N{}, N{});
    }
}

__device__ __forceinline__ void stage_lut(const TreeDev &T, float *s_lut) {
    if (T.n_codes <= kLutLds)
        for (int q = threadIdx.x; q < T.n_codes * NS; q += blockDim.x) s_lut[q] = T.lut[q];
    __syncthreads();
}

// Wide postorder level: block = 4 consecutive nodes of the level (one per
// wave) x tile group g; the block writes the level's partial row of its node
// group for its tiles.
__global__ __launch_bounds__(TT_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_post20(TreeDev T, LevelDev L, int npt, int G) {
    constexpr int kW = TT_BLOCK / TT_WAVE;
    __shared__ float s_lut[kLutLds * NS];
    __shared__ double s_f[kW][kTilesMax][32];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int h = lane >> 5, i = lane & 31, si = row_state(i);
    const int ng = blockIdx.x / G, g = blockIdx.x % G;
    const int nt = (npt - g + G - 1) / G;
    stage_lut(T, s_lut);
    const int k = ng * kW + w;
    if (k < L.n_nodes) {
        const PostDesc dd = load_post_desc(L.desc + kPostDesc * wave_uniform(k));
        if (T.n_codes <= kLutLds)
            post_tiles20(T, dd, npt, g, G, nt, h, i, si, (const float TT_L *)s_lut, s_f[w]);
        else
            post_tiles20(T, dd, npt, g, G, nt, h, i, si, T.lut, s_f[w]);
    } else if (h == 0) {
        for (int j = 0; j < nt; ++j) s_f[w][j][i] = 1.0;
    }
    __syncthreads();
    for (int j = w; j < nt; j += kW) {
        const int p = (g + j * G) * 32 + i;
        if (h == 0 && p < T.P) {
            LogAcc acc;
#pragma unroll
            for (int q = 0; q < kW; ++q) acc.mul(s_f[q][j][i]);
            T.part[(size_t)(L.row_off + ng) * T.P + p] = acc.log_value();
        }
    }
}

// Preorder step (treeanc.py:889-930) of node c for one tile, rows fetched.
template <bool TIP, bool FULL, class LP>
__device__ __forceinline__ void pre_tile20(const TreeDev &T, int c, int p, bool act, int h, const AOp &Ain,
                                           const AOp &Aout, const float (&raw)[kHalf], const float (&pp)[kHalf],
                                           LP lut) {
    float ro[kHalf];
    expand_kind<TIP>(raw, lut, h, ro);
    // child message m_j = sum_i row_i P[i][j]
    const v16f m = contract(Ain, ro);
    // outgroup q_j = max(TINY, parent_j) / m_j (its normalisation cancels)
    float q[kHalf];
#pragma unroll
    for (int s = 0; s < kHalf; ++s) q[s] = fmaxf((float)TT_TINY, pp[s]) * __builtin_amdgcn_rcpf(m[s]);
    // evolve: msg_k = sum_j P[k][j] q_j
    const v16f msg = contract(Aout, q);
    float x[kHalf];
    float part = 0.0f;
#pragma unroll
    for (int s = 0; s < kHalf; ++s) {
        x[s] = msg[s] * ro[s];
        part += x[s];
    }
    const float tot = part + __shfl_xor(part, 32, TT_WAVE);
    float best;
    bool mine;
    const int jm = pair_argmax(x, h, best, mine);
    const float xi = __builtin_amdgcn_rcpf(tot);
    float o[kHalf];
#pragma unroll
    for (int s = 0; s < kHalf; ++s) o[s] = x[s] * xi;
    // both lanes of a pattern store its state (the same byte)
    if (FULL || act) {
        store_half(T.prof + ((size_t)c * T.P + p) * NS, h, o);
        T.states[(size_t)c * T.P + p] = (uint8_t)jm;
    }
}

template <bool TIP>
__device__ __forceinline__ void pre_fetch20(const TreeDev &T, const PreDesc &qd, int cref, int pc, int h,
                                            float (&raw)[kHalf], float (&pp)[kHalf]) {
    fetch_kind<TIP>(T, cref, pc, h, raw);
    load_half(T.prof + ((size_t)qd.par * T.P + pc) * NS, h, pp);
}

// One preorder node over tiles g, g + G, ... (two alternating register sets
// over the full tiles, the partial last tile after the loop, as
// post_bin_tiles20).
template <bool TIP, class LP>
__device__ __forceinline__ void pre_kind_tiles20(const TreeDev &T, const PreDesc &qd, int npt, int g, int G, int nt,
                                                 int h, int i, int si, LP lut) {
    const float TT_C *M = T.P32 + (size_t)qd.c * (NS * NS);
    AOp Ain, Aout;
    load_a_in(M, si, h, Ain);
    load_a_out(M, si, h, Aout);
    const int cref = qd.ir >= 0 ? qd.ir : -(qd.tr + 1);
    const bool partial = (T.P & 31) != 0 && g + (nt - 1) * G == npt - 1;
    const int nf = nt - (partial ? 1 : 0);
    auto tile_p = [&](int j) { return (g + min(j, nt - 1) * G) * 32 + i; };
    float ac[kHalf], ap[kHalf], bc[kHalf], bp[kHalf];
    pre_fetch20<TIP>(T, qd, cref, min(tile_p(0), T.P - 1), h, ac, ap);
    for (int j = 0; j < nf; j += 2) {
        pre_fetch20<TIP>(T, qd, cref, min(tile_p(j + 1), T.P - 1), h, bc, bp);
        pre_tile20<TIP, true>(T, qd.c, tile_p(j), true, h, Ain, Aout, ac, ap, lut);
        if (j + 1 >= nf) {
#pragma unroll
            for (int s = 0; s < kHalf; ++s) {
                ac[s] = bc[s];
                ap[s] = bp[s];
            }
            break;
        }
        pre_fetch20<TIP>(T, qd, cref, min(tile_p(j + 2), T.P - 1), h, ac, ap);
        pre_tile20<TIP, true>(T, qd.c, tile_p(j + 1), true, h, Ain, Aout, bc, bp, lut);
    }
    if (partial) {
        const int p = tile_p(nt - 1);
        pre_tile20<TIP, false>(T, qd.c, p, p < T.P, h, Ain, Aout, ac, ap, lut);
    }
}

template <class LP>
__device__ __forceinline__ void pre_tiles20(const TreeDev &T, const PreDesc &qd, int npt, int g, int G, int nt, int h,
                                            int i, int si, LP lut) {
    if (qd.ir >= 0)
        pre_kind_tiles20<false>(T, qd, npt, g, G, nt, h, i, si, lut);
    else
        pre_kind_tiles20<true>(T, qd, npt, g, G, nt, h, i, si, lut);
}

// Wide preorder level: wave = (node, tile group).
__global__ __launch_bounds__(TT_BLOCK) void k_pre20(TreeDev T, LevelDev L, int npt, int G) {
    __shared__ float s_lut[kLutLds * NS];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int h = lane >> 5, i = lane & 31, si = row_state(i);
    stage_lut(T, s_lut);
    const int task = blockIdx.x * (TT_BLOCK / TT_WAVE) + w;
    if (task >= L.n_nodes * G) return;  // whole wave exits together
    const int k = task / G, g = task % G;
    const int nt = (npt - g + G - 1) / G;
    const PreDesc qd = load_pre_desc(L.desc + kPreDesc * wave_uniform(k));
    if (T.n_codes <= kLutLds)
        pre_tiles20(T, qd, npt, g, G, nt, h, i, si, (const float TT_L *)s_lut);
    else
        pre_tiles20(T, qd, npt, g, G, nt, h, i, si, T.lut);
}

// One postorder node of a narrow run for this block's tile.
template <class LP>
__device__ __forceinline__ double post_node20(const TreeDev &T, const PostDesc &dd, int p, int h, int si, LP lut) {
    const bool act = p < T.P;
    if (dd.nchild != 2) return post_poly20(T, dd, p, act, h, si, lut);
    AOp A0, A1;
    load_a_in(T.D32 + (size_t)dd.cnode0 * (NS * NS), si, h, A0);
    load_a_in(T.D32 + (size_t)dd.cnode1 * (NS * NS), si, h, A1);
    float c0[kHalf], c1[kHalf];
    const int pc = act ? p : T.P - 1;
    auto go = [&](auto tip0, auto tip1) {
        fetch_kind<decltype(tip0)::value>(T, dd.cref0, pc, h, c0);
        fetch_kind<decltype(tip1)::value>(T, dd.cref1, pc, h, c1);
        constexpr bool t0 = decltype(tip0)::value, t1 = decltype(tip1)::value;
        if (T.offs) return post_bin20<t0, t1, false, true>(T, dd.v, p, act, h, A0, A1, c0, c1, lut);
        return post_bin20<t0, t1, false, false>(T, dd.v, p, act, h, A0, A1, c0, c1, lut);
    };
    using Y = std::true_type;
    using N = std::false_type;
    if (dd.cref0 < 0) return dd.cref1 < 0 ? go(Y{}, Y{}) : go(Y{}, N{});
    return dd.cref1 < 0 ? go(N{}, Y{}) : go(N{}, N{});
}

// Runs of narrow levels in one launch (see k_post_run): a block of 8 waves
// takes one 32-pattern tile through every level of the run, the level's
// nodes dealt to the waves, a block barrier between levels.
template <class LP>
__device__ __forceinline__ void post_run_levels(const TreeDev &T, const RunDev &R, int p, int h, int si, int w,
                                                LP lut, LogAcc &acc) {
    for (int l = R.l0; l < R.l1; ++l) {
        const int kb = R.off[l], ke = R.off[l + 1];
        for (int k = wave_uniform(kb + w); k < ke; k += kRunWaves)
            acc.mul(post_node20(T, load_post_desc(R.desc + kPostDesc * k), p, h, si, lut));
        __syncthreads();
    }
}

__global__ __launch_bounds__(kRunWaves * TT_WAVE) void k_post_run20(TreeDev T, RunDev R) {
    __shared__ float s_lut[kLutLds * NS];
    __shared__ double s_acc[kRunWaves][32];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int h = lane >> 5, i = lane & 31, si = row_state(i);
    const int p = blockIdx.x * 32 + i;
    stage_lut(T, s_lut);
    LogAcc acc;
    if (T.n_codes <= kLutLds)
        post_run_levels(T, R, p, h, si, w, (const float TT_L *)s_lut, acc);
    else
        post_run_levels(T, R, p, h, si, w, T.lut, acc);
    if (h == 0) s_acc[w][i] = acc.log_value();
    __syncthreads();
    if (w == 0 && h == 0 && p < T.P) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < kRunWaves; ++q) s += s_acc[q][i];
        T.part[(size_t)R.row_first * T.P + p] = s;
        for (int r = R.row_first + 1; r < R.row_end; ++r) T.part[(size_t)r * T.P + p] = 0.0;
    }
}

template <class LP>
__device__ __forceinline__ void pre_run_levels(const TreeDev &T, const RunDev &R, int p, int h, int si, int w,
                                               LP lut) {
    const bool act = p < T.P;
    const int pc = act ? p : T.P - 1;
    for (int l = R.l0; l < R.l1; ++l) {
        const int kb = R.off[l];
        const int ke = R.tips ? R.off[l + 1] : kb + R.nint[l];
        for (int k = wave_uniform(kb + w); k < ke; k += kRunWaves) {
            const PreDesc qd = load_pre_desc(R.desc + kPreDesc * k);
            const float TT_C *M = T.P32 + (size_t)qd.c * (NS * NS);
            AOp Ain, Aout;
            load_a_in(M, si, h, Ain);
            load_a_out(M, si, h, Aout);
            float c[kHalf], pp[kHalf];
            if (qd.ir >= 0) {
                pre_fetch20<false>(T, qd, qd.ir, pc, h, c, pp);
                pre_tile20<false, false>(T, qd.c, p, act, h, Ain, Aout, c, pp, lut);
            } else {
                pre_fetch20<true>(T, qd, -(qd.tr + 1), pc, h, c, pp);
                pre_tile20<true, false>(T, qd.c, p, act, h, Ain, Aout, c, pp, lut);
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kRunWaves * TT_WAVE) void k_pre_run20(TreeDev T, RunDev R) {
    __shared__ float s_lut[kLutLds * NS];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int h = lane >> 5, i = lane & 31, si = row_state(i);
    const int p = blockIdx.x * 32 + i;
    stage_lut(T, s_lut);
    if (T.n_codes <= kLutLds)
        pre_run_levels(T, R, p, h, si, w, (const float TT_L *)s_lut);
    else
        pre_run_levels(T, R, p, h, si, w, T.lut);
}

// Tile groups per node of a wide level: enough waves to fill the device
// (about 8 per SIMD), at most kTilesMax tiles per wave.
int tile_groups(int n_nodes, int npt) {
    const int waves = std::max(1, n_nodes);
    int G = (8192 + waves - 1) / waves;
    G = std::max(G, (npt + kTilesMax - 1) / kTilesMax);
    return std::min(std::max(G, 1), npt);
}

}  // namespace

namespace ttk {

hipError_t post_run20(const TreeDev &T, const RunDev &R, hipStream_t s) {
    hipLaunchKernelGGL(k_post_run20, dim3((T.P + 31) / 32), dim3(kRunWaves * TT_WAVE), 0, s, T, R);
    return hipGetLastError();
}

hipError_t pre_run20(const TreeDev &T, const RunDev &R, hipStream_t s) {
    hipLaunchKernelGGL(k_pre_run20, dim3((T.P + 31) / 32), dim3(kRunWaves * TT_WAVE), 0, s, T, R);
    return hipGetLastError();
}

// The level's partial rows are one per 4 nodes (the level was set up with
// one node per chunk for this kernel).
hipError_t post_marginal20(const TreeDev &T, const LevelDev &L, hipStream_t s) {
    if (L.chunk != 1) return hipErrorInvalidValue;
    const int npt = (T.P + 31) / 32;
    const int groups = (L.n_nodes + 3) / 4;
    const int G = tile_groups(groups * 4, npt);
    hipLaunchKernelGGL(k_post20, dim3(groups * G), dim3(TT_BLOCK), 0, s, T, L, npt, G);
    return hipGetLastError();
}

hipError_t pre_marginal20(const TreeDev &T, const LevelDev &L, hipStream_t s) {
    if (L.n_nodes <= 0) return hipSuccess;
    const int npt = (T.P + 31) / 32;
    const int G = tile_groups(L.n_nodes, npt);
    const long tasks = (long)L.n_nodes * G;
    hipLaunchKernelGGL(k_pre20, dim3((unsigned)((tasks + 3) / 4)), dim3(TT_BLOCK), 0, s, T, L, npt, G);
    return hipGetLastError();
}

}  // namespace ttk

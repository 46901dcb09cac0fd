// One-launch marginal (or log-likelihood) pass of one tree with its internal
// rows resident in LDS.
//
// The level kernels pay a launch per level, and a pass over a small or
// medium tree is a chain of ~2 x depth dependent levels whose work is a few
// microseconds in all.  Here a workgroup takes TP patterns through the whole
// tree: the internal rows of those patterns live in LDS (n_int x TP x n
// floats), a level is one or more rounds of the workgroup's threads over
// (node, pattern) pairs, and rounds are separated by a workgroup barrier
// instead of a launch and a round trip through HBM.  Patterns are
// independent, so workgroups never wait for each other; the last one to
// finish sums the per-pattern log-likelihoods (fixed order) into the total.
//
// What a thread needs from memory for one (node, pattern) step -- the
// node's descriptor and its branches' P(t) or P - I -- is packed into one
// record per level entry (k_tile_pack, refreshed whenever P(t) changes).
// When they fit, a phase's records are staged in LDS before its first round
// (the postorder's at the start, the preorder's during the root step), so
// the rounds issue no global load at all: on this architecture loads and
// stores share one counter that completes in order, and a load waited on in
// every round would also wait for the previous round's result stores to be
// acknowledged.  Otherwise the records stream from memory.  Either way the
// record of a thread's next round is read into registers while the current
// round runs, and a round costs LDS traffic, arithmetic and a barrier.
//
// Rows are written to HBM as well (subtree rows, profiles, states: the
// results), but never read back (except, with reused slots, the subtree
// rows the preorder needs, read a round ahead).  The per-(node, pattern)
// arithmetic is the level kernels' own device code (tt_device.h) on values
// read from LDS and registers, so profiles and states are identical to
// theirs; the per-pattern log normalisers are summed per thread and then
// across the threads of a pattern (wave shuffles, then one value per
// wave), so logL agrees to rounding.  The workgroups' sums meet in a
// fence-free partial/count protocol (tile_total).
//
// Postorder levels are node heights (treeanc.py:840-878), the root step
// treeanc.py:814-829, preorder levels node depths (treeanc.py:880-932): a
// node's slot first holds its subtree row and, once its own preorder step
// is done, its marginal profile, which its children read as the parent's.
#include "tt_device.h"

namespace {
using namespace ttd;

constexpr int kTileThreads = 512;

__host__ __device__ inline size_t align4(size_t x) { return (x + 3) & ~(size_t)3; }

// record sizes in 32-bit words, multiples of 4 (16-byte loads)
__host__ __device__ constexpr int post_rec_words(int n) { return (8 + 2 * n * n + 3) & ~3; }
__host__ __device__ constexpr int pre_rec_words(int n) { return (8 + n * n + 3) & ~3; }

struct TileDev {
    const int32_t *post_rec;    // per height-ordered internal row: {row, slot, children, child 0}, {child 1, ..}, P - I of both children
    const int32_t *pre_rec;     // per depth-ordered node: {node, parent's slot, own slot, tip row}, {row, row's slot}, P(t)
    const int32_t *rounds;      // rounds as [first, end) entry pairs: post [Rp], preorder internal [Ri], all [Ra]
    int32_t Rp, Ri, Ra;
    int32_t TP;                 // patterns per workgroup (power of two)
    int32_t n_blocks;
    int32_t n_post, n_pre;      // records of each phase (staged into LDS when the kernel is built STAGED)
    int32_t n_slots;            // LDS rows per pattern
    int32_t root_post, root_pre;  // the root's slot in each phase
    const int32_t *post_slot;   // slot of every internal row (polytomies' child lists), or null: slot = row
    double *partial;            // [n_blocks] sum over the workgroup's patterns of mult * logL
    unsigned *counter;          // workgroups finished; reset to 0 by the last one
    int32_t pre_off = -1;       // STAGED: word offset of the preorder's records in the staged region when
                                // they have their own (staged at the start), or -1: they replace the
                                // postorder's at the root step
};

// Workgroup barrier that orders LDS only: rounds talk through LDS, and their
// HBM writes (results nobody in the launch reads back) need not drain first.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Sum of v over the 64 lanes of a wave, in lane 0 (a fixed shuffle tree).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = TT_WAVE / 2; off > 0; off >>= 1) v += __shfl_down(v, off);
    return v;
}

// Sum of v over the threads of one pattern (threads q, q + TP, q + 2 TP, ...,
// q = t & (TP - 1)), returned to threads t < TP: the lanes of a wave first
// (shuffles), then the waves' partials from LDS (s_red, one barrier), in a
// fixed order.  Every thread of the workgroup calls it.
__device__ __forceinline__ double pattern_sum(double v, double *s_red, int TP) {
    const int t = threadIdx.x;
    for (int off = TT_WAVE / 2; off >= TP; off >>= 1) v += __shfl_down(v, off);
    s_red[t] = v;
    lds_barrier();
    double s = 0.0;
    if (t < TP) {
        const int step = max(TP, TT_WAVE);  // between the partials of one pattern
#pragma unroll
        for (int j = 0; j < kTileThreads / TT_WAVE; ++j) {
            const int k = t + j * step;
            const double x = s_red[min(k, kTileThreads - 1)];
            s += k < kTileThreads ? x : 0.0;
        }
    }
    return s;
}

// The workgroup's sum of s_red[k] over its patterns k < TP that exist (mult
// * logL per pattern) goes to partial[blockIdx.x]; the last workgroup to
// count itself sums the partials (every thread a strided share, then wave
// shuffles and the waves' sums in a fixed order) into total[0], and puts
// the counter and the partials back to their empty state for the next pass.
// The partial and the count are device-scope atomics with no fence between
// them (a release here would write back the whole L2 of the die): the last
// workgroup may see the count before a partial, and then waits for that
// partial -- whose store was issued before its count -- in a bounded loop
// (a partial still empty at the bound makes the total a NaN).  Every
// thread of the workgroup calls it.
constexpr unsigned long long kPartialEmpty = ~0ull;  // the bits of an empty partial (a NaN)
constexpr int kPartialSpin = 1 << 16;

__device__ __forceinline__ void tile_total(const TreeDev &T, const TileDev &S, double *s_red, int TP) {
    __shared__ bool s_last;
    const int t = threadIdx.x;
    if (t < TT_WAVE) {
        double a = 0.0;
        for (int k = t; k < TP; k += TT_WAVE) a += (blockIdx.x * TP + k < T.P) ? s_red[k] : 0.0;
        a = wave_sum(a);
        if (t == 0) {
            __hip_atomic_store(&S.partial[blockIdx.x], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned done = __hip_atomic_fetch_add(S.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_last = done == (unsigned)S.n_blocks - 1;
        }
    }
    lds_barrier();
    if (s_last) {
        auto *bits = reinterpret_cast<unsigned long long *>(S.partial);
        double a = 0.0;
        for (int b = t; b < S.n_blocks; b += kTileThreads) {
            unsigned long long x = __hip_atomic_load(&bits[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (int spin = 0; x == kPartialEmpty && spin < kPartialSpin; ++spin)
                x = __hip_atomic_load(&bits[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            a += __builtin_bit_cast(double, x);
            __hip_atomic_store(&bits[b], kPartialEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        a = wave_sum(a);
        if ((t & (TT_WAVE - 1)) == 0) s_red[t / TT_WAVE] = a;
        lds_barrier();
        if (t == 0) {
            double tot = 0.0;
#pragma unroll
            for (int w = 0; w < kTileThreads / TT_WAVE; ++w) tot += s_red[w];
            T.total[0] = tot;
            __hip_atomic_store(S.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <int N>
struct PostRec {
    int4 a, b;  // v, e0, nchild, cref0 | cref1, cnode0, cnode1, vn
    float D0[N * N], D1[N * N];
};

template <int N>
struct PreRec {
    int4 a;  // node, parent's slot, own slot (or -1), tip row (or -1)
    int4 b;  // internal row (or -1), the slot holding its subtree row (when not read from memory)
    float M[N * N];
};

template <int N>
__device__ __forceinline__ void load_rec(const int32_t *base, PostRec<N> &r) {
    const int4 *s = reinterpret_cast<const int4 *>(base);
    r.a = s[0];
    r.b = s[1];
    const float *f = reinterpret_cast<const float *>(base + 8);
#pragma unroll
    for (int i = 0; i < N * N; ++i) {
        r.D0[i] = f[i];
        r.D1[i] = f[N * N + i];
    }
}

template <int N>
__device__ __forceinline__ void load_rec(const int32_t *base, PreRec<N> &r) {
    r.a = reinterpret_cast<const int4 *>(base)[0];
    r.b = reinterpret_cast<const int4 *>(base)[1];
    const float *f = reinterpret_cast<const float *>(base + 8);
#pragma unroll
    for (int i = 0; i < N * N; ++i) r.M[i] = f[i];
}

// dst[i] = f(i) for i in [start, n), dst in LDS, f a read from memory: each
// thread issues kStageLoads reads (clamped index) before it writes any, so
// staging an array costs one memory latency per kStageLoads * kTileThreads
// elements instead of one per kTileThreads.
constexpr int kStageLoads = 8;
template <class V, class F>
__device__ __forceinline__ void stage_in(V *dst, int n, F f, int start = 0) {
    const int t = threadIdx.x;
    for (int b = start; b < n; b += kStageLoads * kTileThreads) {
        V v[kStageLoads];
#pragma unroll
        for (int u = 0; u < kStageLoads; ++u) v[u] = f(min(b + u * kTileThreads + t, n - 1));
#pragma unroll
        for (int u = 0; u < kStageLoads; ++u) {
            const int i = b + u * kTileThreads + t;
            if (i < n) dst[i] = v[u];
        }
    }
}

// Staging in two phases, so that the first reads of several arrays are all
// in flight before any of them is written to LDS: load() issues a thread's
// first U reads (clamped index), store() writes them and stages whatever
// lies beyond the first U * kTileThreads elements.
template <class V, int U>
struct Stager {
    V v[U];
    template <class F>
    __device__ __forceinline__ void load(int n, F f) {
        if (n <= 0) return;
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = f(min(u * kTileThreads + (int)threadIdx.x, n - 1));
    }
    template <class F>
    __device__ __forceinline__ void store(V *dst, int n, F f) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = u * kTileThreads + (int)threadIdx.x;
            if (i < n) dst[i] = v[u];
        }
        stage_in(dst, n, f, U * kTileThreads);
    }
};

// Reader of the workgroup's tip characters: element r TP + j is the code of
// tip row r at pattern p0 + j (0 past the last pattern).
__device__ __forceinline__ auto tip_reader(const TreeDev &T, int TP, int p0) {
    return [&T, TP, p0](int i) {
        const int r = i / TP, pp = p0 + (i & (TP - 1));
        const uint8_t c = T.tip_codes[(size_t)r * T.P + min(pp, T.P - 1)];
        return pp < T.P ? c : (uint8_t)0;
    };
}

// STAGED: each phase's records copied into LDS; SUBG: slots are reused once
// their rows are dead (fewer rows per pattern in LDS), so the preorder reads
// each node's subtree row back from memory instead of from its slot.
template <int N, bool STAGED, bool SUBG>
__global__ __launch_bounds__(kTileThreads) void k_tile_marginal(TreeDev T, TileDev S, int tips, int pre) {
    extern __shared__ __align__(16) unsigned char s_dyn[];
    const int TP = S.TP;
    // LDS layout (tile_lds_bytes): rows, profile map, reduction, round
    // offsets, tip codes, staged records; each part starts on a 16-byte
    // boundary
    float *rows = reinterpret_cast<float *>(s_dyn);                                      // [n_slots][TP][N]
    float *s_lut = rows + align4((size_t)S.n_slots * TP * N);                            // [n_codes][N]
    double *s_red = reinterpret_cast<double *>(s_lut + align4((size_t)T.n_codes * N));  // [kTileThreads]
    int *s_rounds = reinterpret_cast<int *>(s_red + kTileThreads);
    const int n_rounds = 2 * (S.Rp + S.Ri + S.Ra);
    uint8_t *s_tip = reinterpret_cast<uint8_t *>(s_rounds + align4(n_rounds));  // [n_tips][TP]
    int32_t *s_rec = reinterpret_cast<int32_t *>(s_tip + (((size_t)T.n_tips * TP + 15) & ~(size_t)15));

    const int t = threadIdx.x;
    const int q = t & (TP - 1);  // this thread's pattern within the tile
    const int k0 = t / TP;       // this thread's entry within a round
    const int p = blockIdx.x * TP + q;
    const bool act = p < T.P;
    // read now what the root step needs from memory: a vector load there
    // would wait for every result store the postorder issued (loads and
    // stores complete in order)
    const double mult_p = T.mult[min(p, T.P - 1)];
    double pi_r[N];
#pragma unroll
    for (int j = 0; j < N; ++j) pi_r[j] = T.pi[j];

    // stage the profile map, the round offsets, the tile's tip characters and
    // the records (the preorder's too when they have their own region)
    const uint4 *pre4 = reinterpret_cast<const uint4 *>(S.pre_rec);
    auto f_pre = [&](int i) { return pre4[i]; };
    const int n_pre4 = S.n_pre * (pre_rec_words(N) / 4);
    {
        const int n_lut = T.n_codes * N, n_tipb = T.n_tips * TP;
        const int n_rec = STAGED ? S.n_post * (post_rec_words(N) / 4) : 0;
        const int n_pre_now = STAGED && pre && S.pre_off >= 0 ? n_pre4 : 0;
        auto f_lut = [&](int i) { return T.lut[i]; };
        auto f_rd = [&](int i) { return S.rounds[i]; };
        auto f_tip = tip_reader(T, TP, blockIdx.x * TP);
        const uint4 *rec4 = reinterpret_cast<const uint4 *>(S.post_rec);
        auto f_rec = [&](int i) { return rec4[i]; };
        Stager<float, 1> g_lut;
        Stager<int, 1> g_rd;
        Stager<uint8_t, kStageLoads> g_tip;
        Stager<uint4, kStageLoads> g_rec, g_pre;
        g_lut.load(n_lut, f_lut);
        g_rd.load(n_rounds, f_rd);
        g_tip.load(n_tipb, f_tip);
        g_rec.load(n_rec, f_rec);
        g_pre.load(n_pre_now, f_pre);
        g_lut.store(s_lut, n_lut, f_lut);
        g_rd.store(s_rounds, n_rounds, f_rd);
        g_tip.store(s_tip, n_tipb, f_tip);
        g_rec.store(reinterpret_cast<uint4 *>(s_rec), n_rec, f_rec);
        g_pre.store(reinterpret_cast<uint4 *>(s_rec + max(S.pre_off, 0)), n_pre_now, f_pre);
    }
    lds_barrier();  // the staging loads have landed in LDS (each thread waited for its own)

    auto slot = [&](int row) { return rows + ((size_t)row * TP + q) * N; };

    // ---- postorder: rounds over height-ordered internal rows ---------------------
    LogAcc acc;
    // a, b: the record's descriptor words; D0, D1: the children's P - I.
    // A binary entry runs in two steps: post_gather reads both children's
    // rows from LDS (a tip's row is its code's row of the profile map, the
    // same values bin_finish_m would read), then post_finish computes from
    // registers -- so that the next round's record can be requested between
    // the two, behind this round's reads.
    auto post_gather = [&](const int4 &a, const int4 &b, BinRaw<N> &R) {
        if (a.z != 2) return;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int cref = c ? b.x : a.w;
            const float *src = cref < 0 ? s_lut + s_tip[max(-cref - 1, 0) * TP + q] * N : slot(max(cref, 0));
#pragma unroll
            for (int i = 0; i < N; ++i) R.row[c][i] = src[i];
        }
    };
    auto post_finish = [&](const int4 &a, const int4 &b, auto D0, auto D1, const BinRaw<N> &R) {
        // row v, its slot vs; children as slots (>= 0) or tips (< 0)
        const int v = a.x, vs = a.y, nchild = a.z;
        if (nchild == 2) {
            bin_finish_m<N>(T, v, 0, 0, D0, D1, p, R, s_lut, acc, slot(vs));  // both rows gathered
        } else {  // polytomies and unary nodes read their child lists from memory
            auto rowf = [&](int cref, float (&row)[N]) {
                if (cref < 0) {
                    const int code = s_tip[(-cref - 1) * TP + q];
#pragma unroll
                    for (int i = 0; i < N; ++i) row[i] = s_lut[code * N + i];
                } else {
                    const float *src = slot(S.post_slot ? S.post_slot[cref] : cref);
#pragma unroll
                    for (int i = 0; i < N; ++i) row[i] = src[i];
                }
            };
            post_node_f<N>(T, v, p, s_lut, acc, rowf, [](int e) { return e; }, slot(vs));
        }
    };
    {
        constexpr int W = post_rec_words(N);
        const int *rd = s_rounds;
        // two record buffers in turn, loaded unconditionally (clamped index)
        // one round ahead -- from the staged copy in LDS or from memory -- so
        // a round's dependent reads start from its record in registers and
        // it waits only for the record it uses, never for the loads just
        // issued for the next round
        const int32_t *src;
        if constexpr (STAGED)
            src = s_rec;
        else
            src = S.post_rec;
        PostRec<N> ra, rb;
        // the staged copy is read only for the entries a thread runs (idle
        // waves take no LDS bandwidth); memory is read unconditionally, so
        // that the loads' counter never waits on a branch
        auto fetch = [&](int r, bool &valid, PostRec<N> &rec) {
            const int k = r < S.Rp ? rd[2 * r] + k0 : 0;
            valid = r < S.Rp && k < rd[2 * r + 1] && act;
            if (!STAGED || valid) load_rec<N>(src + (size_t)W * min(k, S.n_post - 1), rec);
        };
        bool va, vb;
        BinRaw<N> R;
        if (S.Rp > 0) fetch(0, va, ra);
        for (int r = 0; r < S.Rp; r += 2) {
            if (va) post_gather(ra.a, ra.b, R);
            fetch(r + 1, vb, rb);
            if (va) post_finish(ra.a, ra.b, ra.D0, ra.D1, R);
            lds_barrier();
            if (r + 1 >= S.Rp) break;
            if (vb) post_gather(rb.a, rb.b, R);
            fetch(r + 2, va, ra);
            if (vb) post_finish(rb.a, rb.b, rb.D0, rb.D1, R);
            lds_barrier();
        }
    }
    // the preorder's records replace the postorder's while the root step
    // runs, unless they were staged at the start
    if constexpr (STAGED) {
        if (pre && S.pre_off < 0) stage_in(reinterpret_cast<uint4 *>(s_rec), n_pre4, f_pre);
    }

    // ---- root (treeanc.py:814-829) ------------------------------------------
    const double off_sum = pattern_sum(act ? acc.log_value() : 0.0, s_red, TP);
    if (t < TP && act) {
        const double off = off_sum;
        const float *rr = slot(S.root_post);
        double x[N], s = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            x[j] = pi_r[j] * (double)rr[j];
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
        if (pre) {
            store_row<N>(T.prof + ((size_t)T.root * T.P + p) * N, out);
            T.states[(size_t)T.root * T.P + p] = (uint8_t)jm;
            float *rp = slot(S.root_pre);
#pragma unroll
            for (int j = 0; j < N; ++j) rp[j] = out[j];
        }
        s_red[t] = mult_p * lp;
    }
    // the preorder reads the subtree rows back from memory when slots were
    // reused: a full barrier makes the postorder's row stores visible
    if constexpr (SUBG)
        __syncthreads();
    else
        lds_barrier();

    // ---- preorder: rounds over depth-ordered nodes -------------------------------
    if (pre) {
        constexpr int W = pre_rec_words(N);
        const int *rd = s_rounds + 2 * (tips ? S.Rp + S.Ri : S.Rp);
        const int R = tips ? S.Ra : S.Ri;
        // sub: the subtree row of an internal node, read from memory a round
        // ahead (SUBG)
        auto pre_entry = [&](const int4 &a, const int4 &b, auto M, const float (&sub)[N]) {
            const int c = a.x, pslot = a.y, oslot = a.z, tr = a.w, ir = b.x, sslot = b.y;
            PreRaw<N> Rw;
            if (ir >= 0) {
                Rw.code = 0;
                if constexpr (SUBG) {
#pragma unroll
                    for (int i = 0; i < N; ++i) Rw.sub[i] = sub[i];
                } else {
                    const float *src = slot(sslot);
#pragma unroll
                    for (int i = 0; i < N; ++i) Rw.sub[i] = src[i];
                }
            } else {
                Rw.code = s_tip[tr * TP + q];
            }
            const float *pp = slot(pslot);
#pragma unroll
            for (int i = 0; i < N; ++i) Rw.pp[i] = pp[i];
            pre_finish_m<N>(T, c, ir, p, Rw, s_lut, M, oslot >= 0 ? slot(oslot) : nullptr);
        };
        const int32_t *src;  // as in the postorder: records one round ahead
        if constexpr (STAGED)
            src = s_rec + max(S.pre_off, 0);
        else
            src = S.pre_rec;
        PreRec<N> ra, rb;
        auto fetch = [&](int r, bool &valid, PreRec<N> &rec) {
            const int k = r < R ? rd[2 * r] + k0 : 0;
            valid = r < R && k < rd[2 * r + 1] && act;
            if (!STAGED || valid) load_rec<N>(src + (size_t)W * min(k, S.n_pre - 1), rec);
        };
        if constexpr (SUBG) {
            // records two rounds ahead and the subtree row one round ahead
            // (its address is in the record), three buffers in turn; a
            // subtree row is read unconditionally (row 0 for tips and
            // entries a thread does not run, the last pattern for threads
            // past it)
            PreRec<N> rc;
            float sa[N], sb[N], sc[N];
            auto fetch_sub = [&](const PreRec<N> &rec, bool valid, float (&sub)[N]) {
                const int ir = valid ? max(rec.b.x, 0) : 0;
                load_rowf<N>(T.sub + ((size_t)ir * T.P + min(p, T.P - 1)) * N, sub);
            };
            bool va = false, vb = false, vc = false;
            fetch(0, va, ra);
            fetch(1, vb, rb);
            fetch_sub(ra, va, sa);
            for (int r = 0; r < R; r += 3) {
                fetch(r + 2, vc, rc);
                fetch_sub(rb, vb, sb);
                if (va) pre_entry(ra.a, ra.b, ra.M, sa);
                lds_barrier();
                if (r + 1 >= R) break;
                fetch(r + 3, va, ra);
                fetch_sub(rc, vc, sc);
                if (vb) pre_entry(rb.a, rb.b, rb.M, sb);
                lds_barrier();
                if (r + 2 >= R) break;
                fetch(r + 4, vb, rb);
                fetch_sub(ra, va, sa);
                if (vc) pre_entry(rc.a, rc.b, rc.M, sc);
                lds_barrier();
            }
        } else {
            const float none[N] = {};
            bool va, vb;
            if (R > 0) fetch(0, va, ra);
            for (int r = 0; r < R; r += 2) {
                fetch(r + 1, vb, rb);
                if (va) pre_entry(ra.a, ra.b, ra.M, none);
                lds_barrier();
                if (r + 1 >= R) break;
                fetch(r + 2, va, ra);
                if (vb) pre_entry(rb.a, rb.b, rb.M, none);
                lds_barrier();
            }
        }
    }

    // ---- workgroup sum of mult * logL, then the total over workgroups -------
    tile_total(T, S, s_red, TP);
}

// One-launch joint pass (treeanc.py:934-1080) over the same tiles, slots and
// rounds: the postorder rounds form each internal row's max-product message
// (joint_children / joint_store arithmetic: the same products, comparisons
// and first-max rule, so messages, argmax rows and states are the level
// kernels' own), the message rows live in the LDS slots, the argmax rows and
// the decoded states in LDS as well (and all of them go to memory as
// results); the root step picks the root states, the backtrack rounds over
// the depth-ordered nodes give every node its state and add its FP64 score
// term, and the workgroups' sums of mult * logL make the total as in the
// marginal pass.  STAGED: the descriptors of both phases are staged in LDS;
// otherwise each thread loads its next round's descriptor while the current
// round runs (two register sets in turn, as the marginal pass streams its
// records).
template <int N, bool STAGED>
__global__ __launch_bounds__(kTileThreads) void k_tile_joint(TreeDev T, TileDev S, const int32_t *pre_desc,
                                                             int tips) {
    using R = JR<N>;
    constexpr int CB = (int)cx_row_bytes(N);
    extern __shared__ __align__(16) unsigned char s_dyn[];
    const int TP = S.TP;
    // LDS layout (tile_joint_lds_bytes): message rows, reduction, round
    // offsets, postorder descriptors, preorder descriptors, tip codes,
    // argmax rows, states
    float *rows = reinterpret_cast<float *>(s_dyn);                                   // [n_slots][TP][N]
    double *s_red = reinterpret_cast<double *>(rows + align4((size_t)S.n_slots * TP * N));  // [kTileThreads]
    int *s_rounds = reinterpret_cast<int *>(s_red + kTileThreads);
    const int n_rounds = 2 * (S.Rp + S.Ri + S.Ra);
    int4 *s_post = reinterpret_cast<int4 *>(s_rounds + align4(n_rounds));  // [n_post][2] (STAGED)
    int4 *s_pre = s_post + (STAGED ? 2 * (size_t)S.n_post : 0);           // [n_pre] (STAGED)
    uint8_t *s_tip = reinterpret_cast<uint8_t *>(s_pre + (STAGED ? S.n_pre : 0));  // [n_tips][TP]
    uint8_t *s_cx = s_tip + (((size_t)T.n_tips * TP + 15) & ~(size_t)15);             // [n_nodes][TP][CB]
    uint8_t *s_st = s_cx + (((size_t)T.n_nodes * TP * CB + 15) & ~(size_t)15);       // [n_nodes][TP]

    const int t = threadIdx.x;
    const int q = t & (TP - 1);
    const int k0 = t / TP;
    const int p = blockIdx.x * TP + q;
    const bool act = p < T.P;
    const double mult_p = T.mult[min(p, T.P - 1)];  // read now, as in the marginal pass

    constexpr int W = post_rec_words(N);
    {
        const int n_post = STAGED ? 2 * S.n_post : 0, n_pre = STAGED ? S.n_pre : 0, n_tipb = T.n_tips * TP;
        auto f_rd = [&](int i) { return S.rounds[i]; };
        auto f_post = [&](int i) { return reinterpret_cast<const int4 *>(S.post_rec + (size_t)W * (i >> 1))[i & 1]; };
        auto f_pre = [&](int i) { return reinterpret_cast<const int4 *>(pre_desc)[i]; };
        auto f_tip = tip_reader(T, TP, blockIdx.x * TP);
        Stager<int, 1> g_rd;
        Stager<int4, 4> g_post, g_pre;
        Stager<uint8_t, kStageLoads> g_tip;
        g_rd.load(n_rounds, f_rd);
        g_post.load(n_post, f_post);
        g_pre.load(n_pre, f_pre);
        g_tip.load(n_tipb, f_tip);
        g_rd.store(s_rounds, n_rounds, f_rd);
        g_post.store(s_post, n_post, f_post);
        g_pre.store(s_pre, n_pre, f_pre);
        g_tip.store(s_tip, n_tipb, f_tip);
    }
    lds_barrier();

    auto slot = [&](int sl) { return rows + ((size_t)sl * TP + q) * N; };
    const auto PJ = joint_P<N>(T);
    const auto lutj = joint_lut<N>(T);
    auto put_cx = [&](int node, const int (&cx)[N]) {
        uint8_t *d = s_cx + ((size_t)node * TP + q) * CB;
        if constexpr (cx_packed(N)) {
            uint32_t b = 0;
#pragma unroll
            for (int j = 0; j < N; ++j) b |= (uint32_t)cx[j] << (2 * j);
            d[0] = (uint8_t)b;
        } else {
#pragma unroll
            for (int j = 0; j < N; ++j) d[j] = (uint8_t)cx[j];
        }
    };
    auto get_cx = [&](int node, int ps) -> int {
        const uint8_t *d = s_cx + ((size_t)node * TP + q) * CB;
        if constexpr (cx_packed(N))
            return (d[0] >> (2 * ps)) & 3;
        else
            return d[ps];
    };
    // tip child (tip_joint_code): msg_j *= max_i max(TINY, P_ij) max(profile_i, TINY)
    auto tip_child = [&](int cnode, int tr, R (&msg)[N]) {
        const auto lj = lutj + s_tip[tr * TP + q] * N;
        const auto P0 = PJ + (size_t)cnode * (N * N);
        R tl[N];
#pragma unroll
        for (int i = 0; i < N; ++i) tl[i] = lj[i];
        int cx[N];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            R best = P0[j] * tl[0];
            int bi = 0;
#pragma unroll
            for (int i = 1; i < N; ++i) {
                const R val = P0[i * N + j] * tl[i];
                if (val > best) {
                    best = val;
                    bi = i;
                }
            }
            msg[j] *= best;
            cx[j] = bi;
        }
        put_cx(cnode, cx);
        if (tips) cx_store<N>(T, cnode, p, cx);
    };
    // product over the children of the descriptor's row (joint_children)
    auto children = [&](const int4 &a, const int4 &b, R (&msg)[N]) {
#pragma unroll
        for (int i = 0; i < N; ++i) msg[i] = R(1);
        if (a.z == 2) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int cref = c ? b.x : a.w;
                if (cref < 0) {
                    tip_child(c ? b.z : b.y, -cref - 1, msg);
                } else {
                    const float *r = slot(cref);
#pragma unroll
                    for (int i = 0; i < N; ++i) msg[i] *= (R)r[i];
                }
            }
        } else {
            const int e0 = T.child_ptr[a.x], e1 = T.child_ptr[a.x + 1];
            for (int e = e0; e < e1; ++e) {
                const int cref = T.child_ref[e];
                if (cref < 0) {
                    tip_child(T.child_node[e], -cref - 1, msg);
                } else {
                    const float *r = slot(S.post_slot ? S.post_slot[cref] : cref);
#pragma unroll
                    for (int i = 0; i < N; ++i) msg[i] *= (R)r[i];
                }
                if (e - e0 >= 2) {
                    R mx = msg[0];
#pragma unroll
                    for (int i = 1; i < N; ++i) mx = fmax(mx, msg[i]);
                    int ex;
                    (void)frexp(mx, &ex);
#pragma unroll
                    for (int i = 0; i < N; ++i) msg[i] = ldexp(msg[i], -ex);
                }
            }
        }
    };

    // ---- postorder: max-product rounds (the root is the root step's) ---------
    auto post_step = [&](const int4 &a, const int4 &b) {
        if (a.x != T.root_row) {
            R msg[N];
            children(a, b, msg);
            // joint_store: Lx_j = max_i max(TINY, P_ij) msg_i over the row's own branch
            const auto P1 = PJ + (size_t)b.w * (N * N);
            R best[N];
            int cx[N];
            R off = R(0);
            int jm = 0;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                R bb = P1[j] * msg[0];
                int bi = 0;
#pragma unroll
                for (int i = 1; i < N; ++i) {
                    const R val = P1[i * N + j] * msg[i];
                    if (val > bb) {
                        bb = val;
                        bi = i;
                    }
                }
                best[j] = bb;
                if (bb > off) {
                    off = bb;
                    jm = j;
                }
                cx[j] = bi;
            }
            const R inv = R(1) / off;
            float out[N];
#pragma unroll
            for (int j = 0; j < N; ++j) out[j] = (j == jm) ? 1.0f : (float)(best[j] * inv);
            float *dst = slot(a.y);
#pragma unroll
            for (int j = 0; j < N; ++j) dst[j] = out[j];
            store_row<N>(T.Lx + ((size_t)a.x * T.P + p) * N, out);
            cx_store<N>(T, b.w, p, cx);
            put_cx(b.w, cx);
        }
    };
    auto post_desc = [&](int k, int4 &a, int4 &b) {
        if constexpr (STAGED) {
            a = s_post[2 * k];
            b = s_post[2 * k + 1];
        } else {
            const int4 *g = reinterpret_cast<const int4 *>(S.post_rec + (size_t)W * k);
            a = g[0];
            b = g[1];
        }
    };
    {
        const int *rd = s_rounds;
        if constexpr (STAGED) {
            for (int r = 0; r < S.Rp; ++r) {
                const int k = rd[2 * r] + k0;
                if (k < rd[2 * r + 1] && act) {
                    int4 a, b;
                    post_desc(k, a, b);
                    post_step(a, b);
                }
                lds_barrier();
            }
        } else {
            // two descriptor sets in turn, loaded unconditionally (clamped index)
            auto entry = [&](int r, bool &valid) {
                const int k = r < S.Rp ? rd[2 * r] + k0 : 0;
                valid = r < S.Rp && k < rd[2 * r + 1];
                return min(k, S.n_post - 1);
            };
            int4 a0, b0, a1, b1;
            bool va = false, vb = false;
            if (S.Rp > 0) post_desc(entry(0, va), a0, b0);
            for (int r = 0; r < S.Rp; r += 2) {
                post_desc(entry(r + 1, vb), a1, b1);
                if (va && act) post_step(a0, b0);
                lds_barrier();
                if (r + 1 >= S.Rp) break;
                post_desc(entry(r + 2, va), a0, b0);
                if (vb && act) post_step(a1, b1);
                lds_barrier();
            }
        }
    }

    // ---- root (treeanc.py:1015-1037): root_joint_state ----------------------
    if (t < TP && act) {
        int4 a, b;
        post_desc(S.n_post - 1, a, b);  // height order: the root comes last
        R msg[N];
        children(a, b, msg);
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
        s_st[(size_t)T.root * TP + q] = (uint8_t)jm;
    }
    lds_barrier();

    // ---- backtrack over every node (treeanc.py:1048-1063) ---------------------
    // d: node, parent, internal row or -1, tip row or -1
    auto back_step = [&](const int4 &d) {
        const int c = d.x, sp = s_st[(size_t)d.y * TP + q];
        if (d.z >= 0) {
            const int sc = get_cx(c, sp);
            s_st[(size_t)c * TP + q] = (uint8_t)sc;
            T.states[(size_t)c * T.P + p] = (uint8_t)sc;
        } else if (tips) {
            T.states[(size_t)c * T.P + p] = (uint8_t)get_cx(c, sp);
        }
    };
    {
        const int *rd = s_rounds + 2 * (S.Rp + S.Ri);
        if constexpr (STAGED) {
            for (int r = 0; r < S.Ra; ++r) {
                const int k = rd[2 * r] + k0;
                if (k < rd[2 * r + 1] && act) back_step(s_pre[k]);
                lds_barrier();
            }
        } else {
            const int4 *g = reinterpret_cast<const int4 *>(pre_desc);
            auto entry = [&](int r, bool &valid) {
                const int k = r < S.Ra ? rd[2 * r] + k0 : 0;
                valid = r < S.Ra && k < rd[2 * r + 1];
                return min(k, S.n_pre - 1);
            };
            bool va = false, vb = false;
            int4 d0 = S.Ra > 0 ? g[entry(0, va)] : int4{0, 0, 0, 0}, d1;
            for (int r = 0; r < S.Ra; r += 2) {
                d1 = g[entry(r + 1, vb)];
                if (va && act) back_step(d0);
                lds_barrier();
                if (r + 1 >= S.Ra) break;
                d0 = g[entry(r + 2, va)];
                if (vb && act) back_step(d1);
                lds_barrier();
            }
        }
    }
    // ---- the score of the decoded states -----------------------------------------
    // Every non-root node's term from the states in LDS, the thread's share
    // of the preorder entries in increasing order; the entries of a chunk
    // are independent, so their reads of log P are in flight together.
    double term = 0.0;
    if (act) {
        const int4 *g;
        if constexpr (STAGED)
            g = s_pre;
        else
            g = reinterpret_cast<const int4 *>(pre_desc);
        constexpr int U = 8;
        const int step = kTileThreads / TP;
        for (int k = k0; k < S.n_pre; k += U * step) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = k + u * step;
                const int4 d = g[min(kk, S.n_pre - 1)];
                const int sp = s_st[(size_t)d.y * TP + q];
                const double x = d.z >= 0 ? joint_int_term<N>(T, d.x, min((int)s_st[(size_t)d.x * TP + q], N - 1), sp)
                                          : joint_tip_term<N>(T, T.loglut, d.x, s_tip[max(d.w, 0) * TP + q], sp);
                v[u] = kk < S.n_pre ? x : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) term += v[u];
        }
    }
    const double term_sum = pattern_sum(term, s_red, TP);
    if (t < TP && act) {
        const double lp = log(T.pi[s_st[(size_t)T.root * TP + q]]) + term_sum;
        T.jseqlh[p] = lp;
        s_red[t] = mult_p * lp;
    }
    lds_barrier();

    // ---- workgroup sum of mult * logL, then the total over workgroups -------
    tile_total(T, S, s_red, TP);
}

// Records of the tile pass from the level lists and the current P(t): one
// thread per post entry (g < n_post) or pre entry.
// post_slot: slot of every internal row; pre_slot: slot of every node's
// profile (-1 for tips).
template <int N>
__global__ void k_tile_pack(TreeDev T, const int32_t *post_desc, int n_post, const int32_t *pre_desc, int n_pre,
                            const int32_t *post_slot, const int32_t *pre_slot, int subg, int32_t *post_rec,
                            int32_t *pre_rec) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n_post) {
        const int32_t *d = post_desc + kPostDesc * g;  // {row, first edge, children, child 0, child 1, ...}
        int32_t *o = post_rec + (size_t)post_rec_words(N) * g;
        for (int i = 0; i < kPostDesc; ++i) o[i] = d[i];
        o[1] = post_slot[d[0]];
        o[3] = d[3] >= 0 ? post_slot[d[3]] : d[3];
        o[4] = d[4] >= 0 ? post_slot[d[4]] : d[4];
        float *f = reinterpret_cast<float *>(o + 8);
        const bool bin = d[2] == 2;
        for (int i = 0; i < N * N; ++i) {
            f[i] = bin ? T.D32[(size_t)d[5] * (N * N) + i] : 0.0f;
            f[N * N + i] = bin ? T.D32[(size_t)d[6] * (N * N) + i] : 0.0f;
        }
    } else if (g < n_post + n_pre) {
        const int k = g - n_post;
        const int32_t *d = pre_desc + kPreDesc * k;  // {node, parent node, row or -1, tip row or -1}
        int32_t *o = pre_rec + (size_t)pre_rec_words(N) * k;
        o[0] = d[0];
        o[1] = pre_slot[d[1]];
        o[2] = d[2] >= 0 ? pre_slot[d[0]] : -1;
        o[3] = d[3];
        o[4] = d[2];
        o[5] = d[2] >= 0 && !subg ? post_slot[d[2]] : -1;
        o[6] = 0;
        o[7] = 0;
        float *f = reinterpret_cast<float *>(o + 8);
        for (int i = 0; i < N * N; ++i) f[i] = T.P32[(size_t)d[0] * (N * N) + i];
    }
}

}  // namespace

namespace ttk {

int tile_threads() { return kTileThreads; }
int tile_post_rec_words(int n) { return post_rec_words(n); }
int tile_pre_rec_words(int n) { return pre_rec_words(n); }

// Words of the staged-record region: the larger phase's records, or both
// phases' (sep: the preorder's records in their own region).
size_t tile_stage_words(int n, int n_post, int n_pre, bool sep) {
    const size_t a = (size_t)n_post * post_rec_words(n), b = (size_t)n_pre * pre_rec_words(n);
    return sep ? a + b : std::max(a, b);
}

// LDS bytes of k_tile_marginal for tp patterns per workgroup, n_slots rows
// per pattern and n_rounds round bounds (two per round) in all, with
// stage_words of staged records.
size_t tile_lds_bytes(const TreeDev &T, int tp, int n_slots, int n_rounds, size_t stage_words) {
    const size_t rows = align4((size_t)n_slots * tp * T.n) * sizeof(float);
    const size_t lut = align4((size_t)T.n_codes * T.n) * sizeof(float);
    const size_t red = (size_t)kTileThreads * sizeof(double);
    const size_t rounds = align4((size_t)n_rounds) * sizeof(int);
    const size_t tips = ((size_t)T.n_tips * tp + 15) & ~(size_t)15;
    return rows + lut + red + rounds + tips + stage_words * sizeof(int32_t);
}

// LDS bytes of k_tile_joint (layout in the kernel).
size_t tile_joint_lds_bytes(const TreeDev &T, int tp, int n_slots, int n_rounds, int n_post, int n_pre, bool staged) {
    const size_t rows = align4((size_t)n_slots * tp * T.n) * sizeof(float);
    const size_t red = (size_t)kTileThreads * sizeof(double);
    const size_t rounds = align4((size_t)n_rounds) * sizeof(int);
    const size_t desc = staged ? (size_t)n_post * 32 + (size_t)n_pre * 16 : 0;
    const size_t tips = ((size_t)T.n_tips * tp + 15) & ~(size_t)15;
    const size_t cx = ((size_t)T.n_nodes * tp * cx_row_bytes(T.n) + 15) & ~(size_t)15;
    const size_t st = (size_t)T.n_nodes * tp;
    return rows + red + rounds + desc + tips + cx + st;
}

#define TT_TILE_DISPATCH(n, CALL)                        \
    switch (n) {                                         \
        case 2: { constexpr int N_ = 2; CALL; } break;   \
        case 3: { constexpr int N_ = 3; CALL; } break;   \
        case 4: { constexpr int N_ = 4; CALL; } break;   \
        case 5: { constexpr int N_ = 5; CALL; } break;   \
        default: return hipErrorInvalidValue;            \
    }

template <int N>
const void *tile_kernel(bool staged, bool subg) {
    if (subg)
        return staged ? (const void *)k_tile_marginal<N, true, true> : (const void *)k_tile_marginal<N, false, true>;
    return staged ? (const void *)k_tile_marginal<N, true, false> : (const void *)k_tile_marginal<N, false, false>;
}

// Workgroups of k_tile_marginal one CU holds at once (registers and LDS).
int tile_blocks_per_cu(int n, size_t lds, bool staged, bool subg) {
    const void *fn = n == 2 ? tile_kernel<2>(staged, subg) : n == 3 ? tile_kernel<3>(staged, subg)
                   : n == 4 ? tile_kernel<4>(staged, subg) : n == 5 ? tile_kernel<5>(staged, subg) : nullptr;
    if (!fn) return 0;
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, kTileThreads, lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return b;
}

template <int N>
const void *tile_joint_kernel(bool staged) {
    return staged ? (const void *)k_tile_joint<N, true> : (const void *)k_tile_joint<N, false>;
}

// Workgroups of k_tile_joint one CU holds at once.
int tile_joint_blocks_per_cu(int n, size_t lds, bool staged) {
    const void *fn = n == 2 ? tile_joint_kernel<2>(staged) : n == 3 ? tile_joint_kernel<3>(staged)
                   : n == 4 ? tile_joint_kernel<4>(staged) : n == 5 ? tile_joint_kernel<5>(staged) : nullptr;
    if (!fn) return 0;
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, kTileThreads, lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return b;
}

hipError_t tile_joint(const TreeDev &T, const TileArgs &A, const int32_t *pre_desc, int tips, bool staged,
                      hipStream_t s) {
    TileDev S{A.post_rec, A.pre_rec, A.rounds, A.Rp, A.Ri, A.Ra, A.tp, (T.P + A.tp - 1) / A.tp, A.n_post, A.n_pre,
              A.n_slots, A.root_post, A.root_pre, A.subg ? A.post_slot : nullptr, A.partial, A.counter};
    const size_t lds = tile_joint_lds_bytes(T, A.tp, A.n_slots, 2 * (A.Rp + A.Ri + A.Ra), A.n_post, A.n_pre, staged);
    const dim3 grid((unsigned)S.n_blocks), block(kTileThreads);
    if (staged) {
        TT_TILE_DISPATCH(T.n, hipLaunchKernelGGL((k_tile_joint<N_, true>), grid, block, lds, s, T, S, pre_desc, tips));
    } else {
        TT_TILE_DISPATCH(T.n, hipLaunchKernelGGL((k_tile_joint<N_, false>), grid, block, lds, s, T, S, pre_desc, tips));
    }
    return hipGetLastError();
}

hipError_t tile_pack(const TreeDev &T, const int32_t *post_desc, int n_post, const int32_t *pre_desc, int n_pre,
                     const int32_t *post_slot, const int32_t *pre_slot, bool subg, int32_t *post_rec,
                     int32_t *pre_rec, hipStream_t s) {
    const int total = n_post + n_pre;
    if (!total) return hipSuccess;
    TT_TILE_DISPATCH(T.n, hipLaunchKernelGGL(k_tile_pack<N_>, dim3((total + 255) / 256), dim3(256), 0, s, T,
                                             post_desc, n_post, pre_desc, n_pre, post_slot, pre_slot, subg ? 1 : 0,
                                             post_rec, pre_rec));
    return hipGetLastError();
}

hipError_t tile_marginal(const TreeDev &T, const TileArgs &A, int tips, int pre, hipStream_t s) {
    TileDev S{A.post_rec, A.pre_rec, A.rounds, A.Rp, A.Ri, A.Ra, A.tp, (T.P + A.tp - 1) / A.tp, A.n_post, A.n_pre,
              A.n_slots, A.root_post, A.root_pre, A.subg ? A.post_slot : nullptr, A.partial, A.counter};
    const bool sep = A.staged && A.pre_sep;
    S.pre_off = sep ? A.n_post * post_rec_words(T.n) : -1;
    const size_t lds = tile_lds_bytes(T, A.tp, A.n_slots, 2 * (A.Rp + A.Ri + A.Ra),
                                      A.staged ? tile_stage_words(T.n, A.n_post, A.n_pre, sep) : 0);
    const dim3 grid((unsigned)S.n_blocks), block(kTileThreads);
    if (A.staged && A.subg) {
        TT_TILE_DISPATCH(T.n, hipLaunchKernelGGL((k_tile_marginal<N_, true, true>), grid, block, lds, s, T, S, tips, pre));
    } else if (A.staged) {
        TT_TILE_DISPATCH(T.n, hipLaunchKernelGGL((k_tile_marginal<N_, true, false>), grid, block, lds, s, T, S, tips, pre));
    } else if (A.subg) {
        TT_TILE_DISPATCH(T.n, hipLaunchKernelGGL((k_tile_marginal<N_, false, true>), grid, block, lds, s, T, S, tips, pre));
    } else {
        TT_TILE_DISPATCH(T.n, hipLaunchKernelGGL((k_tile_marginal<N_, false, false>), grid, block, lds, s, T, S, tips, pre));
    }
    return hipGetLastError();
}

}  // namespace ttk

// One-launch marginal (or log-likelihood) pass of one tree with its internal
// rows resident in LDS.
//
// The level kernels pay a launch per level, and a pass over a small or
// medium tree is a chain of ~2 x depth dependent levels whose work is a few
// microseconds in all.  Here a workgroup takes TP patterns through the whole
// tree: the internal rows of those patterns live in LDS (n_int x TP x n
// floats), every level is one round of the workgroup's threads over
// (node, pattern) pairs, and levels are separated by a workgroup barrier
// instead of a launch and a round trip through HBM.  Patterns are
// independent, so workgroups never wait for each other; the last one to
// finish sums the per-pattern log-likelihoods (fixed order) into the total.
//
// Rows are written to HBM as well (subtree rows, profiles, states: the
// results), but never read back.  The per-(node, pattern) arithmetic is the
// level kernels' own device code (tt_device.h) on values read from LDS, so
// profiles and states are identical to theirs; the per-pattern log
// normalisers are summed per thread and then across the threads of a
// pattern, so logL agrees to rounding.
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

// level-list descriptors read per lane (a wave covers several nodes)
__device__ __forceinline__ PostDesc post_desc_at(const int32_t *d) {
    const int4 a = *reinterpret_cast<const int4 *>(d), b = *reinterpret_cast<const int4 *>(d + 4);
    return PostDesc{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}
__device__ __forceinline__ PreDesc pre_desc_at(const int32_t *d) {
    const int4 a = *reinterpret_cast<const int4 *>(d);
    return PreDesc{a.x, a.y, a.z, a.w};
}

struct TileDev {
    const int32_t *post_desc;  // kPostDesc ints per entry of the height-ordered internal rows
    const int32_t *post_off;   // [H + 1] level offsets (the last level is the root)
    const int32_t *pre_desc;   // kPreDesc ints per entry of the depth-ordered nodes
    const int32_t *pre_off;    // [D + 1]
    const int32_t *pre_nint;   // [D] internal nodes per depth (listed first)
    int32_t H, D;
    int32_t TP;                     // patterns per workgroup (power of two)
    int32_t n_blocks;
    double *partial;                // [n_blocks] sum over the workgroup's patterns of mult * logL
    unsigned *counter;              // workgroups finished; reset to 0 by the last one
};

template <int N>
__global__ __launch_bounds__(kTileThreads) void k_tile_marginal(TreeDev T, TileDev S, int tips, int pre) {
    extern __shared__ __align__(16) unsigned char s_dyn[];
    const int TP = S.TP;
    // LDS layout (tile_lds_bytes): rows, profile map, reduction, tip codes;
    // each part starts on a 16-byte boundary
    float *rows = reinterpret_cast<float *>(s_dyn);                               // [n_int][TP][N]
    float *s_lut = rows + align4((size_t)T.n_int * TP * N);                       // [n_codes][N]
    double *s_red = reinterpret_cast<double *>(s_lut + align4((size_t)T.n_codes * N));  // [kTileThreads]
    uint8_t *s_tip = reinterpret_cast<uint8_t *>(s_red + kTileThreads);             // [n_tips][TP]
    __shared__ double s_sum[TT_WAVE];
    __shared__ bool s_last;

    const int t = threadIdx.x;
    const int q = t & (TP - 1);            // this thread's pattern within the tile
    const int k0 = t / TP, kstep = kTileThreads / TP;
    const int p = blockIdx.x * TP + q;
    const bool act = p < T.P;

    // stage the profile map and the tile's tip characters
    for (int i = t; i < T.n_codes * N; i += kTileThreads) s_lut[i] = T.lut[i];
    for (int i = t; i < T.n_tips * TP; i += kTileThreads) {
        const int r = i / TP, pp = blockIdx.x * TP + (i & (TP - 1));
        s_tip[i] = pp < T.P ? T.tip_codes[(size_t)r * T.P + pp] : 0;
    }
    __syncthreads();

    auto slot = [&](int row) { return rows + ((size_t)row * TP + q) * N; };
    auto rowf = [&](int cref, float (&r)[N]) {
        if (cref < 0) {
            const int code = s_tip[(-cref - 1) * TP + q];
#pragma unroll
            for (int i = 0; i < N; ++i) r[i] = s_lut[code * N + i];
        } else {
            const float *src = slot(cref);
#pragma unroll
            for (int i = 0; i < N; ++i) r[i] = src[i];
        }
    };

    // ---- postorder: one round per height -----------------------------------
    LogAcc acc;
    for (int h = 0; h < S.H; ++h) {
        const int kb = S.post_off[h], ke = S.post_off[h + 1];
        for (int k = kb + k0; k < ke; k += kstep) {
            const PostDesc d = post_desc_at(S.post_desc + kPostDesc * k);
            if (!act) continue;
            if (d.nchild == 2) {
                BinRaw<N> R;
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const int cref = c ? d.cref1 : d.cref0;
                    if (cref < 0) {
                        R.code[c] = s_tip[(-cref - 1) * TP + q];
                    } else {
                        R.code[c] = 0;
                        const float *src = slot(cref);
#pragma unroll
                        for (int i = 0; i < N; ++i) R.row[c][i] = src[i];
                    }
                }
                bin_finish_refs<N>(T, d.v, d.cref0, d.cref1, d.cnode0, d.cnode1, p, R, s_lut, acc, slot(d.v));
            } else {
                post_node_f<N>(T, d.v, p, s_lut, acc, rowf, [](int e) { return e; }, slot(d.v));
            }
        }
        __syncthreads();
    }

    // ---- root (treeanc.py:814-829) ------------------------------------------
    s_red[t] = act ? acc.log_value() : 0.0;
    __syncthreads();
    if (t < TP && act) {
        double off = 0.0;
        for (int k = q; k < kTileThreads; k += TP) off += s_red[k];
        float *rr = slot(T.root_row);
        double x[N], s = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            x[j] = T.pi[j] * (double)rr[j];
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
#pragma unroll
            for (int j = 0; j < N; ++j) rr[j] = out[j];
        }
        s_red[t] = T.mult[p] * lp;
    }
    __syncthreads();

    // ---- preorder: one round per depth ----------------------------------------
    if (pre) {
        for (int dd = 0; dd < S.D; ++dd) {
            const int kb = S.pre_off[dd];
            const int ke = tips ? S.pre_off[dd + 1] : kb + S.pre_nint[dd];
            for (int k = kb + k0; k < ke; k += kstep) {
                const PreDesc c = pre_desc_at(S.pre_desc + kPreDesc * k);
                if (!act) continue;
                PreRaw<N> R;
                const float *pp = slot(T.int_row[c.par]);
#pragma unroll
                for (int i = 0; i < N; ++i) R.pp[i] = pp[i];
                if (c.ir >= 0) {
                    R.code = 0;
                    const float *src = slot(c.ir);
#pragma unroll
                    for (int i = 0; i < N; ++i) R.sub[i] = src[i];
                } else {
                    R.code = s_tip[c.tr * TP + q];
                }
                pre_finish_ir<N>(T, c.c, c.ir, p, R, s_lut, c.ir >= 0 ? slot(c.ir) : nullptr);
            }
            __syncthreads();
        }
    }

    // ---- workgroup sum of mult * logL, then the total over workgroups -------
    if (t < TT_WAVE) {
        double a = 0.0;
        for (int k = t; k < TP; k += TT_WAVE) a += (k < TP && blockIdx.x * TP + k < T.P) ? s_red[k] : 0.0;
        s_sum[t] = a;
    }
    __syncthreads();
    if (t == 0) {
        double a = 0.0;
        for (int k = 0; k < TT_WAVE; ++k) a += s_sum[k];
        __hip_atomic_store(&S.partial[blockIdx.x], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const unsigned done = atomicAdd(S.counter, 1u);
        s_last = done == (unsigned)S.n_blocks - 1;
    }
    __syncthreads();
    if (s_last && t == 0) {  // the last workgroup: sum the partials in workgroup order
        __threadfence();
        double a = 0.0;
        for (int b = 0; b < S.n_blocks; ++b)
            a += __hip_atomic_load(&S.partial[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        T.total[0] = a;
        __hip_atomic_store(S.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace

namespace ttk {

// LDS bytes of k_tile_marginal for tp patterns per workgroup.
size_t tile_lds_bytes(const TreeDev &T, int tp) {
    const size_t rows = align4((size_t)T.n_int * tp * T.n) * sizeof(float);
    const size_t lut = align4((size_t)T.n_codes * T.n) * sizeof(float);
    const size_t red = (size_t)kTileThreads * sizeof(double);
    const size_t tips = ((size_t)T.n_tips * tp + 15) & ~(size_t)15;
    return rows + lut + red + tips;
}

int tile_threads() { return kTileThreads; }

hipError_t tile_marginal(const TreeDev &T, const int32_t *post_desc, const int32_t *post_off, int H,
                         const int32_t *pre_desc, const int32_t *pre_off, const int32_t *pre_nint, int D, int tp,
                         double *partial, unsigned *counter, int tips, int pre, hipStream_t s) {
    TileDev S{post_desc, post_off, pre_desc, pre_off, pre_nint, H, D, tp, (T.P + tp - 1) / tp, partial, counter};
    const size_t lds = tile_lds_bytes(T, tp);
    switch (T.n) {
#define TT_TILE_CASE(NN)                                                                                          \
    case NN:                                                                                                      \
        hipLaunchKernelGGL(k_tile_marginal<NN>, dim3((unsigned)S.n_blocks), dim3(kTileThreads), lds, s, T, S, tips, \
                           pre);                                                                                  \
        break;
        TT_TILE_CASE(2)
        TT_TILE_CASE(3)
        TT_TILE_CASE(4)
        TT_TILE_CASE(5)
#undef TT_TILE_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ttk

// Persistent batched marginal reconstruction of many trees in one launch.
//
// Work queue: items (tree, 64-pattern tile), sorted on the host longest
// first.  A persistent block of 4 waves pops an item with one atomic, then
// walks that tile through the tree's postorder levels (heights), the root and
// the preorder levels (depths); the 4 waves split the nodes of each level and
// meet at a block barrier between levels.  Patterns are independent, so an
// item never waits for another block: there are no cross-block dependencies
// and no spin-waits, and the critical path of the largest tree is about
// (internal nodes / 4) node steps per pass.
//
// Per (node, pattern) arithmetic is the same device code as the level
// kernels (tt_device.h), so profiles and states equal tt_marginal's; the
// per-pattern log normalisers are summed in registers instead of in partial
// rows.
#include "tt_device.h"

#include <cstdlib>

namespace {
using namespace ttd;

constexpr int kWaves = TT_BLOCK / TT_WAVE;

__global__ void k_reset_queue(int *queue) {
    if (threadIdx.x < kXcds) queue[threadIdx.x * kQueueStride] = 0;
}

// Next item for the calling thread's workgroup (-1: all queues empty): its
// die's queue first, then the others; *start is the first queue (relative
// to the own die) still worth trying.
__device__ __forceinline__ int pop_item(int *queue, const int *xoff, int &start) {
    const int x0 = blockIdx.x % kXcds;
    for (; start < kXcds; ++start) {
        const int x = (x0 + start) % kXcds;
        const int lo = xoff[x], n = xoff[x + 1] - lo;
        if (n > 0) {
            const int i = atomicAdd(queue + x * kQueueStride, 1);
            if (i < n) return lo + i;
        }
    }
    return -1;
}

// TILES 64-pattern tiles per item: every lane carries TILES patterns of the
// same nodes, so each node step keeps 2 * TILES row fetches in flight and the
// wave-uniform work (node ids, child lists, P(t)) is shared by all of them.
// The second tile's state is held in separately named variables (no arrays
// of register structs, which the compiler may place in scratch).
template <int N, int TL, class LP>
__device__ __forceinline__ void post_pair_t(const TreeDev &T, const PostDesc &d0, const PostDesc &d1, bool has1,
                                            int p0, int p1, bool act0, bool act1, LP lut, LogAcc &acc0,
                                            LogAcc &acc1) {
    if (d0.nchild == 2 && d1.nchild == 2) {
        BinRaw<N> RA0, RB0, RA1, RB1;
        if (act0) {
            bin_fetch_d<N>(T, d0, p0, RA0);
            bin_fetch_d<N>(T, d1, p0, RB0);
        }
        if (TL == 2 && act1) {
            bin_fetch_d<N>(T, d0, p1, RA1);
            bin_fetch_d<N>(T, d1, p1, RB1);
        }
        if (act0) {
            bin_finish_d<N>(T, d0, p0, RA0, lut, acc0);
            if (has1) bin_finish_d<N>(T, d1, p0, RB0, lut, acc0);
        }
        if (TL == 2 && act1) {
            bin_finish_d<N>(T, d0, p1, RA1, lut, acc1);
            if (has1) bin_finish_d<N>(T, d1, p1, RB1, lut, acc1);
        }
    } else {
        // each node by its own arity (see post_pair)
        if (act0) {
            post_one_d<N>(T, d0, p0, lut, acc0);
            if (has1) post_one_d<N>(T, d1, p0, lut, acc0);
        }
        if (TL == 2 && act1) {
            post_one_d<N>(T, d0, p1, lut, acc1);
            if (has1) post_one_d<N>(T, d1, p1, lut, acc1);
        }
    }
}

template <int N, int TL, class LP>
__device__ __forceinline__ void pre_pair_t(const TreeDev &T, const PreDesc &q0, const PreDesc &q1, bool has1, int p0,
                                           int p1, bool act0, bool act1, LP lut) {
    PreRaw<N> RA0, RB0, RA1, RB1;
    if (act0) {
        pre_fetch_d<N>(T, q0, p0, RA0);
        pre_fetch_d<N>(T, q1, p0, RB0);
    }
    if (TL == 2 && act1) {
        pre_fetch_d<N>(T, q0, p1, RA1);
        pre_fetch_d<N>(T, q1, p1, RB1);
    }
    if (act0) {
        pre_finish_ir<N>(T, q0.c, q0.ir, p0, RA0, lut);
        if (has1) pre_finish_ir<N>(T, q1.c, q1.ir, p0, RB0, lut);
    }
    if (TL == 2 && act1) {
        pre_finish_ir<N>(T, q0.c, q0.ir, p1, RA1, lut);
        if (has1) pre_finish_ir<N>(T, q1.c, q1.ir, p1, RB1, lut);
    }
}

// WPE: waves per SIMD the register allocator must allow (1 = no constraint);
// the two-tile kernel needs 100 VGPRs (4 waves), at 5 it spills a few.
template <int N, int TL, int WPE>
__global__ __launch_bounds__(TT_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE))) void k_batch_marginal(const BatchTreeDev *__restrict__ trees_arg,
                                                             const BatchItem *__restrict__ items_arg, int n_items,
                                                             int *queue, const int *xoff, double *item_out,
                                                             int mode) {
    // mode bit 0: reconstruct tip states; bit 1: postorder and root only (loglik)
    const bool tips = mode & 1, pre = !(mode & 2);
    // tree records and items are read-only: scalar loads through the constant address space
    const BatchTreeDev TT_C *trees = (const BatchTreeDev TT_C *)trees_arg;
    const BatchItem TT_C *items = (const BatchItem TT_C *)items_arg;
    __shared__ int s_item;
    __shared__ double s_acc[kWaves][TL][TT_WAVE];
    __shared__ float s_lut[256 * N];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    int cur_tree = -1, qstart = 0;
    for (;;) {
        if (threadIdx.x == 0) s_item = pop_item(queue, xoff, qstart);
        __syncthreads();
        const int it = s_item;
        if (it < 0) break;
        const BatchItem item = items[it];
        const BatchTreeDev TT_C &B = trees[item.tree];
        const TreeDev TT_C &T = B.T;
        if (item.tree != cur_tree) {  // stage the tree's profile map in LDS
            for (int i = threadIdx.x; i < T.n_codes * N; i += TT_BLOCK) s_lut[i] = T.lut[i];
            cur_tree = item.tree;
        }
        __syncthreads();
        const int p0 = item.tile * TT_WAVE + lane;  // item.tile counts 64-pattern tiles
        const int p1 = p0 + TT_WAVE;
        const bool act0 = p0 < T.P, act1 = TL == 2 && p1 < T.P;
        // postorder: heights 1..H, the last level is the root
        LogAcc acc0, acc1;
        for (int h = 0; h < B.H; ++h) {
            const int kb = B.post_off[h], ke = B.post_off[h + 1];
            for (int k = wave_uniform(kb + w); k < ke; k += 2 * kWaves) {
                const bool has1 = k + kWaves < ke;
                const PostDesc d0 = load_post_desc(B.post_desc + kPostDesc * k);
                const PostDesc d1 = has1 ? load_post_desc(B.post_desc + kPostDesc * (k + kWaves)) : d0;
                post_pair_t<N, TL>(T, d0, d1, has1, p0, p1, act0, act1, s_lut, acc0, acc1);
            }
            __syncthreads();
        }
        // root: per-pattern logL from the 4 waves' normaliser products
        s_acc[w][0][lane] = act0 ? acc0.log_value() : 0.0;
        if (TL == 2) s_acc[w][TL - 1][lane] = act1 ? acc1.log_value() : 0.0;
        __syncthreads();
        double contrib = 0.0;
        if (w == 0) {
            double off0 = 0.0, off1 = 0.0;
#pragma unroll
            for (int q = 0; q < kWaves; ++q) {
                off0 += s_acc[q][0][lane];
                off1 += s_acc[q][TL - 1][lane];
            }
            if (act0) contrib += root_step<N>(T, p0, off0);
            if (act1) contrib += root_step<N>(T, p1, off1);
        }
        const double tot = block_sum(contrib);
        if (threadIdx.x == 0) item_out[it] = tot;
        if (!pre) continue;  // loglik: the next item (its first barrier orders the LDS reuse)
        if (!tips && B.fill_tips)
            for (int k = w; k < T.n_tips; k += kWaves) {
                const size_t row = (size_t)B.tip_nodes[k] * T.P;
                if (act0) T.states[row + p0] = 255;
                if (act1) T.states[row + p1] = 255;
            }
        __syncthreads();
        // preorder: depths 1..D (tips only when reconstructing them)
        for (int d = 0; d < B.D; ++d) {
            const int kb = B.pre_off[d];
            const int ke = tips ? B.pre_off[d + 1] : kb + B.pre_nint[d];
            for (int k = wave_uniform(kb + w); k < ke; k += 2 * kWaves) {
                const bool has1 = k + kWaves < ke;
                const PreDesc q0 = load_pre_desc(B.pre_desc + kPreDesc * k);
                const PreDesc q1 = has1 ? load_pre_desc(B.pre_desc + kPreDesc * (k + kWaves)) : q0;
                pre_pair_t<N, TL>(T, q0, q1, has1, p0, p1, act0, act1, s_lut);
            }
            __syncthreads();
        }
    }
}

// Wave-granular variant: every wave pops its own item and walks all of the
// tile's levels alone (program order replaces the block barriers, so sparse
// deep levels cost one step instead of idling three waves at a barrier).
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = TT_WAVE / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, TT_WAVE);
    return x;
}

template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_batch_marginal_w(const BatchTreeDev *__restrict__ trees_arg,
                                                               const BatchItem *__restrict__ items_arg,
                                                               int n_items, int *queue, const int *xoff,
                                                               double *item_out, int mode) {
    const bool tips = mode & 1, pre = !(mode & 2);
    const BatchTreeDev TT_C *trees = (const BatchTreeDev TT_C *)trees_arg;
    const BatchItem TT_C *items = (const BatchItem TT_C *)items_arg;
    __shared__ float s_lut[kWaves][256 * N];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    float *lut = s_lut[w];
    int cur_tree = -1, qstart = 0;
    for (;;) {
        int it = 0;
        if (lane == 0) it = pop_item(queue, xoff, qstart);
        it = __shfl(it, 0, TT_WAVE);
        if (it < 0) break;
        const BatchItem item = items[it];
        const BatchTreeDev TT_C &B = trees[item.tree];
        const TreeDev TT_C &T = B.T;
        if (item.tree != cur_tree) {  // stage the tree's profile map in this wave's LDS
            for (int i = lane; i < T.n_codes * N; i += TT_WAVE) lut[i] = T.lut[i];
            cur_tree = item.tree;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int p = item.tile * TT_WAVE + lane;
        const bool act = p < T.P;
        LogAcc acc;
        for (int h = 0; h < B.H; ++h) {  // a pair never spans two heights: a parent may follow its child
            const int kb = B.post_off[h], ke = B.post_off[h + 1];
            for (int k = kb; k < ke; k += 2) {
                const bool has1 = k + 1 < ke;
                const int v0 = wave_uniform(B.post_nodes[k]);
                const int v1 = has1 ? wave_uniform(B.post_nodes[k + 1]) : v0;
                if (act) post_pair<N>(T, v0, v1, has1, p, lut, acc);
            }
        }
        double contrib = 0.0;
        if (act) contrib = root_step<N>(T, p, acc.log_value());
        const double tot = wave_sum(contrib);
        if (lane == 0) item_out[it] = tot;
        if (!pre) continue;
        if (!tips && B.fill_tips && act)
            for (int k = 0; k < T.n_tips; ++k) T.states[(size_t)B.tip_nodes[k] * T.P + p] = 255;
        for (int d = 0; d < B.D; ++d) {  // depth order: parents before children
            const int kb = B.pre_off[d];
            const int ke = tips ? B.pre_off[d + 1] : kb + B.pre_nint[d];
            for (int k = kb; k < ke; k += 2) {
                const bool has1 = k + 1 < ke;
                const int c0 = wave_uniform(B.pre_nodes[k]);
                const int c1 = has1 ? wave_uniform(B.pre_nodes[k + 1]) : c0;
                if (act) pre_pair<N>(T, c0, c1, has1, p, lut);
            }
        }
    }
}

// Joint reconstruction of many trees in one persistent launch
// (treeanc.py:934-1080 per tree): a block of 4 waves pops a (tree, 64-pattern
// tile) item and walks its max-product postorder levels (FP32 messages, FP64
// arithmetic, argmax table), the root and the integer backtrack down the
// depths, the level's nodes split over the waves with a block barrier between
// levels.  The backtrack also sums the FP64 log-space score of the decoded
// states (every branch once, tips included); the per-(node, pattern)
// arithmetic is the level kernels' (post_joint_node, root_joint_state,
// joint_int_term, joint_tip_term), so states are identical to tt_joint and
// logL agrees to the order of the per-pattern sum.  mode bit 0: reconstruct
// tips.
// WPE: waves per SIMD the register allocator must allow (1 = no constraint).
// TL: 64-pattern tiles per item; with two, each node step is taken by two
// waves, one per tile, so the node's tables and the byte rows (states,
// argmax rows, tip codes: 64 bytes per tile) are fetched for both tiles at
// once, and 4 / TL nodes are in flight per block step.
template <int N, int WPE, int TL>
__global__ __launch_bounds__(TT_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE))) void k_batch_joint(const BatchTreeDev *__restrict__ trees_arg,
                                                         const BatchItem *__restrict__ items_arg, int n_items,
                                                         int *queue, const int *xoff, double *item_out,
                                                         int mode) {
    const BatchTreeDev TT_C *trees = (const BatchTreeDev TT_C *)trees_arg;
    const BatchItem TT_C *items = (const BatchItem TT_C *)items_arg;
    const int tips = mode & 1;
    constexpr bool kStage = N <= 8;  // profile maps staged in LDS (n_codes <= 256 rows of N doubles each)
    __shared__ int s_item;
    __shared__ double s_sc[kWaves][TT_WAVE];
    __shared__ double s_lj[kStage ? 256 * N : 1];
    __shared__ double s_ll[kStage ? 256 * N : 1];
    constexpr int NW = kWaves / TL;  // nodes per block step
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int wn = w % NW, wt = w / NW;  // this wave's node slot and tile
    int cur_tree = -1, qstart = 0;
    for (;;) {
        if (threadIdx.x == 0) s_item = pop_item(queue, xoff, qstart);
        __syncthreads();
        const int it = s_item;
        if (it < 0) break;
        const BatchItem item = items[it];
        const BatchTreeDev TT_C &B = trees[item.tree];
        const TreeDev TT_C &T = B.T;
        if constexpr (kStage) {
            if (item.tree != cur_tree) {
                for (int i = threadIdx.x; i < T.n_codes * N; i += TT_BLOCK) {
                    s_lj[i] = T.lutj[i];
                    s_ll[i] = T.loglut[i];
                }
                cur_tree = item.tree;
            }
            __syncthreads();
        }
        auto tip_term = [&](int c, int code, int sp) {
            if constexpr (kStage)
                return joint_tip_term_sel<N>(T, s_ll, c, code, sp);
            else
                return joint_tip_term<N>(T, T.loglut, c, code, sp);
        };
        const int p = (item.tile + wt) * TT_WAVE + lane;
        const bool act = p < T.P;
        // postorder: heights 1..H-1 (the last level is the root), two nodes per step
        for (int h = 0; h + 1 < B.H; ++h) {
            const int kb = B.post_off[h], ke = B.post_off[h + 1];
            for (int k = wave_uniform(kb + wn); k < ke; k += 2 * NW) {
                const bool has1 = k + NW < ke;
                const PostDesc d0 = load_post_desc(B.post_desc + kPostDesc * k);
                const PostDesc d1 = has1 ? load_post_desc(B.post_desc + kPostDesc * (k + NW)) : d0;
                if (!act) continue;
                bool pair = false;
                if constexpr (kStage) pair = d0.nchild == 2 && d1.nchild == 2;
                if (pair) {
                    if constexpr (kStage) {
                        JointRaw<N> RA, RB;  // both fetched unconditionally so they stay in registers
                        joint_fetch_d<N>(T, d0, p, RA);
                        joint_fetch_d<N>(T, d1, p, RB);
                        joint_finish_d<N>(T, d0, p, RA, s_lj, tips);
                        if (has1) joint_finish_d<N>(T, d1, p, RB, s_lj, tips);
                    }
                } else {
                    post_joint_node<N>(T, d0.v, d0.vn, p, tips);
                    if (has1) post_joint_node<N>(T, d1.v, d1.vn, p, tips);
                }
            }
            __syncthreads();
        }
        if (wn == 0 && act) (void)root_joint_state<N>(T, p, tips);
        if (!tips && B.fill_tips && act)
            for (int k = wn; k < T.n_tips; k += NW) T.states[(size_t)B.tip_nodes[k] * T.P + p] = 255;
        __syncthreads();
        // backtrack (treeanc.py:1048-1063): depths 1..D, parents before
        // children, kStep nodes per wave step; the score of the decoded states
        // (every non-root branch once) rides along for internal children
        constexpr int kStep = 4;
        double sc = 0.0;
        for (int d = 0; d < B.D; ++d) {
            const int kb = B.pre_off[d], kint = kb + B.pre_nint[d];
            const int ke = tips ? B.pre_off[d + 1] : kint;
            for (int k = wave_uniform(kb + wn); k < ke; k += kStep * NW) {
                PreDesc q[kStep];
#pragma unroll
                for (int m = 0; m < kStep; ++m)
                    q[m] = load_pre_desc(B.pre_desc + kPreDesc * (k + m * NW < ke ? k + m * NW : k));
                if (!act) continue;
                int ps[kStep], cs[kStep];
#pragma unroll
                for (int m = 0; m < kStep; ++m) ps[m] = T.states[(size_t)q[m].par * T.P + p];
#pragma unroll
                for (int m = 0; m < kStep; ++m) {
                    if constexpr (cx_packed(N))
                        cs[m] = (T.Cx[(size_t)q[m].c * T.P + p] >> (2 * ps[m])) & 3;
                    else
                        cs[m] = T.Cx[((size_t)q[m].c * T.P + p) * N + ps[m]];
                }
#pragma unroll
                for (int m = 0; m < kStep; ++m)
                    if (k + m * NW < ke) T.states[(size_t)q[m].c * T.P + p] = cs[m];
#pragma unroll
                for (int m = 0; m < kStep; ++m)
                    if (k + m * NW < kint) sc += joint_int_term<N>(T, q[m].c, cs[m], ps[m]);
            }
            __syncthreads();
        }
        // tip branches (states final: no barriers)
        if (act) {
            for (int d = 0; d < B.D; ++d) {
                const int kt = B.pre_off[d] + B.pre_nint[d], ke = B.pre_off[d + 1];
                for (int k = wave_uniform(kt + wn); k < ke; k += kStep * NW) {
                    PreDesc q[kStep];
#pragma unroll
                    for (int m = 0; m < kStep; ++m)
                        q[m] = load_pre_desc(B.pre_desc + kPreDesc * (k + m * NW < ke ? k + m * NW : k));
                    int ps[kStep], code[kStep];
#pragma unroll
                    for (int m = 0; m < kStep; ++m) {
                        ps[m] = T.states[(size_t)q[m].par * T.P + p];
                        code[m] = T.tip_codes[(size_t)q[m].tr * T.P + p];
                    }
#pragma unroll
                    for (int m = 0; m < kStep; ++m)
                        if (k + m * NW < ke) sc += tip_term(q[m].c, code[m], ps[m]);
                }
            }
        }
        s_sc[w][lane] = sc;
        __syncthreads();
        double contrib = 0.0;
        if (wn == 0 && act) {
            double lp = log(T.pi[T.states[(size_t)T.root * T.P + p]]);
#pragma unroll
            for (int q = 0; q < NW; ++q) lp += s_sc[wt * NW + q][lane];
            T.jseqlh[p] = lp;
            contrib = T.mult[p] * lp;
        }
        const double tot = block_sum(contrib);
        if (threadIdx.x == 0) item_out[it] = tot;
        __syncthreads();
    }
}

// Wide alphabets (n = 20: 256 VGPRs with spills already at one tile) get only
// the one-tile block kernel; the other shapes are instantiated for n <= 8.
constexpr bool all_shapes(int n) { return n <= 8; }
// the joint kernel fits 8 waves per SIMD only up to this alphabet size
constexpr int kJoint8MaxN = 4;

template <int N>
int max_blocks_per_cu(const BatchShape &c) {
    int per_cu = 0;
    if constexpr (!all_shapes(N)) {
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_marginal<N, 1, 1>, TT_BLOCK, 0);
    } else {
        if (c.wave_items)
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_marginal_w<N>, TT_BLOCK, 0);
        else if (c.tiles == 2 && c.waves_per_simd == 5)
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_marginal<N, 2, 5>, TT_BLOCK, 0);
        else if (c.tiles == 2)
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_marginal<N, 2, 1>, TT_BLOCK, 0);
        else
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_marginal<N, 1, 1>, TT_BLOCK, 0);
    }
    return std::max(1, per_cu);
}

}  // namespace

#define TT_BATCH_DISPATCH(n, CALL)                        \
    switch (n) {                                          \
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
__global__ void k_batch_totals(const BatchTreeDev *trees, const int32_t *tot, int m, const double *item_out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int t = tot[3 * e], first = tot[3 * e + 1], cnt = tot[3 * e + 2];
    double sum = 0.0;
    for (int j = first; j < first + cnt; ++j) sum += item_out[j];
    trees[t].T.total[0] = sum;
}
}  // namespace

hipError_t ttk::batch_totals(const BatchTreeDev *trees, const int32_t *tot, int m, const double *item_out,
                             hipStream_t s) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_totals, dim3((m + 255) / 256), dim3(256), 0, s, trees, tot, m, item_out);
    return hipGetLastError();
}

int ttk::device_cus() {
    static int cus = [] {
        int dev = 0, c = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
        return std::max(1, c);
    }();
    return cus;
}

bool ttk::batch_all_shapes(int n) { return all_shapes(n); }

bool ttk::batch_joint_budget(int n, int waves_per_simd) {
    return waves_per_simd == 0 || (waves_per_simd == 6 && all_shapes(n)) || (waves_per_simd == 8 && n <= kJoint8MaxN);
}

hipError_t ttk::batch_joint(int n, const BatchTreeDev *trees, const BatchItem *items, int n_items, int *queue,
                            const int *xoff, double *item_out, int mode, const BatchShape &shape, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    const int resident = device_cus() * std::max(1, shape.blocks_per_cu);
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(kXcds), 0, s, queue);
    const int grid = std::min(resident, n_items);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(TT_BLOCK), 0, s, trees, items, n_items, queue, xoff, item_out, mode);
    };
    TT_BATCH_DISPATCH(n, {
        const bool two = shape.tiles == 2;
        if constexpr (all_shapes(N_)) {
            if (shape.waves_per_simd == 6) {
                two ? go(k_batch_joint<N_, 6, 2>) : go(k_batch_joint<N_, 6, 1>);
                return hipGetLastError();
            }
            if (N_ <= kJoint8MaxN && shape.waves_per_simd == 8) {
                two ? go(k_batch_joint<N_, N_ <= kJoint8MaxN ? 8 : 1, 2>)
                    : go(k_batch_joint<N_, N_ <= kJoint8MaxN ? 8 : 1, 1>);
                return hipGetLastError();
            }
        }
        two ? go(k_batch_joint<N_, 1, 2>) : go(k_batch_joint<N_, 1, 1>);
    });
    return hipGetLastError();
}

int ttk::batch_joint_max_blocks_per_cu(int n, int waves_per_simd, int tiles) {
    auto occ = [](const void *fn) {
        int per_cu = 0;
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, TT_BLOCK, 0);
        return std::max(1, per_cu);
    };
    auto f = [&]() -> int {
        TT_BATCH_DISPATCH(n, {
            const bool two = tiles == 2;
            if constexpr (all_shapes(N_)) {
                if (waves_per_simd == 6)
                    return occ(two ? (const void *)k_batch_joint<N_, 6, 2> : (const void *)k_batch_joint<N_, 6, 1>);
                if (N_ <= kJoint8MaxN && waves_per_simd == 8)
                    return occ(two ? (const void *)k_batch_joint<N_, N_ <= kJoint8MaxN ? 8 : 1, 2>
                                   : (const void *)k_batch_joint<N_, N_ <= kJoint8MaxN ? 8 : 1, 1>);
            }
            return occ(two ? (const void *)k_batch_joint<N_, 1, 2> : (const void *)k_batch_joint<N_, 1, 1>);
        });
        return 1;
    };
    return f();
}

int ttk::batch_max_blocks_per_cu(int n, const BatchShape &c) {
    auto f = [&]() -> int {
        TT_BATCH_DISPATCH(n, return max_blocks_per_cu<N_>(c));
        return 0;
    };
    return f();
}

hipError_t ttk::batch_marginal(int n, const BatchTreeDev *trees, const BatchItem *items, int n_items, int *queue,
                               const int *xoff, double *item_out, int tips, const BatchShape &shape, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    const int resident = device_cus() * std::max(1, shape.blocks_per_cu);
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(kXcds), 0, s, queue);
    TT_BATCH_DISPATCH(n, {
        if constexpr (!all_shapes(N_)) {
            const int grid = std::min(resident, n_items);
            hipLaunchKernelGGL((k_batch_marginal<N_, 1, 1>), dim3(grid), dim3(TT_BLOCK), 0, s, trees, items,
                               n_items, queue, xoff, item_out, tips);
        } else if (shape.wave_items) {
            const int grid = std::min(resident, (n_items + kWaves - 1) / kWaves);
            hipLaunchKernelGGL(k_batch_marginal_w<N_>, dim3(grid), dim3(TT_BLOCK), 0, s, trees, items, n_items,
                               queue, xoff, item_out, tips);
        } else if (shape.tiles == 2 && shape.waves_per_simd == 5) {
            const int grid = std::min(resident, n_items);
            hipLaunchKernelGGL((k_batch_marginal<N_, 2, 5>), dim3(grid), dim3(TT_BLOCK), 0, s, trees, items,
                               n_items, queue, xoff, item_out, tips);
        } else if (shape.tiles == 2) {
            const int grid = std::min(resident, n_items);
            hipLaunchKernelGGL((k_batch_marginal<N_, 2, 1>), dim3(grid), dim3(TT_BLOCK), 0, s, trees, items,
                               n_items, queue, xoff, item_out, tips);
        } else {
            const int grid = std::min(resident, n_items);
            hipLaunchKernelGGL((k_batch_marginal<N_, 1, 1>), dim3(grid), dim3(TT_BLOCK), 0, s, trees, items,
                               n_items, queue, xoff, item_out, tips);
        }
    });
    return hipGetLastError();
}

// 20-state (amino-acid) level kernels on FP32 MFMA (v_mfma_f32_32x32x2f32).
//
// A wave handles one node for 32 patterns.  The child message is computed
// transposed, states x patterns = D^T (32 x 20, rows 20..31 zero) times
// R^T (20 x 32 patterns).  In the 32x32 accumulator lane l holds column
// (pattern) l % 32 and the 16 rows
//   state(r, h) = 8 * (r / 4) + 4 * h + r % 4,  h = l / 32,
// so the two lanes l and l ^ 32 together own one pattern's 20 states, and
// per-pattern reductions (argmax, sums) are a within-lane pass plus one
// exchange with the partner lane.  Operand layouts: lane l supplies
// A[row l % 32][k = l / 32] and B[k = l / 32][col l % 32].
//
// Every vector lives in that accumulator layout, including the inputs: step
// s of the contraction takes input state state(s, h) from lane half h, i.e.
// the value the lane already holds in slot s.  Half 0 holds 12 real states
// and half 1 holds 8 (its slots 8..11 are padding, zero), so a contraction
// is 12 MFMA steps with no data exchanged between lanes, and a lane loads
// only its own states of a row (3 or 2 float4).
//
// Arithmetic follows tt_device.h: deficit messages m = r + d with
// d = r (P - I), normaliser factor formed in FP64 and multiplied into a
// LogAcc; preorder by the division trick.  Polytomies sum per-state log
// messages over their children (the reference's own log-space form).
#include "tt_device.h"

namespace {
using namespace ttd;

typedef float v16f __attribute__((ext_vector_type(16)));
constexpr int NS = 20;

__device__ __forceinline__ int acc_state(int r, int h) { return (r >> 2) * 8 + h * 4 + (r & 3); }

constexpr int kSteps = 12;  // accumulator slots that hold real states in half 0

template <class Ptr>
__device__ __forceinline__ void load_groups(Ptr row, int h, float (&o)[16]);

// This lane's states of the subtree row of child edge cref for pattern p,
// in accumulator layout (zeros for padding and when !act).
template <class LP>
__device__ __forceinline__ void child_slots(const TreeDev &T, int cref, int p, bool act, int h, LP lut,
                                            float (&xo)[16]) {
    if (!act) {
#pragma unroll
        for (int r = 0; r < 16; ++r) xo[r] = 0.0f;
        return;
    }
    if (cref < 0) {
        const int code = T.tip_codes[(size_t)(-cref - 1) * T.P + p];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int st = acc_state(r, h);
            xo[r] = st < NS ? lut[code * NS + (st < NS ? st : 0)] : 0.0f;
        }
    } else {
        load_groups(T.sub + ((size_t)cref * T.P + p) * NS, h, xo);
    }
}

// out_i = sum_k x_k M[k][i] (A[i][k] = M[k * NS + i]).
template <class MP>
__device__ __forceinline__ v16f contract_in(const float (&xo)[16], int h, int i, MP M) {
    v16f acc = {0.0f};
#pragma unroll
    for (int s = 0; s < kSteps; ++s) {
        const int k = acc_state(s, h);
        const float a = (i < NS && k < NS) ? M[(k < NS ? k : 0) * NS + (i < NS ? i : 0)] : 0.0f;
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, xo[s], acc, 0, 0, 0);
    }
    return acc;
}

// out_i = sum_k M[i][k] x_k (A[i][k] = M[i * NS + k]).
template <class MP>
__device__ __forceinline__ v16f contract_out(const float (&xo)[16], int h, int i, MP M) {
    v16f acc = {0.0f};
#pragma unroll
    for (int s = 0; s < kSteps; ++s) {
        const int k = acc_state(s, h);
        const float a = (i < NS && k < NS) ? M[(i < NS ? i : 0) * NS + (k < NS ? k : 0)] : 0.0f;
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, xo[s], acc, 0, 0, 0);
    }
    return acc;
}

// First maximum over the pattern's 20 states held by lanes l and l^32:
// returns the winning state and whether this lane owns it.
__device__ __forceinline__ int pair_argmax(const float (&x)[16], int h, float &best, bool &mine) {
    best = -INFINITY;
    int bi = 1 << 20;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int st = acc_state(r, h);
        if (st < NS && x[r] > best) {
            best = x[r];
            bi = st;
        }
    }
    const float ob = __shfl_xor(best, 32, TT_WAVE);
    const int obi = __shfl_xor(bi, 32, TT_WAVE);
    mine = !(ob > best || (ob == best && obi < bi));
    if (!mine) best = ob;
    return mine ? bi : obi;
}

__device__ __forceinline__ double pair_pick(double mine_value, bool mine) {
    const double other = __shfl_xor(mine_value, 32, TT_WAVE);
    return mine ? mine_value : other;
}

// Store this lane's part (float4 groups) of a 20-state row.
template <class Ptr>
__device__ __forceinline__ void store_groups(Ptr row, int h, const float (&o)[16]) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int base = 8 * g + 4 * h;
        if (base < NS)
            *reinterpret_cast<float4 TT_G *>(row + base) =
                make_float4(o[4 * g], o[4 * g + 1], o[4 * g + 2], o[4 * g + 3]);
    }
}

template <class Ptr>
__device__ __forceinline__ void load_groups(Ptr row, int h, float (&o)[16]) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int base = 8 * g + 4 * h;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (base < NS) v = *reinterpret_cast<const float4 TT_G *>(row + base);
        o[4 * g] = v.x;
        o[4 * g + 1] = v.y;
        o[4 * g + 2] = v.z;
        o[4 * g + 3] = v.w;
    }
}

// Store the node's row normalised by its entry at state jm (value xm).
__device__ __forceinline__ void store_post20(const TreeDev &T, int v, int p, bool act, int h, const float (&x)[16],
                                             int jm, float xm) {
    float o[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) o[r] = (acc_state(r, h) == jm) ? 1.0f : x[r] / xm;
    if (act) store_groups(T.sub + ((size_t)v * T.P + p) * NS, h, o);
}

// Postorder step of internal row v for the 32 patterns of this wave.
template <class LP>
__device__ __forceinline__ void post_node20(const TreeDev &T, const PostDesc &dd, int p, bool act, int h, int i,
                                            LP lut, LogAcc &acc) {
    const int v = dd.v, e0 = dd.e0, e1 = dd.e0 + dd.nchild;
    if (dd.nchild == 2) {
        float ro0[16], ro1[16];  // both children's rows fetched before any arithmetic
        child_slots(T, dd.cref0, p, act, h, lut, ro0);
        child_slots(T, dd.cref1, p, act, h, lut, ro1);
        const v16f d0 = contract_in(ro0, h, i, T.D32 + (size_t)dd.cnode0 * (NS * NS));
        const v16f d1 = contract_in(ro1, h, i, T.D32 + (size_t)dd.cnode1 * (NS * NS));
        float x[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = (ro0[r] + d0[r]) * (ro1[r] + d1[r]);
        // first maximum of this lane's states; the children's (r, d) ride along
        float best = -INFINITY, a0 = 0.f, b0 = 0.f, a1 = 0.f, b1 = 0.f;
        int bi = 1 << 20;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int st = acc_state(r, h);
            if (st < NS && x[r] > best) {
                best = x[r];
                bi = st;
                a0 = ro0[r];
                b0 = d0[r];
                a1 = ro1[r];
                b1 = d1[r];
            }
        }
        const float ob = __shfl_xor(best, 32, TT_WAVE);
        const int obi = __shfl_xor(bi, 32, TT_WAVE);
        const bool mine = !(ob > best || (ob == best && obi < bi));
        const int jm = mine ? bi : obi;
        const float xm = mine ? best : ob;
        const double f = pair_pick(msg64(a0, b0) * msg64(a1, b1), mine);
        acc.mul(f);
        if (T.offs && act && h == 0) T.offs[(size_t)v * T.P + p] = log(f);
        store_post20(T, v, p, act, h, x, jm, xm);
        return;
    }
    // polytomy: per-state sums of log messages, as the reference does
    float L[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) L[r] = 0.0f;
    for (int e = e0; e < e1; ++e) {
        float ro[16];
        child_slots(T, wave_uniform(T.child_ref[e]), p, act, h, lut, ro);
        const v16f d = contract_in(ro, h, i, T.D32 + (size_t)wave_uniform(T.child_node[e]) * (NS * NS));
#pragma unroll
        for (int r = 0; r < 16; ++r) L[r] += (acc_state(r, h) < NS) ? log_msg(ro[r], d[r]) : 0.0f;
    }
    float best;
    bool mine;
    const int jm = pair_argmax(L, h, best, mine);
    acc.mul(exp((double)best));
    if (T.offs && act && h == 0) T.offs[(size_t)v * T.P + p] = (double)best;
    float o[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) o[r] = (acc_state(r, h) == jm) ? 1.0f : expf(L[r] - best);
    if (act) store_groups(T.sub + ((size_t)v * T.P + p) * NS, h, o);
}

__global__ __launch_bounds__(TT_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_post_marginal20(TreeDev T, LevelDev L, int n_ptiles32) {
    __shared__ double s_acc[TT_BLOCK / TT_WAVE][TT_WAVE];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int h = lane >> 5, i = lane & 31;
    const int pt = blockIdx.x % n_ptiles32, cg = blockIdx.x / n_ptiles32;
    const int p = pt * 32 + i;
    const bool act = p < T.P;
    const int chunk = cg * (TT_BLOCK / TT_WAVE) + w;
    LogAcc acc;
    if (chunk < L.n_chunks) {
        const int kb = chunk * L.chunk;
        const int ke = min(kb + L.chunk, L.n_nodes);
        for (int k = wave_uniform(kb); k < ke; ++k)
            post_node20(T, load_post_desc(L.desc + kPostDesc * k), p, act, h, i, T.lut, acc);
    }
    s_acc[w][lane] = (act && h == 0) ? acc.log_value() : 0.0;
    __syncthreads();
    if (w == 0 && act && h == 0) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < TT_BLOCK / TT_WAVE; ++q) s += s_acc[q][lane];
        T.part[(size_t)(L.row_off + cg) * T.P + p] = s;
    }
}

// Preorder step (treeanc.py:889-930) of node c for this wave's 32 patterns.
__device__ __forceinline__ void pre_node20(const TreeDev &T, const PreDesc &qd, int p, bool act, int h, int i) {
    const int c = qd.c, par = qd.par, ir = qd.ir;
    const auto P32 = T.P32 + (size_t)c * (NS * NS);
    float ro[16], pp[16];
    child_slots(T, ir >= 0 ? ir : -(qd.tr + 1), p, act, h, T.lut, ro);
    if (act) {
        load_groups(T.prof + ((size_t)par * T.P + p) * NS, h, pp);
    } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) pp[r] = 0.0f;
    }
    // child message m_j = sum_i row_i P[i][j]
    const v16f m = contract_in(ro, h, i, P32);
    // outgroup q_j = max(TINY, parent_j) / m_j (its normalisation cancels)
    float q[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) q[r] = acc_state(r, h) < NS ? fmaxf((float)TT_TINY, pp[r]) / m[r] : 0.0f;
    // evolve: msg_k = sum_j P[k][j] q_j
    const v16f msg = contract_out(q, h, i, P32);
    float x[16];
    float part = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        x[r] = (acc_state(r, h) < NS) ? ro[r] * msg[r] : 0.0f;
        part += x[r];
    }
    const float tot = part + __shfl_xor(part, 32, TT_WAVE);
    float best;
    bool mine;
    const int jm = pair_argmax(x, h, best, mine);
    float o[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) o[r] = x[r] / tot;
    if (act) {
        store_groups(T.prof + ((size_t)c * T.P + p) * NS, h, o);
        if (h == 0) T.states[(size_t)c * T.P + p] = (uint8_t)jm;
    }
}

__global__ __launch_bounds__(TT_BLOCK) void k_pre_marginal20(TreeDev T, LevelDev L, int n_ptiles32) {
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int h = lane >> 5, i = lane & 31;
    const int task = blockIdx.x * (TT_BLOCK / TT_WAVE) + w;
    if (task >= L.n_nodes * n_ptiles32) return;  // whole wave exits together
    const int k = task / n_ptiles32, pt = task % n_ptiles32;
    const int p = pt * 32 + i;
    pre_node20(T, load_pre_desc(L.desc + kPreDesc * wave_uniform(k)), p, p < T.P, h, i);
}

// Runs of narrow levels in one launch (see k_post_run): a block of 4 waves
// takes one 32-pattern tile through every level of the run, the level's
// nodes split over the waves, a block barrier between levels.
__global__ __launch_bounds__(TT_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_post_run20(TreeDev T, RunDev R) {
    constexpr int kW = TT_BLOCK / TT_WAVE;
    __shared__ double s_acc[kW][TT_WAVE];
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int h = lane >> 5, i = lane & 31;
    const int p = blockIdx.x * 32 + i;
    const bool act = p < T.P;
    LogAcc acc;
    for (int l = R.l0; l < R.l1; ++l) {
        const int kb = R.off[l], ke = R.off[l + 1];
        for (int k = wave_uniform(kb + w); k < ke; k += kW)
            post_node20(T, load_post_desc(R.desc + kPostDesc * k), p, act, h, i, T.lut, acc);
        __syncthreads();
    }
    s_acc[w][lane] = (act && h == 0) ? acc.log_value() : 0.0;
    __syncthreads();
    if (w == 0 && act && h == 0) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < kW; ++q) s += s_acc[q][lane];
        T.part[(size_t)R.row_first * T.P + p] = s;
        for (int r = R.row_first + 1; r < R.row_end; ++r) T.part[(size_t)r * T.P + p] = 0.0;
    }
}

__global__ __launch_bounds__(TT_BLOCK) void k_pre_run20(TreeDev T, RunDev R) {
    constexpr int kW = TT_BLOCK / TT_WAVE;
    const int lane = threadIdx.x & (TT_WAVE - 1), w = threadIdx.x / TT_WAVE;
    const int h = lane >> 5, i = lane & 31;
    const int p = blockIdx.x * 32 + i;
    const bool act = p < T.P;
    for (int l = R.l0; l < R.l1; ++l) {
        const int kb = R.off[l];
        const int ke = R.tips ? R.off[l + 1] : kb + R.nint[l];
        for (int k = wave_uniform(kb + w); k < ke; k += kW)
            pre_node20(T, load_pre_desc(R.desc + kPreDesc * k), p, act, h, i);
        __syncthreads();
    }
}

}  // namespace

namespace ttk {

hipError_t post_run20(const TreeDev &T, const RunDev &R, hipStream_t s) {
    hipLaunchKernelGGL(k_post_run20, dim3((T.P + 31) / 32), dim3(TT_BLOCK), 0, s, T, R);
    return hipGetLastError();
}

hipError_t pre_run20(const TreeDev &T, const RunDev &R, hipStream_t s) {
    hipLaunchKernelGGL(k_pre_run20, dim3((T.P + 31) / 32), dim3(TT_BLOCK), 0, s, T, R);
    return hipGetLastError();
}

hipError_t post_marginal20(const TreeDev &T, const LevelDev &L, hipStream_t s) {
    const int npt = (T.P + 31) / 32;
    const int groups = (L.n_chunks + 3) / 4;
    hipLaunchKernelGGL(k_post_marginal20, dim3(groups * npt), dim3(TT_BLOCK), 0, s, T, L, npt);
    return hipGetLastError();
}

hipError_t pre_marginal20(const TreeDev &T, const LevelDev &L, hipStream_t s) {
    if (L.n_nodes <= 0) return hipSuccess;
    const int npt = (T.P + 31) / 32;
    const long tasks = (long)L.n_nodes * npt;
    hipLaunchKernelGGL(k_pre_marginal20, dim3((unsigned)((tasks + 3) / 4)), dim3(TT_BLOCK), 0, s, T, L, npt);
    return hipGetLastError();
}

}  // namespace ttk

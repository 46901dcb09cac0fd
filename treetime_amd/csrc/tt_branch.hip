// Joint branch-length optimisation on the device.
//
// Reference path (treeanc.py:1176-1242 optimize_branch_lengths_joint, per
// branch :1244-1270 optimal_branch_length -> :1148-1163 add_branch_state ->
// gtr.py:631-712 state_pair -> :816-920 optimal_t_compressed with the
// objective of :714-747 prob_t_compressed).  Every branch is independent:
// its objective depends only on the (parent state, child state) counts of
// the current reconstruction, so all branches of a tree run in one launch,
// one workgroup per branch:
//   1. the workgroup counts the multiplicity-weighted state pairs of its
//      branch into an LDS histogram (FP64 LDS atomics; multiplicities are
//      integer-valued column counts, so the sums are exact in any order);
//   2. wave 0 runs Brent's method in s = sqrt(t) on
//        f(s) = -sum_ij trunc(count_ij) * log(max(P(s^2)[child j][parent i], 1e-24))
//      with the lanes spread over the n x n state pairs.  Every lane holds
//      the whole Brent state; the objective is reduced with an xor
//      butterfly, which leaves the same bits in every lane, so the control
//      flow is uniform.
#include <math.h>

#include "tt_common.h"

namespace {

// Brent's minimiser (R. P. Brent, Algorithms for Minimization without
// Derivatives, 1973, ch. 5; the form of Numerical Recipes 10.2), with the
// conventions of scipy.optimize.minimize_scalar(method='brent'), which the
// reference calls (scipy 1.15): a given bracket (xa, xb, xc) must satisfy
// xa < xb < xc and f(xb) < min(f(xa), f(xc)); tol1 = tol*|x| + 1e-11; the
// golden-section fraction 0.3819660; at most 500 iterations, and a result
// that took all of them counts as failed.
constexpr double kGoldenFrac = 0.3819660;
constexpr double kAbsTol = 1.0e-11;
constexpr int kMaxIter = 500;
constexpr double kSuperTiny = 1.0e-24;  // config.py SUPERTINY_NUMBER

__device__ inline double wave_total(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    return x;
}

template <int N>
struct PairObjective {
    static constexpr int kBins = N * N;
    static constexpr int kPerLane = (kBins + TT_WAVE - 1) / TT_WAVE;
    double cnt[kPerLane];  // truncated counts of this lane's bins (0 outside)
    int pi_[kPerLane], ch_[kPerLane];
    double lam[N];
    const double *v, *vinv;
    double mu;

    // f(s) = -sum count * log(max(P(s^2)[child][parent], SUPERTINY)),
    // P(t) = v diag(exp(mu t lambda)) v_inv clipped at 0 (gtr.py:1027-1067).
    __device__ double operator()(double s) const {
        const double mt = mu * (s * s);
        double e[N];
#pragma unroll
        for (int m = 0; m < N; ++m) e[m] = exp(mt * lam[m]);
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < kPerLane; ++q) {
            if (cnt[q] != 0.0) {
                const int j = ch_[q], i = pi_[q];
                double p = 0.0;
#pragma unroll
                for (int m = 0; m < N; ++m) p = fma(v[j * N + m], e[m] * vinv[m * N + i], p);
                acc = fma(log(fmax(p, kSuperTiny)), cnt[q], acc);
            }
        }
        return -wave_total(acc);
    }
};

template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_branch_opt(TreeDev T, const int16_t *code_state, int gap,
                                                        const double *eig, const double *v, const double *vinv,
                                                        double mu, double max_t, double tol, double *counts,
                                                        double *t_out, int32_t *status) {
    constexpr int NN = N * N;
    __shared__ double s_cnt[NN];
    const int node = blockIdx.x;
    const int par = T.parent[node];
    if (par < 0) {  // the root has no branch
        if (threadIdx.x == 0) {
            t_out[node] = 0.0;
            status[node] = 0;
        }
        return;
    }
    for (int k = threadIdx.x; k < NN; k += blockDim.x) s_cnt[k] = 0.0;
    __syncthreads();

    // 1. state-pair histogram (gtr.py:672-712).  Tip rows without
    //    reconstructed states (255) use their characters' alphabet index
    //    (code_state, -1 = not counted).
    const int P = T.P;
    const int trow = T.tip_row[node];
    const uint8_t TT_G *cs_row = T.states + (size_t)node * P;
    const uint8_t TT_G *ps_row = T.states + (size_t)par * P;
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        const int ps = ps_row[p];
        int cs = cs_row[p];
        if (cs == 255 && trow >= 0) cs = code_state[T.tip_codes[(size_t)trow * P + p]];
        if (ps >= N || cs < 0 || cs >= N) continue;
        if (ps == gap || cs == gap) continue;
        atomicAdd(&s_cnt[ps * N + cs], T.mult[p]);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NN; k += blockDim.x) counts[(size_t)node * NN + k] = s_cnt[k];
    if (threadIdx.x >= TT_WAVE) return;

    // 2. Brent in s on wave 0 (gtr.py:816-920)
    const int lane = threadIdx.x;
    PairObjective<N> f;
    f.v = v;
    f.vinv = vinv;
    f.mu = mu;
#pragma unroll
    for (int m = 0; m < N; ++m) f.lam[m] = eig[m];
    double diff = 0.0, all = 0.0;
#pragma unroll
    for (int q = 0; q < PairObjective<N>::kPerLane; ++q) {
        const int k = lane + q * TT_WAVE;
        const double c = k < NN ? trunc(s_cnt[k]) : 0.0;  // the reference keeps int counts
        f.cnt[q] = c;
        f.pi_[q] = k < NN ? k / N : 0;
        f.ch_[q] = k < NN ? k % N : 0;
        all += c;
        if (k < NN && k / N != k % N) diff += c;
    }
    all = wave_total(all);
    diff = wave_total(diff);
    const double hamming = diff / all;  // NaN without any counted pair

    const double xa = -sqrt(max_t), xb = sqrt(hamming), xc = sqrt(max_t);
    int st = 0;
    double x = xb;
    if (!(xa < xb && xb < xc)) {
        st = 1;
    } else {
        const double fa = f(xa), fb = f(xb), fc = f(xc);
        if (!(fb < fa && fb < fc)) {
            st = 1;
        } else {
            double a = xa, b = xc;
            double w = xb, vv = xb;
            double fx = fb, fw = fb, fv = fb;
            double dx = 0.0, step = 0.0;
            int it = 0;
            for (; it < kMaxIter; ++it) {
                const double tol1 = tol * fabs(x) + kAbsTol, tol2 = 2.0 * tol1;
                const double mid = 0.5 * (a + b);
                if (fabs(x - mid) < tol2 - 0.5 * (b - a)) break;
                bool golden = fabs(dx) <= tol1;
                if (!golden) {
                    // parabola through (x, fx), (w, fw), (vv, fv)
                    const double r = (x - w) * (fx - fv);
                    double q = (x - vv) * (fx - fw);
                    double p = (x - vv) * q - (x - w) * r;
                    q = 2.0 * (q - r);
                    if (q > 0.0) p = -p;
                    q = fabs(q);
                    const double dx_prev = dx;
                    dx = step;
                    if (p > q * (a - x) && p < q * (b - x) && fabs(p) < fabs(0.5 * q * dx_prev)) {
                        step = p * 1.0 / q;
                        const double u = x + step;
                        if (u - a < tol2 || b - u < tol2) step = mid - x >= 0 ? tol1 : -tol1;
                    } else {
                        golden = true;
                    }
                }
                if (golden) {
                    dx = x >= mid ? a - x : b - x;
                    step = kGoldenFrac * dx;
                }
                const double u = fabs(step) < tol1 ? (step >= 0 ? x + tol1 : x - tol1) : x + step;
                const double fu = f(u);
                if (fu > fx) {
                    if (u < x) a = u;
                    else b = u;
                    if (fu <= fw || w == x) {
                        vv = w;
                        fv = fw;
                        w = u;
                        fw = fu;
                    } else if (fu <= fv || vv == x || vv == w) {
                        vv = u;
                        fv = fu;
                    }
                } else {
                    if (u >= x) a = x;
                    else b = x;
                    vv = w;
                    fv = fw;
                    w = x;
                    fw = fx;
                    x = u;
                    fx = fu;
                }
            }
            if (it >= kMaxIter || isnan(x) || isnan(fx)) st = 2;
        }
    }
    if (lane == 0) {
        // failed searches fall back to the Hamming distance (gtr.py:916-918)
        t_out[node] = st == 0 ? x * x : (st == 2 ? hamming : NAN);
        status[node] = st;
    }
}

}  // namespace

#define TT_DISPATCH_ALL(n, CALL)                          \
    switch (n) {                                         \
        case 2: { constexpr int N_ = 2; CALL; } break;   \
        case 3: { constexpr int N_ = 3; CALL; } break;   \
        case 4: { constexpr int N_ = 4; CALL; } break;   \
        case 5: { constexpr int N_ = 5; CALL; } break;   \
        case 20: { constexpr int N_ = 20; CALL; } break; \
        case 21: { constexpr int N_ = 21; CALL; } break; \
        case 22: { constexpr int N_ = 22; CALL; } break; \
        default: return hipErrorInvalidValue;            \
    }

namespace ttk {
hipError_t branch_opt(const TreeDev &T, const int16_t *code_state, int gap, const double *eig, const double *v,
                      const double *vinv, double mu, double max_t, double tol, double *counts, double *t_out,
                      int32_t *status, hipStream_t s) {
    TT_DISPATCH_ALL(T.n, hipLaunchKernelGGL(k_branch_opt<N_>, dim3((unsigned)T.n_nodes), dim3(TT_BLOCK), 0, s, T,
                                            code_state, gap, eig, v, vinv, mu, max_t, tol, counts, t_out, status));
    return hipGetLastError();
}
}  // namespace ttk

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

// Minimise f in s from the bracket (xa, xb, xc); every caller thread runs the
// same iterates (f returns identical bits in all of them).  st: 0 ok, 1 the
// bracket condition fails, 2 no convergence.
template <class F>
__device__ inline double brent_minimize(F &f, double xa, double xb, double xc, double tol, int &st) {
    st = 0;
    double x = xb;
    if (!(xa < xb && xb < xc)) {
        st = 1;
        return x;
    }
    const double fa = f(xa), fb = f(xb), fc = f(xc);
    if (!(fb < fa && fb < fc)) {
        st = 1;
        return x;
    }
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
    return x;
}

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
                                                        double *t_out, int32_t *status, int search) {
    constexpr int NN = N * N;
    __shared__ double s_cnt[NN];
    const int node = blockIdx.x;
    const int par = T.parent[node];
    if (par < 0) {  // the root has no branch
        if (threadIdx.x == 0 && search) {
            t_out[node] = 0.0;
            status[node] = 0;
        }
        for (int k = threadIdx.x; k < NN; k += blockDim.x) counts[(size_t)node * NN + k] = 0.0;
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
    if (threadIdx.x >= TT_WAVE || !search) return;  // counts only

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

    int st = 0;
    const double x = brent_minimize(f, -sqrt(max_t), sqrt(hamming), sqrt(max_t), tol, st);
    if (lane == 0) {
        // failed searches fall back to the Hamming distance (gtr.py:916-918)
        t_out[node] = st == 0 ? x * x : (st == 2 ? hamming : NAN);
        status[node] = st;
    }
}

// ---------------------------------------------------------------------------
// Marginal branch lengths (treeanc.py:1272-1295 optimal_marginal_branch_length,
// :1297-1358 optimize_tree_marginal): per branch, Brent in s on
//   f(s) = -sum_p mult_p w_p log(pc_p . P(s^2) pp_p + 1e-24) + exp(s^4 / 1e4)
// (gtr.py:816-920 with profiles=True, gtr.py:922-963 prob_t_profiles), where
// pc is the child's marginal_subtree_LH (tips: the character's profile,
// internal nodes: the subtree row normalised to sum 1) and pp its
// marginal_outgroup_LH = normalise(max(TINY, parent profile) / m), m the
// child's message to the parent at the current length (treeanc.py:895-899);
// w_p = (1 - pp_gap)(1 - pc_gap) when the alphabet has a gap, else 1.  For
// the children n1, n2 of a bifurcating root the reference optimises their
// summed length with pc = n1's subtree and pp = normalise(n2's subtree * pi)
// (treeanc.py:1318-1331); that pair is the extra workgroup n_nodes.
//
// One workgroup per branch.  Every evaluation rebuilds P(t) in LDS (FP64,
// clipped at 0) and streams the branch's patterns (profiles recomputed from
// the FP32 rows, arithmetic in FP64); the per-pattern terms are summed in a
// fixed order (per thread, then a fixed LDS tree), so every thread holds the
// same bits and runs the same Brent iterates.

constexpr int kOptThreads = TT_BLOCK;

__device__ inline double block_total(double x, double *s_red) {
    s_red[threadIdx.x] = x;
    __syncthreads();
    for (int w = kOptThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s_red[threadIdx.x] += s_red[threadIdx.x + w];
        __syncthreads();
    }
    const double t = s_red[0];
    __syncthreads();
    return t;
}

// Branch profiles of pattern p: pc (child side) and pp (parent side), FP64;
// M is the branch's P(t) (row i = child state), indexed M[i * N + j].
template <int N, class MP>
__device__ inline void branch_profiles_m(const TreeDev &T, int node, int pair_other, int p, MP M, double (&pc)[N],
                                         double (&pp)[N]) {
    const int ir = T.int_row[node];
    if (ir >= 0) {
        const float TT_G *r = T.sub + ((size_t)ir * T.P + p) * N;
        double z = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            pc[i] = (double)r[i];
            z += pc[i];
        }
#pragma unroll
        for (int i = 0; i < N; ++i) pc[i] /= z;
    } else {
        const int code = T.tip_codes[(size_t)T.tip_row[node] * T.P + p];
#pragma unroll
        for (int i = 0; i < N; ++i) pc[i] = (double)T.lut[code * N + i];
    }
    double z = 0.0;
    if (pair_other >= 0) {  // bifurcating root: the sibling's subtree times pi
        const int jr = T.int_row[pair_other];
        if (jr >= 0) {
            const float TT_G *r = T.sub + ((size_t)jr * T.P + p) * N;
#pragma unroll
            for (int j = 0; j < N; ++j) pp[j] = (double)r[j] * (T.s_pi ? T.s_pi[(size_t)p * N + j] : T.pi[j]);
        } else {
            const int code = T.tip_codes[(size_t)T.tip_row[pair_other] * T.P + p];
#pragma unroll
            for (int j = 0; j < N; ++j)
                pp[j] = (double)T.lut[code * N + j] * (T.s_pi ? T.s_pi[(size_t)p * N + j] : T.pi[j]);
        }
    } else {  // outgroup: max(TINY, parent profile) / (child message at the current length)
        const float TT_G *pr = T.prof + ((size_t)T.parent[node] * T.P + p) * N;
        double r[N];
        if (ir >= 0) {
            const float TT_G *sr = T.sub + ((size_t)ir * T.P + p) * N;
#pragma unroll
            for (int i = 0; i < N; ++i) r[i] = (double)sr[i];
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) r[i] = pc[i];
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double m = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) m = fma(r[i], M[i * N + j], m);
            pp[j] = fmax(TT_TINY, (double)pr[j]) / fmax(TT_TINY, m);
        }
    }
#pragma unroll
    for (int j = 0; j < N; ++j) z += pp[j];
#pragma unroll
    for (int j = 0; j < N; ++j) pp[j] /= z;
}

template <int N>
__device__ inline void branch_profiles(const TreeDev &T, int node, int pair_other, int p, double (&pc)[N],
                                       double (&pp)[N]) {
    branch_profiles_m<N>(T, node, pair_other, p, T.P64 + (size_t)node * (N * N), pc, pp);
}

// P(t) of branch c for pattern p under a site-specific model, FP64, without
// clipping (gtr_site_specific.py:350-371): I + V diag(delta) V^-1 with delta
// interpolated between the branch's grid points.
template <int N>
__device__ inline void site_P64(const TreeDev &T, int c, int p, double (&E)[N * N]) {
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
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double a = (i == j) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k) a = fma(V[i * N + k] * d[k], Vi[k * N + j], a);
            E[i * N + j] = a;
        }
}

// Branch profiles with the branch's own P(t) at its current length, also
// under a site-specific model.
template <int N>
__device__ inline void branch_profiles_any(const TreeDev &T, int node, int pair_other, int p, double (&pc)[N],
                                           double (&pp)[N]) {
    if constexpr (N <= 5) {  // site-specific models exist for up to 5 states
        if (T.s_V64) {
            double E[N * N];
            if (pair_other < 0) site_P64<N>(T, node, p, E);
            branch_profiles_m<N>(T, node, pair_other, p, E, pc, pp);
            return;
        }
    }
    branch_profiles<N>(T, node, pair_other, p, pc, pp);
}

// (t_lo, t_hi, w) of a length t under the site-specific model's
// interpolator (scipy interp1d 'linear': searchsorted left, clipped to
// [1, G - 1]) when t * rate_scale < 10, else (t, t, 0): the exact P(t).
__device__ inline void site_interp(const TreeDev &T, double t, double &ta, double &tb, double &w) {
    const int G = T.s_G;
    if (G >= 2 && t * T.s_rate < 10.0) {
        int lo = 0, hi = G;  // first index with grid >= t
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (T.s_grid[mid] < t)
                lo = mid + 1;
            else
                hi = mid;
        }
        const int h = min(max(lo, 1), G - 1);
        ta = T.s_grid[h - 1];
        tb = T.s_grid[h];
        w = (t - ta) / (tb - ta);
    } else {
        ta = tb = t;
        w = 0.0;
    }
}

// pc . P_p(t) pp for one pattern in the eigenbasis:
// sum_i pc_i pp_i + sum_k (pc V)_k delta_k (V^-1 pp)_k.
template <int N>
__device__ inline double site_pair_lh(const TreeDev &T, int p, const double (&pc)[N], const double (&pp)[N],
                                      double ta, double tb, double w) {
    const auto rho = T.s_rho64 + (size_t)p * N;
    const auto V = T.s_V64 + (size_t)p * (N * N);
    const auto Vi = T.s_Vi64 + (size_t)p * (N * N);
    double res = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) res = fma(pc[i], pp[i], res);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            a = fma(pc[i], V[i * N + k], a);
            b = fma(Vi[k * N + i], pp[i], b);
        }
        double d = expm1(ta * rho[k]);
        if (w != 0.0) d += w * (expm1(tb * rho[k]) - d);
        res = fma(a * d, b, res);
    }
    return res;
}

// Per-pattern statistics for GTR_site_specific.infer (treeanc.py:1556-1589
// with site_specific=True), summed over the branches in node order for each
// pattern p (one thread per pattern):
//   marginal (mode 0): M = normalised pc_i pp_j (P_ij + 1e-24) of the branch
//     (get_branch_mutation_matrix, treeanc.py:1085-1120; the pattern's own P
//     under a site-specific model), n[p][i][j] += mult M_ij,
//     T[p][i] += t/2 mult (sum_j M_ij + sum_j M_ji);
//   joint (mode 1): child state i and parent state j of the reconstruction
//     (tips without reconstructed states: code_state of their character, -1
//     = not an alphabet state), T[p][i] += t mult; for i != j both valid:
//     n[p][i][j] += mult, T[p][j] += t/2 mult, T[p][i] -= t/2 mult.
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_site_counts(TreeDev T, const double *t_eff, int mode,
                                                         const int16_t *code_state, double *n_out, double *t_out) {
    constexpr int NN = N * N;
    const int p = blockIdx.x * TT_BLOCK + threadIdx.x;
    if (p >= T.P) return;
    double n_acc[NN], t_acc[N];
#pragma unroll
    for (int k = 0; k < NN; ++k) n_acc[k] = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) t_acc[k] = 0.0;
    const double mult = T.mult[p];
    for (int c = 0; c < T.n_nodes; ++c) {
        const int par = T.parent[c];
        if (par < 0) continue;
        const double tc = t_eff[c];
        if (mode == 0) {
            double E[NN], pc[N], pp[N];
            if (T.s_V64) {
                site_P64<N>(T, c, p, E);
            } else {
#pragma unroll
                for (int k = 0; k < NN; ++k) E[k] = T.P64[(size_t)c * NN + k];
            }
            branch_profiles_m<N>(T, c, -1, p, E, pc, pp);
            double m[NN], z = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i)
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    m[i * N + j] = pc[i] * pp[j] * (E[i * N + j] + kSuperTiny);
                    z += m[i * N + j];
                }
            const double w = mult / z;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double row = 0.0, col = 0.0;
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    n_acc[i * N + j] = fma(m[i * N + j], w, n_acc[i * N + j]);
                    row += m[i * N + j];
                    col += m[j * N + i];
                }
                t_acc[i] = fma(0.5 * tc * w, row + col, t_acc[i]);
            }
        } else {
            int cs = T.states[(size_t)c * T.P + p];
            const int trow = T.tip_row[c];
            if (cs == 255 && trow >= 0) cs = code_state[T.tip_codes[(size_t)trow * T.P + p]];
            if (cs == 255) cs = -1;
            const int ps = T.states[(size_t)par * T.P + p];
            const bool mut = cs >= 0 && ps < N && cs != ps;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double dt = 0.0;
                if (i == cs) dt += tc * mult - (mut ? 0.5 * tc * mult : 0.0);
                if (mut && i == ps) dt += 0.5 * tc * mult;
                t_acc[i] += dt;
#pragma unroll
                for (int j = 0; j < N; ++j)
                    if (mut && i == cs && j == ps) n_acc[i * N + j] += mult;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NN; ++k) n_out[(size_t)p * NN + k] = n_acc[k];
#pragma unroll
    for (int k = 0; k < N; ++k) t_out[(size_t)p * N + k] = t_acc[k];
}

// P(t) = v diag(exp(mu t lambda)) v_inv clipped at 0, into LDS (gtr.py:1027-1067).
template <int N>
__device__ inline void build_expqt(double t, const double *eig, const double *v, const double *vinv, double mu,
                                   double *s_P) {
    const double mt = mu * t;
    for (int k = threadIdx.x; k < N * N; k += blockDim.x) {
        const int i = k / N, j = k % N;
        double a = 0.0;
        for (int m = 0; m < N; ++m) a = fma(v[i * N + m], exp(mt * eig[m]) * vinv[m * N + j], a);
        s_P[k] = fmax(0.0, a);
    }
    __syncthreads();
}

// sum_p mult_p w_p log(pc_p . P pp_p + 1e-24) over this branch's patterns.
template <int N>
__device__ inline double profile_loglik(const TreeDev &T, int node, int other, int gap, const double *s_P,
                                        double *s_red) {
    double acc = 0.0;
    for (int p = threadIdx.x; p < T.P; p += blockDim.x) {
        double pc[N], pp[N];
        branch_profiles<N>(T, node, other, p, pc, pp);
        double res = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double row = 0.0;
#pragma unroll
            for (int j = 0; j < N; ++j) row = fma(s_P[i * N + j], pp[j], row);
            res = fma(pc[i], row, res);
        }
        double w = T.mult[p];
        if (gap >= 0) w *= (1.0 - pp[gap]) * (1.0 - pc[gap]);
        acc = fma(w, log(res + kSuperTiny), acc);
    }
    return block_total(acc, s_red);
}

// Site-specific model: the same sum with P_p(t) per pattern
// (prob_t_profiles with a (n, n, L) expQt, gtr.py:950-952).
template <int N>
__device__ inline double profile_loglik_site(const TreeDev &T, int node, int other, int gap, double t,
                                             double *s_red) {
    double ta, tb, w;
    site_interp(T, t, ta, tb, w);
    double acc = 0.0;
    for (int p = threadIdx.x; p < T.P; p += blockDim.x) {
        double pc[N], pp[N];
        branch_profiles_any<N>(T, node, other, p, pc, pp);
        const double res = site_pair_lh<N>(T, p, pc, pp, ta, tb, w);
        double wt = T.mult[p];
        if (gap >= 0) wt *= (1.0 - pp[gap]) * (1.0 - pc[gap]);
        acc = fma(wt, log(res + kSuperTiny), acc);
    }
    return block_total(acc, s_red);
}

template <int N>
__global__ __launch_bounds__(kOptThreads) void k_branch_opt_marginal(TreeDev T, int gap, const double *eig,
                                                                     const double *v, const double *vinv, double mu,
                                                                     double max_t, double tol, int pair_n1,
                                                                     int pair_n2, double *t_out, int32_t *status) {
    __shared__ double s_P[N * N];
    __shared__ double s_red[kOptThreads];
    const int b = blockIdx.x;
    int node = b, other = -1;
    if (b == T.n_nodes) {  // the bifurcating root's pair of children
        if (pair_n1 < 0) return;
        node = pair_n1;
        other = pair_n2;
    } else if (T.parent[b] < 0) {
        if (threadIdx.x == 0) {
            t_out[b] = 0.0;
            status[b] = 0;
        }
        return;
    }
    // Hamming distance of the profile pair (gtr.py:860-863)
    double same = 0.0, all = 0.0;
    for (int p = threadIdx.x; p < T.P; p += blockDim.x) {
        double pc[N], pp[N];
        branch_profiles_any<N>(T, node, other, p, pc, pp);
        double d = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) d = fma(pp[j], pc[j], d);
        same = fma(T.mult[p], d, same);
        all += T.mult[p];
    }
    same = block_total(same, s_red);
    all = block_total(all, s_red);
    const double hamming = 1.0 - same / all;

    auto f = [&](double s) {
        double lp;
        if constexpr (N <= 5) {
            if (T.s_V64) return -profile_loglik_site<N>(T, node, other, gap, s * s, s_red) + exp(s * s * s * s / 10000.0);
        }
        build_expqt<N>(s * s, eig, v, vinv, mu, s_P);
        lp = profile_loglik<N>(T, node, other, gap, s_P, s_red);
        return -lp + exp(s * s * s * s / 10000.0);
    };
    int st = 0;
    const double x = brent_minimize(f, -sqrt(max_t), sqrt(hamming), sqrt(max_t), tol, st);
    if (threadIdx.x == 0) {
        t_out[b] = st == 0 ? x * x : (st == 2 ? hamming : NAN);
        status[b] = st;
    }
}

// prob_t_profiles (return_log=True) of every branch at K lengths t[k]
// (BranchLenInterpolator grids, branch_len_interpolator.py:104-110).
template <int N>
__global__ __launch_bounds__(kOptThreads) void k_branch_profile_lh(TreeDev T, int gap, const double *eig,
                                                                   const double *v, const double *vinv, double mu,
                                                                   const double *t, int K, double *out) {
    __shared__ double s_P[N * N];
    __shared__ double s_red[kOptThreads];
    const int node = blockIdx.x;
    if (T.parent[node] < 0) {
        for (int k = threadIdx.x; k < K; k += blockDim.x) out[(size_t)node * K + k] = 0.0;
        return;
    }
    for (int k = 0; k < K; ++k) {
        const double tk = t[k];
        double lp;
        if (tk < 0.0) {
            lp = -1e10;  // config.py BIG_NUMBER (gtr.py:946-947)
        } else {
            bool done = false;
            if constexpr (N <= 5) {
                if (T.s_V64) {
                    lp = profile_loglik_site<N>(T, node, -1, gap, tk, s_red);
                    done = true;
                }
            }
            if (!done) {
                build_expqt<N>(tk, eig, v, vinv, mu, s_P);
                lp = profile_loglik<N>(T, node, -1, gap, s_P, s_red);
            }
        }
        if (threadIdx.x == 0) out[(size_t)node * K + k] = lp;
    }
}

// ---------------------------------------------------------------------------
// GTR inference counts from the marginal profiles (treeanc.py:1085-1120
// get_branch_mutation_matrix, summed as in infer_gtr, treeanc.py:1556-1572):
// for the branch above node c and pattern p,
//   M_ij = pc_i pp_j (P(t_c)_ij + 1e-24) / Z,  Z = sum_ij pc_i pp_j (P_ij + 1e-24)
// (i the child's state, j the parent's), and per branch
//   n_ij += mult_p M_ij,
//   T_j  += t_c / 2 * mult_p sum_i M_ij,   T_i += t_c / 2 * mult_p sum_j M_ij.
// One workgroup per branch; every thread sums its patterns in registers,
// then the block reduces each entry in a fixed order.  Alphabets up to 5
// states (n^2 + n accumulators per thread).
template <int N>
__global__ __launch_bounds__(kOptThreads) void k_mutation_counts(TreeDev T, const double *t_eff, double *out) {
    constexpr int NN = N * N;
    __shared__ double s_red[kOptThreads];
    const int c = blockIdx.x;
    double *o = out + (size_t)c * (NN + N);
    if (T.parent[c] < 0) {
        for (int k = threadIdx.x; k < NN + N; k += blockDim.x) o[k] = 0.0;
        return;
    }
    const double TT_C *E = T.P64 + (size_t)c * NN;
    const double half_t = 0.5 * t_eff[c];
    double n_acc[NN], t_acc[N];
#pragma unroll
    for (int k = 0; k < NN; ++k) n_acc[k] = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) t_acc[k] = 0.0;
    for (int p = threadIdx.x; p < T.P; p += blockDim.x) {
        double pc[N], pp[N];
        branch_profiles<N>(T, c, -1, p, pc, pp);
        double m[NN], z = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j) {
                m[i * N + j] = pc[i] * pp[j] * (E[i * N + j] + kSuperTiny);
                z += m[i * N + j];
            }
        const double w = T.mult[p] / z;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double row = 0.0, col = 0.0;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                n_acc[i * N + j] = fma(m[i * N + j], w, n_acc[i * N + j]);
                row += m[i * N + j];  // sum over the parent state j: child state i
                col += m[j * N + i];  // sum over the child state: parent state i
            }
            t_acc[i] = fma(half_t * w, row + col, t_acc[i]);
        }
    }
#pragma unroll
    for (int k = 0; k < NN; ++k) {
        const double v = block_total(n_acc[k], s_red);
        if (threadIdx.x == 0) o[k] = v;
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double v = block_total(t_acc[k], s_red);
        if (threadIdx.x == 0) o[NN + k] = v;
    }
}

// Mutation counts for wide alphabets (n > 5), the same sums as
// k_mutation_counts: one workgroup per branch walks the patterns in chunks
// of kOptThreads.  Phase 1 (thread = pattern) forms the branch profiles, the
// weight w_p = mult_p / Z_p and the state-time vector
//   tau_i = half_t w_p (pc_i (E pp)_i + pp_i (pc E)_i)
// into LDS; phase 2 (thread = matrix entry (i, j), or state i for tau) adds
// w_p pc_i pp_j E_ij over the chunk's patterns in pattern order, so every
// entry is summed in a fixed order.
template <int N>
__global__ __launch_bounds__(kOptThreads) void k_mutation_counts_wide(TreeDev T, const double *t_eff, double *out) {
    constexpr int NN = N * N;
    __shared__ double s_pc[kOptThreads][N];
    __shared__ double s_pp[kOptThreads][N];
    __shared__ double s_w[kOptThreads];
    __shared__ double s_tau[kOptThreads][N];
    const int c = blockIdx.x;
    double *o = out + (size_t)c * (NN + N);
    if (T.parent[c] < 0) {
        for (int k = threadIdx.x; k < NN + N; k += blockDim.x) o[k] = 0.0;
        return;
    }
    const double TT_C *E = T.P64 + (size_t)c * NN;
    const double half_t = 0.5 * t_eff[c];
    constexpr int kPer = (NN + N + kOptThreads - 1) / kOptThreads;  // outputs per thread
    double acc[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) acc[k] = 0.0;
    for (int p0 = 0; p0 < T.P; p0 += kOptThreads) {
        const int p = p0 + threadIdx.x;
        if (p < T.P) {
            double pc[N], pp[N];
            branch_profiles<N>(T, c, -1, p, pc, pp);
            double z = 0.0;
            double ep[N], pe[N];
#pragma unroll
            for (int i = 0; i < N; ++i) ep[i] = pe[i] = 0.0;
            for (int i = 0; i < N; ++i) {
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    const double e = E[i * N + j] + kSuperTiny;
                    ep[i] = fma(e, pp[j], ep[i]);
                    pe[j] = fma(pc[i], e, pe[j]);
                }
            }
#pragma unroll
            for (int i = 0; i < N; ++i) z = fma(pc[i], ep[i], z);
            const double w = T.mult[p] / z;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                s_pc[threadIdx.x][i] = pc[i];
                s_pp[threadIdx.x][i] = pp[i];
                s_tau[threadIdx.x][i] = half_t * w * (pc[i] * ep[i] + pp[i] * pe[i]);
            }
            s_w[threadIdx.x] = w;
        }
        __syncthreads();
        const int np = min(kOptThreads, T.P - p0);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int q = threadIdx.x + k * kOptThreads;
            if (q < NN) {
                const int i = q / N, j = q % N;
                const double e = E[q] + kSuperTiny;
                double a = acc[k];
                for (int r = 0; r < np; ++r) a = fma(s_w[r] * s_pc[r][i] * s_pp[r][j], e, a);
                acc[k] = a;
            } else if (q < NN + N) {
                const int i = q - NN;
                double a = acc[k];
                for (int r = 0; r < np; ++r) a += s_tau[r][i];
                acc[k] = a;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int q = threadIdx.x + k * kOptThreads;
        if (q < NN + N) o[q] = acc[k];
    }
}

// Rows derived from the last marginal pass, one thread per (node, pattern),
// for the reference's node attributes:
//   kind 1: marginal_subtree_LH (treeanc.py:857-877), internal rows
//           normalised to sum 1, tips the character's profile (seq2prof,
//           treeanc.py:847-853);
//   kind 2: marginal_outgroup_LH (treeanc.py:895-899), normalise(max(TINY,
//           parent profile) / child message); the root's is pi
//           (treeanc.py:817-820);
//   kind 3: marginal_log_Lx = log(marginal_subtree_LH @ P(t))
//           (treeanc.py:862-866, GTR.propagate_profile with return_log; a
//           site-specific model clamps at TINY, gtr_site_specific.py:404-406);
//           the root has none (zeros).
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_node_rows(TreeDev T, int kind, float *out) {
    const size_t gid = (size_t)blockIdx.x * TT_BLOCK + threadIdx.x;
    if (gid >= (size_t)T.n_nodes * T.P) return;
    const int c = (int)(gid / T.P), p = (int)(gid % T.P);
    float *o = out + gid * N;
    double pc[N], pp[N];
    if (T.parent[c] < 0) {
        if (kind == 1) {
            const float TT_G *r = T.sub + ((size_t)T.root_row * T.P + p) * N;
            double z = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) z += (double)r[i];
#pragma unroll
            for (int i = 0; i < N; ++i) o[i] = (float)((double)r[i] / z);
        } else if (kind == 2) {
#pragma unroll
            for (int i = 0; i < N; ++i) o[i] = (float)(T.s_pi ? T.s_pi[(size_t)p * N + i] : T.pi[i]);
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) o[i] = 0.0f;
        }
        return;
    }
    branch_profiles_any<N>(T, c, -1, p, pc, pp);
    if (kind == 1) {
#pragma unroll
        for (int i = 0; i < N; ++i) o[i] = (float)pc[i];
    } else if (kind == 2) {
#pragma unroll
        for (int i = 0; i < N; ++i) o[i] = (float)pp[i];
    } else {
        bool site = false;
        if constexpr (N <= 5) site = T.s_V64 != nullptr;
        if (site) {
            if constexpr (N <= 5) {
                double E[N * N];
                site_P64<N>(T, c, p, E);
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    double m = 0.0;
#pragma unroll
                    for (int i = 0; i < N; ++i) m = fma(pc[i], E[i * N + j], m);
                    o[j] = (float)log(fmax(TT_TINY, m));
                }
            }
        } else {
            const double TT_C *M = T.P64 + (size_t)c * (N * N);
            for (int j = 0; j < N; ++j) {
                double m = 0.0;
#pragma unroll
                for (int i = 0; i < N; ++i) m = fma(pc[i], M[i * N + j], m);
                o[j] = (float)log(m);
            }
        }
    }
}

// get_branch_mutation_matrix (treeanc.py:1085-1120) of the branch above
// node c: out[p][i][j] = pc_i pp_j (P(t)_ij + 1e-24) / Z_p, i the child's
// state, j the parent's; one thread per pattern.
template <int N>
__global__ __launch_bounds__(TT_BLOCK) void k_branch_mutation_matrix(TreeDev T, int c, double *out) {
    const int p = blockIdx.x * TT_BLOCK + threadIdx.x;
    if (p >= T.P) return;
    double pc[N], pp[N];
    branch_profiles_any<N>(T, c, -1, p, pc, pp);
    double *o = out + (size_t)p * (N * N);
    bool site = false;
    if constexpr (N <= 5) site = T.s_V64 != nullptr;
    double z = 0.0;
    if (site) {
        if constexpr (N <= 5) {
            double E[N * N];
            site_P64<N>(T, c, p, E);
#pragma unroll
            for (int i = 0; i < N; ++i)
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    const double m = pc[i] * pp[j] * (E[i * N + j] + kSuperTiny);
                    o[i * N + j] = m;
                    z += m;
                }
        }
    } else {
        const double TT_C *E = T.P64 + (size_t)c * (N * N);
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const double m = pc[i] * pp[j] * (E[i * N + j] + kSuperTiny);
                o[i * N + j] = m;
                z += m;
            }
    }
    const double inv = 1.0 / z;
    for (int k = 0; k < N * N; ++k) o[k] *= inv;
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
hipError_t mutation_counts(const TreeDev &T, const double *t_eff, double *out, hipStream_t s) {
    switch (T.n) {
        case 2: hipLaunchKernelGGL(k_mutation_counts<2>, dim3(T.n_nodes), dim3(kOptThreads), 0, s, T, t_eff, out); break;
        case 3: hipLaunchKernelGGL(k_mutation_counts<3>, dim3(T.n_nodes), dim3(kOptThreads), 0, s, T, t_eff, out); break;
        case 4: hipLaunchKernelGGL(k_mutation_counts<4>, dim3(T.n_nodes), dim3(kOptThreads), 0, s, T, t_eff, out); break;
        case 5: hipLaunchKernelGGL(k_mutation_counts<5>, dim3(T.n_nodes), dim3(kOptThreads), 0, s, T, t_eff, out); break;
        case 20: hipLaunchKernelGGL(k_mutation_counts_wide<20>, dim3(T.n_nodes), dim3(kOptThreads), 0, s, T, t_eff, out); break;
        case 21: hipLaunchKernelGGL(k_mutation_counts_wide<21>, dim3(T.n_nodes), dim3(kOptThreads), 0, s, T, t_eff, out); break;
        case 22: hipLaunchKernelGGL(k_mutation_counts_wide<22>, dim3(T.n_nodes), dim3(kOptThreads), 0, s, T, t_eff, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t node_rows(const TreeDev &T, int kind, float *out, hipStream_t s) {
    const size_t total = (size_t)T.n_nodes * T.P;
    if (!total) return hipSuccess;
    const unsigned blocks = (unsigned)((total + TT_BLOCK - 1) / TT_BLOCK);
    TT_DISPATCH_ALL(T.n, hipLaunchKernelGGL(k_node_rows<N_>, dim3(blocks), dim3(TT_BLOCK), 0, s, T, kind, out));
    return hipGetLastError();
}

hipError_t branch_mutation_matrix(const TreeDev &T, int c, double *out, hipStream_t s) {
    const unsigned blocks = (unsigned)((T.P + TT_BLOCK - 1) / TT_BLOCK);
    TT_DISPATCH_ALL(T.n, hipLaunchKernelGGL(k_branch_mutation_matrix<N_>, dim3(blocks), dim3(TT_BLOCK), 0, s, T, c,
                                            out));
    return hipGetLastError();
}

hipError_t site_counts(const TreeDev &T, const double *t_eff, int mode, const int16_t *code_state, double *n_out,
                       double *t_out, hipStream_t s) {
    const unsigned blocks = (unsigned)((T.P + TT_BLOCK - 1) / TT_BLOCK);
    switch (T.n) {
        case 2: hipLaunchKernelGGL(k_site_counts<2>, dim3(blocks), dim3(TT_BLOCK), 0, s, T, t_eff, mode, code_state, n_out, t_out); break;
        case 3: hipLaunchKernelGGL(k_site_counts<3>, dim3(blocks), dim3(TT_BLOCK), 0, s, T, t_eff, mode, code_state, n_out, t_out); break;
        case 4: hipLaunchKernelGGL(k_site_counts<4>, dim3(blocks), dim3(TT_BLOCK), 0, s, T, t_eff, mode, code_state, n_out, t_out); break;
        case 5: hipLaunchKernelGGL(k_site_counts<5>, dim3(blocks), dim3(TT_BLOCK), 0, s, T, t_eff, mode, code_state, n_out, t_out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t branch_opt_marginal(const TreeDev &T, int gap, const double *eig, const double *v, const double *vinv,
                               double mu, double max_t, double tol, int pair_n1, int pair_n2, double *t_out,
                               int32_t *status, hipStream_t s) {
    TT_DISPATCH_ALL(T.n, hipLaunchKernelGGL(k_branch_opt_marginal<N_>, dim3((unsigned)T.n_nodes + 1),
                                            dim3(kOptThreads), 0, s, T, gap, eig, v, vinv, mu, max_t, tol, pair_n1,
                                            pair_n2, t_out, status));
    return hipGetLastError();
}

hipError_t branch_profile_lh(const TreeDev &T, int gap, const double *eig, const double *v, const double *vinv,
                             double mu, const double *t, int K, double *out, hipStream_t s) {
    TT_DISPATCH_ALL(T.n, hipLaunchKernelGGL(k_branch_profile_lh<N_>, dim3((unsigned)T.n_nodes), dim3(kOptThreads), 0,
                                            s, T, gap, eig, v, vinv, mu, t, K, out));
    return hipGetLastError();
}

hipError_t branch_opt(const TreeDev &T, const int16_t *code_state, int gap, const double *eig, const double *v,
                      const double *vinv, double mu, double max_t, double tol, double *counts, double *t_out,
                      int32_t *status, int search, hipStream_t s) {
    TT_DISPATCH_ALL(T.n, hipLaunchKernelGGL(k_branch_opt<N_>, dim3((unsigned)T.n_nodes), dim3(TT_BLOCK), 0, s, T,
                                            code_state, gap, eig, v, vinv, mu, max_t, tol, counts, t_out, status,
                                            search));
    return hipGetLastError();
}
}  // namespace ttk

/*
 * ttgpu.h -- C ABI of the MI355X (gfx950) Felsenstein-pruning engine.
 *
 * The reference (TreeTime 0.12.1) is pure Python and has no FFI layer: its
 * boundary is the TreeAnc / GTR method surface.  Each entry point below names
 * the reference call it replaces (paths relative to the reference checkout).
 * The Python package treetime_amd binds these with ctypes and keeps the
 * reference's method names and argument meaning on top of them.
 *
 * Conventions
 *   - Every function returns TT_OK (0) or a negative status; the message of
 *     the last failure on the calling thread is tt_last_error().
 *   - Host input arrays are copied; the caller keeps ownership.
 *   - A tt_tree owns its device state (patterns, messages, profiles).
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *   - Host output pointers are filled before the call returns (the call
 *     synchronises its stream when it has host outputs).
 *   - No C++ exception crosses this boundary.  One host thread per tree at a
 *     time; concurrency comes from streams and from tt_batch_*.
 */
#ifndef TTGPU_H
#define TTGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TT_OK = 0,
    TT_EINVAL = -1,         /* bad argument (maps to ValueError / TypeError) */
    TT_ENOMEM_BUDGET = -2,  /* admission would exceed the engine's HBM budget */
    TT_EDEVICE = -3,        /* HIP runtime error */
    TT_ERANGE = -4,         /* mu*t*lambda > 10: reference gtr.py:1042-1047 raises ValueError */
    TT_ESTATE = -5          /* call order violated (e.g. joint results before tt_joint) */
};

/* flags for tt_marginal / tt_joint */
#define TT_RECONSTRUCT_TIPS 1u  /* reconstruct_tip_states=True (treeanc.py:900-903, 1043-1045) */
#define TT_KEEP_NORMALISERS 2u  /* tt_marginal: from this pass on, the postorder also keeps every node's log
                                   normaliser (debug=True, treeanc.py:803-810; see tt_get_subtree_prefactor);
                                   shared models only */

/* which= values for tt_get_profiles */
#define TT_PROFILE_MARGINAL 0   /* node.marginal_profile (treeanc.py:909-917, 822-824) */
#define TT_PROFILE_SUBTREE 1    /* node.marginal_subtree_LH, row-normalised (treeanc.py:877); tips: the
                                   character's profile (treeanc.py:847-853) */
#define TT_PROFILE_OUTGROUP 2   /* node.marginal_outgroup_LH (treeanc.py:895-899); root: pi (:817-820) */
#define TT_PROFILE_LOG_MSG 3    /* node.marginal_log_Lx, the log message to the parent (treeanc.py:862-866,
                                   kept with debug=True, :803-810); root: zeros */

typedef struct tt_engine tt_engine;
typedef struct tt_tree tt_tree;

/* Thread-local text of the last error. */
const char *tt_last_error(void);
/* ABI version: (major << 16) | minor. */
int32_t tt_version(void);

/* ---- engine: one per device, holds the HBM admission budget ------------ */
int32_t tt_engine_create(int32_t device, uint64_t hbm_budget_bytes, tt_engine **out);
/* TT_ESTATE (nothing released) while trees created on the engine are alive. */
int32_t tt_engine_destroy(tt_engine *eng);
int32_t tt_engine_usage(const tt_engine *eng, uint64_t *bytes_in_use, uint64_t *budget);

/* ---- tree: topology + compressed alignment, resident on the device ------
 * Replaces the per-node attribute set-up of TreeAnc (treeanc.py:316-503,
 * 840-853) and the compressed alignment of SequenceData
 * (sequence_data.py:325-464).  Tips are the nodes that have no children; tip
 * rows of tip_codes follow ascending node index.  Children of a node are
 * visited in ascending node index (callers put ladderized order into the
 * numbering).  code_lut[c] is the profile vector of character code c
 * (seq_utils.py:28-122 profile maps).
 */
typedef struct {
    int32_t n_nodes;            /* tips + internal nodes */
    int32_t n_states;           /* alphabet size n */
    int32_t n_patterns;         /* P, compressed alignment length */
    int32_t n_codes;            /* rows of code_lut (<= 256) */
    const int32_t *parent;      /* [n_nodes], -1 for the root */
    const double *branch_len;   /* [n_nodes] effective length (treeanc.py:752-760); root ignored */
    const uint8_t *tip_codes;   /* [n_tips][P] pattern-major character codes */
    const float *code_lut;      /* [n_codes][n_states] */
    const double *multiplicity; /* [P] (sequence_data.py:302-306) */
    const double *pi;           /* [n_states] equilibrium frequencies */
} tt_tree_desc;

int32_t tt_tree_create(tt_engine *eng, const tt_tree_desc *desc, tt_tree **out);
int32_t tt_tree_destroy(tt_tree *tree);
int32_t tt_tree_info(const tt_tree *tree, int32_t *n_levels_post, int32_t *n_levels_pre, uint64_t *device_bytes);
/* How a single-tree marginal / log-likelihood pass runs: tile_patterns is
 * the pattern tile of the one-launch pass (0: level kernels), records_staged
 * 1 when that pass stages its per-node records in LDS, slots_reused 1 when
 * its LDS rows are reused by lifetime (the preorder then re-reads subtree
 * rows from memory); joint_one_launch 1 when the joint pass also runs in one
 * launch over the same tiles.  Introspection for measurement; no
 * counterpart in the reference.  Any output pointer may be NULL. */
int32_t tt_tree_pass_shape(const tt_tree *tree, int32_t *tile_patterns, int32_t *records_staged,
                           int32_t *slots_reused, int32_t *joint_one_launch);

/* GTR eigensystem (gtr.py:612-629): P(t) = v diag(exp(mu t lambda)) v_inv,
 * clipped at 0 (gtr.py:1051-1067).  Rebuilds P(t) of every branch in FP64 on
 * the device.  Replaces GTR.expQt calls inside the passes. */
int32_t tt_tree_set_model(tt_tree *tree, const double *eigvals, const double *v, const double *v_inv, double mu);

/* Site-specific GTR model (GTR_site_specific, gtr_site_specific.py:43-116,
 * 312-371): pattern p has eigenvalues eigvals[p][n], eigenvectors v[p][n][n]
 * and v_inv[p][n][n] (P_p(t) = v diag(exp(mu[p] t eigvals)) v_inv, the
 * reference's v[:, :, p].T and v_inv[:, :, p].T), rate mu[p] and root prior
 * pi[p][n] (treeanc.py:817-820, 1018).  With G >= 2 grid points t_grid[G],
 * branches with t * rate_scale < 10 use P(t) interpolated linearly between
 * the grid points around t (the approximate=True interpolator,
 * gtr_site_specific.py:331-371); other branches, or G = 0, the exact P(t).
 * Replaces tt_tree_set_model until that is called again; tt_marginal,
 * tt_joint, tt_loglik, tt_optimize_branch_lengths_marginal,
 * tt_branch_profile_lh and tt_site_counts then use per-pattern matrices.
 * Alphabets of up to 5 states.  tt_batch_marginal, tt_update_branch, the
 * joint sweep tt_optimize_branch_lengths and tt_mutation_counts refuse such
 * trees (TT_EINVAL). */
int32_t tt_tree_set_site_model(tt_tree *tree, const double *eigvals, const double *v, const double *v_inv,
                               const double *mu, const double *pi, const double *t_grid, int32_t G,
                               double rate_scale);

/* New effective lengths for k branches (node ids); rebuilds their P(t). */
int32_t tt_tree_set_branch_lengths(tt_tree *tree, const int32_t *nodes, const double *t, int32_t k);

/* ---- passes --------------------------------------------------------------
 * tt_marginal: TreeAnc._ml_anc_marginal (treeanc.py:762-932): postorder,
 * root likelihood, preorder with the division trick, argmax states.
 * *logL receives tree.total_sequence_LH (treeanc.py:828). */
int32_t tt_marginal(tt_tree *tree, uint32_t flags, double *logL, void *stream);

/* tt_loglik: postorder + root only (treeanc.py:840-878 then 814-829 with
 * assign_sequence=False), the reference's way to re-evaluate logL. */
int32_t tt_loglik(tt_tree *tree, double *logL, void *stream);

/* tt_joint: TreeAnc._ml_anc_joint (treeanc.py:934-1080); *logL receives
 * tree.sequence_joint_LH (treeanc.py:1037). */
int32_t tt_joint(tt_tree *tree, uint32_t flags, double *logL, void *stream);

/* tt_get_joint_root: after tt_joint, the root's joint log message less its
 * per-pattern maximum, out[P][n] (host) -- normalized_profile of
 * treeanc.py:1017-1019, from which the reference samples the root state
 * (sample_from_profile='root' or True, treeanc.py:1022-1033). */
int32_t tt_get_joint_root(tt_tree *tree, double *out);

/* tt_get_joint_debug: the debug attributes of the last joint pass
 * (treeanc.py:1066-1077): lx[n_nodes][P][n] (host) = node.joint_Lx, the log
 * message of every non-root node to its parent and, at the root, log pi plus
 * the children's messages (treeanc.py:1017); cx (nullable, host
 * [n_nodes][P][n]) = node.joint_Cx, the argmax state of each node per parent
 * state (the root's row is unused).  Recomputed on the device from the
 * stored messages. */
int32_t tt_get_joint_debug(tt_tree *tree, double *lx, uint8_t *cx);

/* tt_joint_from_root: finish the last tt_joint pass from given root states
 * root_states[P] (host, state indices < n): backtrack every node from them
 * and set sequence_LH to the root's joint log message at those states
 * (treeanc.py:1034-1063, np.choose at :1036); *logL receives
 * tree.sequence_joint_LH (:1037).  flags must equal the tt_joint call's.
 * TT_ESTATE unless the last pass on the tree was a joint pass. */
int32_t tt_joint_from_root(tt_tree *tree, uint32_t flags, const uint8_t *root_states, double *logL, void *stream);

/* tt_update_branch: change one branch length and recompute the marginal logL
 * along the path from that branch to the root only.  Equivalent to setting
 * the length and re-running tt_loglik (treeanc.py:1685-1689 pattern).  The
 * first call switches the tree to update mode (the postorder then also keeps
 * each node's log normaliser) and runs one full pass; later calls touch only
 * the root path.  Profiles and states of earlier passes become stale. */
int32_t tt_update_branch(tt_tree *tree, int32_t node, double t, double *logL, void *stream);

/* tt_optimize_branch_lengths: one sweep of joint branch-length optimisation
 * over every branch, from the states of the last marginal or joint
 * reconstruction.  Replaces TreeAnc.optimize_branch_lengths_joint /
 * optimize_branch_len (treeanc.py:1168-1242) with, per branch,
 * optimal_branch_length (:1244-1270), add_branch_state (:1148-1163),
 * GTR.state_pair (gtr.py:631-712) and GTR.optimal_t_compressed (:816-920).
 *   code_state[256]: alphabet index of each character code, used for tips
 *       whose states were not reconstructed; -1 = not counted (the reference
 *       skips characters outside the alphabet for n < 10 and counts them as
 *       state 1 for n >= 10 -- the caller encodes that).
 *   gap_state: state excluded from the counts (ignore_gaps), -1 for none.
 *   max_t, tol: the Brent bracket is (-sqrt(max_t), sqrt(hamming),
 *       sqrt(max_t)) in s = sqrt(t) (config.py MAX_BRANCH_LENGTH), tol 1e-10.
 * Outputs (host, [n_nodes], root entries 0): t_opt = s*s at the minimum;
 * status 0 ok, 1 the bracket condition fails (scipy raises ValueError in the
 * reference; t_opt NaN), 2 no convergence (t_opt = Hamming distance, as the
 * reference falls back); pair_counts (nullable) [n_nodes][n][n] FP64
 * multiplicity-weighted (parent, child) counts before truncation.
 * Branch lengths of the tree are not changed (call
 * tt_tree_set_branch_lengths with the accepted values). */
int32_t tt_optimize_branch_lengths(tt_tree *tree, const int16_t *code_state, int32_t gap_state, double max_t,
                                   double tol, double *t_opt, int32_t *status, double *pair_counts, void *stream);

/* tt_optimize_branch_lengths_marginal: the optimum of every branch given the
 * marginal profiles at its ends, from the last marginal pass.  Replaces
 * TreeAnc.optimal_marginal_branch_length (treeanc.py:1272-1295) for all
 * branches, i.e. GTR.optimal_t_compressed(profiles=True) (gtr.py:816-920) on
 * prob_t_profiles (gtr.py:922-963) with pc = marginal_subtree_LH and
 * pp = marginal_outgroup_LH (treeanc.py:1122-1146); Brent in s = sqrt(t) on
 * -logP(s^2) + exp(s^4 / 1e4), bracket (-sqrt(max_t), sqrt(hamming),
 * sqrt(max_t)), tolerance tol.
 *   gap_state: the alphabet's gap state (prob_t_profiles weighs patterns by
 *       the non-gap probability), -1 for none.
 * Outputs (host): t_opt [n_nodes] (root 0), status [n_nodes] as in
 * tt_optimize_branch_lengths; root_pair (nullable, [2]): when the root has
 * exactly two children, the optimum of their summed length with
 * pp = normalise(second child's subtree * pi), pc = first child's subtree
 * (optimize_tree_marginal, treeanc.py:1318-1331), and its status (-1 when
 * the root is not bifurcating).  TT_ESTATE unless the last pass that touched
 * the profiles was tt_marginal / tt_batch_marginal with the current lengths. */
int32_t tt_optimize_branch_lengths_marginal(tt_tree *tree, int32_t gap_state, double max_t, double tol,
                                            double *t_opt, int32_t *status, double *root_pair, void *stream);

/* tt_branch_profile_lh: GTR.prob_t_profiles(return_log=True) of every branch
 * (profile pair as above) at K branch lengths t[K]: out[n_nodes][K] (host),
 * root rows 0, t < 0 gives -1e10 (gtr.py:946-947).  The grid evaluation of
 * BranchLenInterpolator in marginal mode (branch_len_interpolator.py:104-110). */
int32_t tt_branch_profile_lh(tt_tree *tree, int32_t gap_state, const double *t, int32_t K, double *out,
                             void *stream);

/* tt_branch_state_counts: the multiplicity-weighted (parent state, child
 * state) counts of every branch from the states of the last reconstruction,
 * pair_counts (host) [n_nodes][n][n] -- the counts step of
 * tt_optimize_branch_lengths without the search.  Replaces the per-branch
 * counting that infer_gtr does for a joint reconstruction
 * (treeanc.py:1573-1589, c.mutations and c.cseq) and GTR.state_pair
 * (gtr.py:631-712); code_state and gap_state as there. */
int32_t tt_branch_state_counts(tt_tree *tree, const int16_t *code_state, int32_t gap_state, double *pair_counts,
                               void *stream);

/* tt_mutation_counts: GTR-inference statistics of a marginal reconstruction
 * (treeanc.py:1556-1572 with get_branch_mutation_matrix, :1085-1120): per
 * branch and pattern the joint distribution of (child state i, parent state
 * j), M_ij = pc_i pp_j (P(t)_ij + 1e-24) / Z, summed with the pattern
 * multiplicities into nij[i][j] (host, [n][n]) and the time in each state
 * Ti[i] (host, [n]: half the branch length for the parent's and half for the
 * child's state); per_branch (nullable, host [n_nodes][n*n + n]) gets each
 * branch's share.  Shared (not site-specific) models, every alphabet size;
 * TT_ESTATE unless the profiles are those of the current lengths. */
int32_t tt_mutation_counts(tt_tree *tree, double *nij, double *Ti, double *per_branch, void *stream);

/* tt_branch_mutation_matrix: get_branch_mutation_matrix (treeanc.py:1085-1120,
 * compressed) of the branch above `node` from the last marginal pass:
 * out[P][n][n] (host) = the normalised joint distribution of (child state i,
 * parent state j) per pattern, pc_i pp_j (P(t)_ij + 1e-24) / Z with
 * pc = marginal_subtree_LH and pp = marginal_outgroup_LH of the node
 * (marginal_branch_profile, :1122-1146).  TT_EINVAL for the root, TT_ESTATE
 * unless the profiles are those of the current lengths. */
int32_t tt_branch_mutation_matrix(tt_tree *tree, int32_t node, double *out);

/* Free the joint-pass buffers (Lx, Cx) and return them to the engine's
 * budget; the next tt_joint allocates them again.  No reference counterpart:
 * it lets a server admit requests against an HBM budget. */
int32_t tt_tree_trim(tt_tree *tree);

/* 1 when all work queued on the tree's own stream (passes launched with
 * stream == NULL) has finished, 0 while it runs; for asynchronous callers. */
int32_t tt_tree_query(tt_tree *tree);

/* tree.total_sequence_LH / sequence_joint_LH of the last pass (waits for the
 * tree's own stream). */
int32_t tt_get_total(tt_tree *tree, double *logL);

/* ---- results (device -> host, synchronous) ------------------------------ */
/* tree.sequence_LH: per-pattern log-likelihood of the last pass
 * (marginal: treeanc.py:827, joint: treeanc.py:1036). */
int32_t tt_get_sequence_lh(tt_tree *tree, double *out);
/* Reconstructed state index per node and pattern, [n_nodes][P]; 255 where a
 * node has no reconstruction (tips without TT_RECONSTRUCT_TIPS). */
int32_t tt_get_states(tt_tree *tree, uint8_t *out);
/* Profiles, [n_nodes][P][n_states] (float32): which = TT_PROFILE_*.
 * TT_PROFILE_MARGINAL: rows the last pass did not produce hold an earlier
 * pass's values (or zero).  The other kinds are formed on the device from
 * the last marginal pass (TT_ESTATE if the profiles are stale). */
int32_t tt_get_profiles(tt_tree *tree, int32_t which, float *out);
/* node.marginal_subtree_LH_prefactor (treeanc.py:857-878, kept with
 * debug=True): out[n_nodes][P] (host), the log factor between a node's
 * row-normalised subtree likelihood and the unnormalised one; tips 0.
 * TT_ESTATE unless the last marginal pass ran with TT_KEEP_NORMALISERS (or
 * after tt_update_branch) at the current lengths. */
int32_t tt_get_subtree_prefactor(tt_tree *tree, double *out);

/* Device pointers of the resident buffers (for zero-copy consumers). */
int32_t tt_get_device_buffers(tt_tree *tree, float **profiles, uint8_t **states, double **sequence_lh);

/* ---- primitives ----------------------------------------------------------
 * GTR.propagate_profile (gtr.py:965-995) for transpose=0, GTR.evolve
 * (gtr.py:997-1025) for transpose=1: out = in @ Q (or in @ Q^T), optionally
 * log.  in/out are device buffers of P*n doubles; Q is a host n*n FP64
 * expQt. */
int32_t tt_propagate_profile(const double *in_dev, int32_t P, int32_t n, const double *Q, int32_t transpose,
                             int32_t return_log, double *out_dev, void *stream);

/* GTR_site_specific.propagate_profile / evolve (gtr_site_specific.py:376-437)
 * with one matrix per site: Q_dev[P][n][n] (device, Q[a][i][j] = P_a(t)[i][j]).
 * transpose = 0: out[a][j] = sum_i in[a][i] Q[a][i][j], then log(max(TINY, x))
 * or max(0, x); transpose = 1: out[a][j] = sum_i in[a][i] Q[a][j][i], then
 * log(x) or x.  in/out are device [P][n] FP64. */
int32_t tt_propagate_profile_site(const double *in_dev, int32_t P, int32_t n, const double *Q_dev,
                                  int32_t transpose, int32_t return_log, double *out_dev, void *stream);

/* GTR.prob_t_profiles (gtr.py:922-963) for K transition matrices at once:
 * out[k] = sum_a mult[a] * log(sum_ij pc[a,i] Q_k[i,j] pp[a,j] + 1e-24)
 * (times the non-gap weight when gap_index >= 0).  pp/pc/mult are device
 * buffers; Q is host [K][n][n] FP64; out is host [K]. */
int32_t tt_prob_t_profiles(const double *pp_dev, const double *pc_dev, const double *mult_dev, int32_t P,
                           int32_t n, const double *Q, int32_t K, int32_t gap_index, double *out, void *stream);

/* tt_prob_t_profiles without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as tt_prob_t_profiles (identical values); the K
 * matrices and the partial sums live in a fixed device array of 4 MiB per
 * GPU, so K * (n*n + ceil(P/256) + 1) must not exceed 524288 and K 65535
 * (TT_EINVAL otherwise).  Calls that overlap on one GPU race on that array;
 * calls ordered on one stream do not. */
int32_t tt_prob_t_profiles_noalloc(const double *pp_dev, const double *pc_dev, const double *mult_dev, int32_t P,
                                   int32_t n, const double *Q, int32_t K, int32_t gap_index, double *out,
                                   void *stream);

/* Per-pattern GTR statistics summed over all branches, for the
 * site-specific GTR.infer of TreeAnc.infer_gtr(site_specific=True)
 * (treeanc.py:1556-1589, 1593-1602): n_out[P][n][n] = n_ija (derived state
 * i, ancestral state j) and t_out[P][n] = T_ia, in node order.  marginal != 0:
 * expected counts from the branch pair posteriors of the last marginal pass
 * (get_branch_mutation_matrix, treeanc.py:1085-1120; per-pattern P(t) under
 * a site-specific model); marginal == 0: state changes of the last
 * reconstruction, tips without reconstructed states through code_state[256]
 * (alphabet index of each character, -1 = skipped).  Multiplicity-weighted.
 * Alphabets of up to 5 states.  n_out and t_out are host arrays. */
int32_t tt_site_counts(tt_tree *tree, int32_t marginal, const int16_t *code_state, double *n_out, double *t_out,
                       void *stream);

/* ---- batching --------------------------------------------------------------
 * Marginal reconstruction of k resident trees in ONE persistent launch
 * driven by a device work queue of (tree, pattern-tile) items; each item
 * walks its tile through every level of its tree (per tree the result
 * equals tt_marginal: states and profiles identical, logL to rounding).
 * logL is host [k].  No reference counterpart (the reference runs one tree
 * at a time, treeanc.py:762-932).  The first call for a list of trees and
 * flags builds a plan: items longest tree first, the launch shape chosen by
 * timing the candidates (autotuner), one captured graph. */
int32_t tt_batch_marginal(tt_engine *eng, tt_tree *const *trees, int32_t k, uint32_t flags, double *logL,
                          void *stream);

/* One batched pass over k resident trees (the batching entry of the build
 * plan; no reference counterpart, the reference runs one tree at a time):
 * mode = pass kind | TT_RECONSTRUCT_TIPS.
 *   TT_BATCH_MARGINAL: tt_batch_marginal.
 *   TT_BATCH_LOGLIK:   postorder and root only, the persistent launch with
 *                      its preorder skipped (per tree as tt_loglik: logL and
 *                      sequence_LH; profiles and states are left stale).
 *   TT_BATCH_JOINT:    joint reconstruction (tt_joint per tree) in one
 *                      persistent launch over (tree, tile) items:
 *                      max-product postorder, root, backtrack and the
 *                      log-likelihood of the decoded states; states identical to
 *                      tt_joint, logL to rounding.  tt_tree_trim on a member
 *                      drops the plan (it holds the joint buffers).
 * logL is host [k] (NULL: no read-back). */
#define TT_BATCH_MARGINAL 0x00u
#define TT_BATCH_JOINT 0x10u
#define TT_BATCH_LOGLIK 0x20u
int32_t tt_batch_run(tt_engine *eng, tt_tree *const *trees, int32_t k, uint32_t mode, double *logL, void *stream);

/* The batched pass of tt_batch_run over the trees pool[i], i in members[0..m)
 * only, through the plan of the whole pool (items and launch shape, built and
 * tuned by the first call for (pool, mode), which runs the tuning passes over
 * every pool tree): the members' items are staged to the device and run in
 * one persistent launch.  For the request server, which fuses the pending
 * requests of several trees into one launch (no reference counterpart: the
 * reference runs one tree at a time, treeanc.py:762-1080).  The run waits for
 * the work queued on each member's stream and each member's stream waits for
 * the run, so tt_tree_query, tt_get_total and later per-tree calls see it.
 * Every tree's total (tt_get_total) is written on the device, by this call
 * and by tt_batch_run.  Members must be distinct; logL is host [m] (NULL: no
 * read-back, the call does not wait).  Calls on one plan run in order on the
 * plan's stream (stream NULL) or on the given stream. */
int32_t tt_batch_run_members(tt_engine *eng, tt_tree *const *pool, int32_t k, uint32_t mode, const int32_t *members,
                             int32_t m, double *logL, void *stream);

/* Launch shape the autotuner chose for the plan of (trees, flags): one
 * 4-wave block (0) or one wave (1) per item of the persistent kernel, or each
 * tree's own level-kernel pass in turn (2); 64-pattern tiles per item;
 * resident blocks per CU; the waves per SIMD the kernel's register budget was
 * compiled for (0 = the compiler's own budget); the measured time per launch
 * while tuning (0 if not timed).  Any output pointer may be NULL.  TT_ESTATE
 * when no plan exists yet.  No reference counterpart. */
/* flags: as the mode of tt_batch_run (TT_BATCH_LOGLIK selects the loglik plan). */
int32_t tt_batch_info(tt_engine *eng, tt_tree *const *trees, int32_t k, uint32_t flags, int32_t *wave_items,
                      int32_t *tiles, int32_t *blocks_per_cu, int32_t *waves_per_simd, double *tuned_ms);

/* ---- alignment compression -------------------------------------------------
 * Replaces SequenceData.make_compressed_alignment (sequence_data.py:325-464,
 * dense alignments) and, with TT_FILL_OVERHANGS, the overhang fill of
 * seq2array (seq_utils.py:196-202).  `chars` holds n_seqs rows of row_stride
 * bytes (column c of sequence s at s * row_stride + c, n_seqs * row_stride
 * bytes in all).  ambiguous: the ambiguous character, or -1 for none.
 * Outputs: *n_patterns = P; pattern_of_column[n_columns] =
 * full_to_compressed_sequence_map (sequence_data.py:453-455);
 * patterns[n_seqs][P] (row stride P) = the compressed alignment, ambiguous
 * characters of constant columns rewritten (sequence_data.py:387-394);
 * multiplicity[P] (sequence_data.py:445-447); char_counts[256] (host,
 * nullable) = occurrences of every byte after the fill, from which the caller
 * plans additional constant sites (sequence_data.py:415-442).  patterns and
 * multiplicity must hold n_seqs * capacity and capacity entries; TT_ERANGE
 * (with *n_patterns set, nothing else written) when P > capacity.
 * TT_COMPRESS_DEVICE: chars, pattern_of_column, patterns and multiplicity
 * are device pointers (otherwise host).  Returns when every output is
 * written.  Serialised per engine (one shared scratch buffer). */
#define TT_FILL_OVERHANGS 1u
#define TT_COMPRESS_DEVICE 2u
int32_t tt_compress_alignment(tt_engine *eng, const uint8_t *chars, int32_t n_seqs, int64_t n_columns,
                              int64_t row_stride, int32_t ambiguous, uint32_t flags, int64_t capacity,
                              int64_t *n_patterns, int32_t *pattern_of_column, uint8_t *patterns,
                              double *multiplicity, int64_t *char_counts, void *stream);

/* ---- measurement -------------------------------------------------------------
 * Streaming HBM ceiling of the device with hand-written kernels (no
 * reference counterpart; the bench reports its rooflines against it): kind 0
 * reads `bytes` of src (dst: a sink of at least as many bytes, not written
 * in practice), 1 copies src to dst, 2 writes dst; 16-byte accesses, every
 * CU busy.  *ms = mean time of one sweep over reps sweeps (after one warm-up),
 * timed with events on `stream`.  Buffers are device memory, 16-byte aligned,
 * bytes a multiple of 16. */
int32_t tt_hbm_probe(int32_t kind, void *dst, const void *src, uint64_t bytes, int32_t reps, double *ms,
                     void *stream);

#ifdef __cplusplus
}
#endif
#endif /* TTGPU_H */

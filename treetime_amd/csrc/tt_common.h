// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TT_WAVE 64
#define TT_BLOCK 256
#define TT_TINY 1e-12
#define TT_MAX_STATES 32

// Device buffers of a tree live in global memory; saying so lets the device
// compiler emit global (not flat) loads for pointers read from memory, e.g.
// the per-tree records of the batched kernel.  Host code sees plain pointers
// with the same layout.
//
// Read-only tree structure (child lists, P(t) matrices, level lists) is read
// through the constant address space: uniform loads from it become scalar
// loads into SGPRs, and branches on them become scalar branches.  Nothing
// behind a TT_C pointer is written while a pass runs.
#if defined(__HIP_DEVICE_COMPILE__)
#define TT_G __attribute__((address_space(1)))
#define TT_C __attribute__((address_space(4)))
#else
#define TT_G
#define TT_C
#endif

// Child reference encoding used in the CSR child lists:
//   ref >= 0  -> internal row (index into the [n_int][P][n] row buffers)
//   ref <  0  -> tip row (-ref - 1) of the [n_tips][P] code matrix
struct TreeDev {
    int32_t n_nodes, n_tips, n_int, n, P, n_codes, root, root_row;
    const int32_t TT_C *child_ptr;   // [n_int + 1]
    const int32_t TT_C *child_ref;   // [n_edges]
    const int32_t TT_C *child_node;  // [n_edges] node id of the child (branch index)
    const int32_t TT_C *parent;      // [n_nodes] node ids, -1 for the root
    const int32_t TT_C *int_row;     // [n_nodes] internal row or -1
    const int32_t TT_C *tip_row;     // [n_nodes] tip row or -1
    const int32_t TT_C *int_node;    // [n_int] node id of each internal row
    const uint8_t TT_C *tip_codes;   // [n_tips][P]
    const float TT_C *lut;           // [n_codes][n]   profile of each character code
    const double TT_C *lutj;         // [n_codes][n]   max(profile, TINY): joint messages
    const double TT_C *loglut;       // [n_codes][n]   log(max(profile, TINY)): joint path scores
    const double TT_C *mult;         // [P]
    const double TT_C *pi;           // [n]
    const double TT_C *P64;          // [n_nodes][n][n]  expQt per branch, FP64
    const double TT_C *Pj64;         // [n_nodes][n][n]  max(TINY, expQt): joint messages
    const double TT_C *logP64;       // [n_nodes][n][n]  log(max(TINY, expQt)): joint path scores
    const float TT_C *Pj32;          // [n_nodes][n][n]  max(TINY, expQt) in FP32: joint messages, n > 3
    const float TT_C *lutj32;        // [n_codes][n]   max(profile, TINY) in FP32: joint messages, n > 3
    const float TT_C *P32;           // [n_nodes][n][n]  expQt, FP32
    const float TT_C *D32;           // [n_nodes][n][n]  expQt - I (computed in FP64), FP32
    float TT_G *sub;                 // [n_int][P][n]   subtree partials, max entry = 1
    float TT_G *prof;                // [n_nodes][P][n] marginal profiles (sum = 1)
    uint8_t TT_G *states;            // [n_nodes][P]
    float TT_G *Lx;                  // [n_int][P][n]   joint messages (max-product), max entry = 1
    double TT_G *Lx64;               // [n_int][P][n]   site-specific joint messages (log), max entry = 0
    uint8_t TT_G *Cx;                // joint argmax table: [n_nodes][P] packed (n <= 4) or [n_nodes][P][n]
    double TT_G *part;               // [R][P]          per-chunk-group FP64 log-normaliser (or joint score) sums
    double TT_G *part2;              // [S][P]          row-slice sums
    double TT_G *seqlh;              // [P]  marginal per-pattern logL (root-path updates start here)
    double TT_G *jseqlh;             // [P]  joint per-pattern logL
    double TT_G *bpart;              // [nblocks]       block partials of sum(mult * seqlh)
    double TT_G *total;              // [1]
    double TT_G *offs;               // [n_int][P] per-node log normaliser, written only in update mode
    // joint pass of alphabets of more than 8 states: the tips' character
    // codes numbered 0..n_cc-1 (the codes in use, in code order), and per tip
    // branch and code the tip's message max_i max(TINY, P_ij) max(profile_i,
    // TINY) with its first argmax i, formed at the start of each joint pass
    // (jtab null: not allocated)
    const uint8_t TT_C *tip_cc;      // [n_tips][P] compact code of each tip character
    const float TT_C *lutj_c;        // [n_cc][n]   max(profile, TINY) in FP32 of each compact code
    float TT_G *jtab;                // [n_tips][n_cc][n]
    uint8_t TT_G *jcx;               // [n_tips][n_cc][n]
    int32_t n_cc;
    // site-specific model (all null for a model shared by every pattern):
    // per pattern p, P_p(t) = V diag(exp(t rho)) V^-1 with V [p][i][k],
    // V^-1 [p][k][j], rho [p][k] = mu_p * eigenvalue_k, root prior pi [p][n];
    // per branch (node id c) s_br[3c..3c+2] = (t_lo, t_hi, w): P(t_lo) + w (P(t_hi) - P(t_lo))
    const float TT_C *s_V32;
    const float TT_C *s_Vi32;
    const float TT_C *s_rho32;
    const double TT_C *s_V64;
    const double TT_C *s_Vi64;
    const double TT_C *s_rho64;
    const double TT_C *s_pi;
    const double TT_C *s_br;
    const double TT_C *s_grid;       // [s_G] interpolation grid (s_G = 0: exact P(t) for every length)
    int32_t s_G;
    double s_rate;                   // rate scale: lengths with t * s_rate < 10 are interpolated
};

// One level of a level-scheduled pass.
// Node descriptors, one per entry of the level lists, so a node step needs
// one scalar load instead of a chain (node -> child range -> children):
//   post entry k: desc[8k..8k+7] = {row, first child edge, children,
//                                   child ref 0, child ref 1, child node 0, child node 1, 0}
//   pre entry k:  desc[4k..4k+3] = {node, parent, internal row or -1, tip row or -1}
constexpr int kPostDesc = 8;
constexpr int kPreDesc = 4;

struct LevelDev {
    const int32_t TT_C *nodes;  // internal rows (post) or node ids (pre)
    const int32_t TT_C *desc;   // descriptors of these entries (post or pre layout)
    int32_t n_nodes;       // nodes in this launch
    int32_t chunk;         // nodes per wave (post)
    int32_t n_chunks;
    int32_t row_off;       // first partial row of this level (post)
};

// A run of consecutive narrow levels walked in one launch (one block per
// pattern tile, a block barrier between levels).
struct RunDev {
    const int32_t TT_C *nodes;  // post: internal rows by height; pre: node ids by depth
    const int32_t TT_C *desc;   // descriptors of the whole level list (post or pre layout)
    const int32_t TT_C *off;    // level offsets into nodes
    const int32_t TT_C *nint;   // pre: internal nodes per depth (listed first); unused for post
    int32_t l0, l1;             // levels [l0, l1)
    int32_t row_first, row_end; // post: partial rows of these levels; the first gets the run's sum
    int32_t tips;               // pre: also the tips of each depth
};

namespace ttk {
// launchers (tt_kernels.hip); n is the alphabet size
hipError_t expqt(int n, const double *t, const int32_t *branches, int k, const double *eig, const double *v,
                 const double *vinv, double mu, double *P64, double *Pj64, double *logP64, float *Pj32, float *P32,
                 float *D32, hipStream_t s);
hipError_t post_marginal(const TreeDev &T, const LevelDev &L, hipStream_t s);
hipError_t rowsum(const TreeDev &T, int R, int S, hipStream_t s);
hipError_t root_marginal(const TreeDev &T, int S, hipStream_t s);
hipError_t final_sum(const TreeDev &T, int nblocks, hipStream_t s);
hipError_t pre_marginal(const TreeDev &T, const LevelDev &L, hipStream_t s);
hipError_t post_run(const TreeDev &T, const RunDev &R, hipStream_t s);
// joint pass (n <= 8): runs of narrow postorder levels (no root) and of
// narrow backtrack levels in one launch each, as post_run / pre_run
hipError_t post_joint_run(const TreeDev &T, const RunDev &R, int store_tip_cx, hipStream_t s);
hipError_t backtrack_run(const TreeDev &T, const RunDev &R, hipStream_t s);
hipError_t pre_run(const TreeDev &T, const RunDev &R, hipStream_t s);
hipError_t post_run20(const TreeDev &T, const RunDev &R, hipStream_t s);  // MFMA, n = 20
hipError_t pre_run20(const TreeDev &T, const RunDev &R, hipStream_t s);   // MFMA, n = 20
hipError_t post_marginal20(const TreeDev &T, const LevelDev &L, hipStream_t s);  // MFMA, n = 20
hipError_t pre_marginal20(const TreeDev &T, const LevelDev &L, hipStream_t s);   // MFMA, n = 20
// jtab / jcx of every tip branch (n > 8, jtab allocated)
hipError_t joint_tiptab(const TreeDev &T, const int32_t *tip_nodes, hipStream_t s);
hipError_t post_joint(const TreeDev &T, const LevelDev &L, int store_tip_cx, hipStream_t s);
// site-specific models (tt_site.hip), alphabets of up to 5 states
hipError_t post_site(const TreeDev &T, const LevelDev &L, hipStream_t s);
hipError_t root_site(const TreeDev &T, int S, hipStream_t s);
hipError_t pre_site(const TreeDev &T, const LevelDev &L, hipStream_t s);
hipError_t post_joint_site(const TreeDev &T, const LevelDev &L, int store_tip_cx, hipStream_t s);
hipError_t root_joint_site(const TreeDev &T, int S, int store_tip_cx, hipStream_t s);
// joint logL of the root states already in the root's state row (site-specific
// model, after a joint pass): root log message at that state + node offsets
hipError_t root_given_site(const TreeDev &T, int S, hipStream_t s);
// debug rows of the last joint pass: rel [n_nodes][P][n] and own [n_nodes][P]
// (k_joint_rows), shared model (tt_kernels.hip) or site-specific (tt_site.hip)
hipError_t joint_rows(const TreeDev &T, double *rel, double *own, hipStream_t s);
hipError_t joint_rows_site(const TreeDev &T, double *rel, double *own, hipStream_t s);
hipError_t propagate_site(const double *in, int P, int n, const double *Q, int transpose, int ret_log, double *out,
                          hipStream_t s);
hipError_t root_joint(const TreeDev &T, int store_tip_cx, hipStream_t s);
// log-likelihood of the decoded joint states: slices of G node ids per row
// of part (S rows), then jseqlh and the block partials from S2 row-slice sums
hipError_t joint_score(const TreeDev &T, int S, int G, hipStream_t s);
hipError_t joint_final(const TreeDev &T, int S2, hipStream_t s);
hipError_t backtrack(const TreeDev &T, const LevelDev &L, hipStream_t s);
hipError_t path_update(const TreeDev &T, int node, hipStream_t s);
hipError_t fill_states(const TreeDev &T, const int32_t *nodes, int n_nodes, uint8_t value, hipStream_t s);
hipError_t propagate(const double *in, int P, int n, const double *Q, int transpose, int ret_log, double *out,
                     hipStream_t s);
hipError_t prob_t_profiles(const double *pp, const double *pc, const double *mult, int P, int n, const double *Q,
                           int K, int gap, double *block_out, int nblk, hipStream_t s);
int root_blocks(int P);
bool mfma_enabled();
// one-launch marginal / loglik pass with LDS-resident rows (tt_tile.hip)
int tile_threads();
int tile_post_rec_words(int n);
int tile_pre_rec_words(int n);
size_t tile_stage_words(int n, int n_post, int n_pre, bool sep = false);
size_t tile_lds_bytes(const TreeDev &T, int tp, int n_slots, int n_rounds, size_t stage_words);
int tile_blocks_per_cu(int n, size_t lds, bool staged, bool subg);
// one-launch joint pass over the same tiles (k_tile_joint)
size_t tile_joint_lds_bytes(const TreeDev &T, int tp, int n_slots, int n_rounds, int n_post, int n_pre, bool staged);
int tile_joint_blocks_per_cu(int n, size_t lds, bool staged);
hipError_t tile_pack(const TreeDev &T, const int32_t *post_desc, int n_post, const int32_t *pre_desc, int n_pre,
                     const int32_t *post_slot, const int32_t *pre_slot, bool subg, int32_t *post_rec,
                     int32_t *pre_rec, hipStream_t s);
// arguments of one one-launch pass
struct TileArgs {
    const int32_t *post_rec, *pre_rec, *rounds;
    int Rp, Ri, Ra, tp, n_post, n_pre;
    int n_slots, root_post, root_pre;  // LDS rows per pattern; the root's slot in each phase
    const int32_t *post_slot;          // slot of every internal row
    bool staged, subg;                 // records staged in LDS; slots reused (subtree rows re-read from memory)
    double *partial;
    unsigned *counter;
    bool pre_sep = false;              // staged: the preorder's records in their own LDS region, staged at the start
};
hipError_t tile_marginal(const TreeDev &T, const TileArgs &A, int tips, int pre, hipStream_t s);
hipError_t tile_joint(const TreeDev &T, const TileArgs &A, const int32_t *pre_desc, int tips, bool staged,
                      hipStream_t s);
// sets tt_last_error()'s text on the calling thread and returns code (tt_api.hip)
int32_t set_error(int32_t code, const char *msg);
// joint branch-length optimisation (tt_branch.hip)
hipError_t branch_opt(const TreeDev &T, const int16_t *code_state, int gap, const double *eig, const double *v,
                      const double *vinv, double mu, double max_t, double tol, double *counts, double *t_out,
                      int32_t *status, int search, hipStream_t s);
// marginal branch lengths: Brent on the profile pair of every branch (+ the root pair as entry n_nodes)
hipError_t branch_opt_marginal(const TreeDev &T, int gap, const double *eig, const double *v, const double *vinv,
                               double mu, double max_t, double tol, int pair_n1, int pair_n2, double *t_out,
                               int32_t *status, hipStream_t s);
// GTR inference counts per branch, out [n_nodes][n*n + n] (n_ij then T_i), n <= 5
hipError_t mutation_counts(const TreeDev &T, const double *t_eff, double *out, hipStream_t s);
// rows derived from the last marginal pass for export, [n_nodes][P][n] FP32:
// kind 1 subtree (sum 1), 2 outgroup, 3 log message to the parent (tt_branch.hip)
hipError_t node_rows(const TreeDev &T, int kind, float *out, hipStream_t s);
// per-pattern (child, parent) state distribution of the branch above node c, [P][n][n] FP64
hipError_t branch_mutation_matrix(const TreeDev &T, int c, double *out, hipStream_t s);
// per-pattern statistics summed over branches (mode 0 marginal, 1 joint), n <= 5
hipError_t site_counts(const TreeDev &T, const double *t_eff, int mode, const int16_t *code_state, double *n_out,
                       double *t_out, hipStream_t s);
// prob_t_profiles of every branch at K lengths, out [n_nodes][K]
hipError_t branch_profile_lh(const TreeDev &T, int gap, const double *eig, const double *v, const double *vinv,
                             double mu, const double *t, int K, double *out, hipStream_t s);
}  // namespace ttk

// Joint argmax rows (Cx): for alphabets of up to 4 states one byte per
// (node, pattern), 2 bits per parent state j at bits 2j..2j+1; wider
// alphabets one byte per parent state.
constexpr int kCxPackedMaxN = 4;
__host__ __device__ constexpr bool cx_packed(int n) { return n <= kCxPackedMaxN; }
__host__ __device__ constexpr size_t cx_row_bytes(int n) { return cx_packed(n) ? 1 : (size_t)n; }

// A tree as seen by the persistent batched kernel: the tree buffers plus its
// level schedule.
struct BatchTreeDev {
    TreeDev T;
    const int32_t TT_C *post_nodes;  // internal rows, grouped by height 1..H (root last)
    const int32_t TT_C *post_desc;   // descriptors of post_nodes
    const int32_t TT_C *pre_desc;    // descriptors of pre_nodes
    const int32_t TT_C *post_off;    // [H + 1]
    const int32_t TT_C *pre_nodes;   // non-root node ids grouped by depth 1..D, internal first
    const int32_t TT_C *pre_off;     // [D + 1]
    const int32_t TT_C *pre_nint;    // [D] internal nodes per depth
    const int32_t TT_C *tip_nodes;   // [n_tips]
    int32_t H, D;
    int32_t fill_tips;               // reset tip states to 255 (an earlier pass reconstructed them)
};

// Work queues of the batched launch, one per XCD (accelerator die; each has
// its own L2): the items are split by tree over the dies, a workgroup pops
// from its own die's queue (workgroups are dispatched round-robin over the
// dies, so workgroup b runs on die b % kXcds) and from the others once its
// own is empty.  Layout of the queue buffer (ints): counter of die x at
// x * kQueueStride; item offsets per die ([kXcds + 1]) of the one-tile and
// of the two-tile item lists at kQueueXoff1 and kQueueXoff2.
constexpr int kXcds = 8;
constexpr int kQueueStride = 32;
constexpr int kQueueXoff1 = kXcds * kQueueStride;
constexpr int kQueueXoff2 = kQueueXoff1 + 16;
constexpr int kQueueInts = kQueueXoff2 + 16;

// One work item of the batched launch: patterns [64 tile, 64 (tile + tiles))
// of one tree, tiles per item given by the launch shape.
struct BatchItem {
    int32_t tree, tile;
};

// Launch shape of the persistent batched kernel, picked per plan by the
// autotuner in tt_api.hip.
struct BatchShape {
    int32_t wave_items;     // 1: one wave per item; 0: one 4-wave block per item
    int32_t blocks_per_cu;  // resident blocks per CU the grid is sized for
    int32_t tiles;          // 64-pattern tiles per item (block items: 1 or 2; wave items: 1)
    int32_t waves_per_simd; // register budget of the kernel: 0 = compiler's choice, 5 = fit 5 waves per SIMD
};

namespace ttk {
hipError_t batch_marginal(int n, const BatchTreeDev *trees, const BatchItem *items, int n_items, int *queue,
                          const int *xoff, double *item_out, int tips, const BatchShape &shape, hipStream_t s);
// joint reconstruction of many trees in one persistent launch (block items of one or two tiles)
hipError_t batch_joint(int n, const BatchTreeDev *trees, const BatchItem *items, int n_items, int *queue,
                       const int *xoff, double *item_out, int mode, const BatchShape &shape, hipStream_t s);
// Per-tree log-likelihood of a batched launch: for each of the m entries
// (tree index, first item, item count) of tot, the tree's total = the sum of
// its items' outputs in item order (the order tt_batch_run sums them in).
hipError_t batch_totals(const BatchTreeDev *trees, const int32_t *tot, int m, const double *item_out, hipStream_t s);
int batch_joint_max_blocks_per_cu(int n, int waves_per_simd, int tiles);
// occupancy limit (blocks per CU) of the batched kernel for n states and item shape; 0 if n unsupported
int batch_max_blocks_per_cu(int n, const BatchShape &shape);
// false when only the one-tile block kernel exists for n states
bool batch_all_shapes(int n);
// a register budget (waves per SIMD: 0, 6 or 8) the joint kernel is built with for alphabet size n
bool batch_joint_budget(int n, int waves_per_simd);
int device_cus();
}  // namespace ttk

// Device scratch of one pattern compression (tt_compress.hip): byte offsets
// of its buffers in one allocation, and the row slicing of the column pass.
struct CompressLayout {
    int rows_per_slice, slices, part_pitch, n_blocks;
    size_t off_ends, off_part, off_key, off_small, off_hist, off_blocks, off_col_of_pat, total;
};

namespace ttk {
CompressLayout compress_layout(int n_rows, int L);
// column statistics, column keys and the pattern number of every column that
// opens a pattern; the pattern count is then at compress_total_ptr()
hipError_t compress_columns(const uint8_t *A, int64_t pitch, int n_rows, int L, int amb, bool fill, bool hist,
                            void *scratch, const CompressLayout &l, int *pat, hipStream_t s);
// remaining pattern numbers, multiplicities and the [n_rows][P] compressed rows
// (the pattern count is read on the device; nothing but the column map is
// written when it exceeds capacity)
hipError_t compress_finish(const uint8_t *A, int64_t pitch, int n_rows, int L, int amb, bool fill, void *scratch,
                           const CompressLayout &l, int64_t capacity, int *pat, double *mult, uint8_t *out,
                           hipStream_t s);
const int *compress_total_ptr(void *scratch, const CompressLayout &l);
const unsigned long long *compress_hist_ptr(void *scratch, const CompressLayout &l);
}  // namespace ttk

// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TT_WAVE 64
#define TT_BLOCK 256
#define TT_TINY 1e-12
#define TT_MAX_STATES 32

// Child reference encoding used in the CSR child lists:
//   ref >= 0  -> internal row (index into the [n_int][P][n] row buffers)
//   ref <  0  -> tip row (-ref - 1) of the [n_tips][P] code matrix
struct TreeDev {
    int32_t n_nodes, n_tips, n_int, n, P, n_codes, root, root_row;
    const int32_t *child_ptr;   // [n_int + 1]
    const int32_t *child_ref;   // [n_edges]
    const int32_t *child_node;  // [n_edges] node id of the child (branch index)
    const int32_t *parent;      // [n_nodes] node ids, -1 for the root
    const int32_t *int_row;     // [n_nodes] internal row or -1
    const int32_t *tip_row;     // [n_nodes] tip row or -1
    const int32_t *int_node;    // [n_int] node id of each internal row
    const uint8_t *tip_codes;   // [n_tips][P]
    const float *lut;           // [n_codes][n]   profile of each character code
    const double *loglut;       // [n_codes][n]   log(max(profile, TINY)) for the joint pass
    const double *mult;         // [P]
    const double *pi;           // [n]
    const double *P64;          // [n_nodes][n][n]  expQt per branch, FP64
    const double *logP64;       // [n_nodes][n][n]  log(max(TINY, expQt))
    float *sub;                 // [n_int][P][n]   subtree partials, max entry = 1
    float *prof;                // [n_nodes][P][n] marginal profiles (sum = 1)
    uint8_t *states;            // [n_nodes][P]
    float *Lx;                  // [n_int][P][n]   joint messages, max entry = 0
    uint8_t *Cx;                // [n_nodes][P][n] joint argmax table
    double *part;               // [R][P]          per-chunk-group FP64 log-normaliser sums
    double *part2;              // [S][P]          row-slice sums
    double *seqlh;              // [P]
    double *bpart;              // [nblocks]       block partials of sum(mult * seqlh)
    double *total;              // [1]
    double *offs;               // [n_int][P] per-node log normalisers (only when kept)
};

// One level of a level-scheduled pass.
struct LevelDev {
    const int32_t *nodes;  // internal rows (post) or node ids (pre)
    int32_t n_nodes;       // nodes in this launch
    int32_t chunk;         // nodes per wave (post)
    int32_t n_chunks;
    int32_t row_off;       // first partial row of this level (post)
};

namespace ttk {
// launchers (tt_kernels.hip); n is the alphabet size
hipError_t expqt(int n, const double *t, const int32_t *branches, int k, const double *eig, const double *v,
                 const double *vinv, double mu, double *P64, double *logP64, hipStream_t s);
hipError_t post_marginal(const TreeDev &T, const LevelDev &L, hipStream_t s);
hipError_t rowsum(const TreeDev &T, int R, int S, hipStream_t s);
hipError_t root_marginal(const TreeDev &T, int S, hipStream_t s);
hipError_t final_sum(const TreeDev &T, int nblocks, hipStream_t s);
hipError_t pre_marginal(const TreeDev &T, const LevelDev &L, hipStream_t s);
hipError_t post_joint(const TreeDev &T, const LevelDev &L, int store_tip_cx, hipStream_t s);
hipError_t root_joint(const TreeDev &T, int S, int store_tip_cx, hipStream_t s);
hipError_t backtrack(const TreeDev &T, const LevelDev &L, hipStream_t s);
hipError_t fill_states(const TreeDev &T, const int32_t *nodes, int n_nodes, uint8_t value, hipStream_t s);
hipError_t propagate(const double *in, int P, int n, const double *Q, int transpose, int ret_log, double *out,
                     hipStream_t s);
hipError_t prob_t_profiles(const double *pp, const double *pc, const double *mult, int P, int n, const double *Q,
                           int K, int gap, double *block_out, int nblk, hipStream_t s);
int root_blocks(int P);
}  // namespace ttk

// A tree as seen by the persistent batched kernel: the tree buffers plus its
// level schedule.
struct BatchTreeDev {
    TreeDev T;
    const int32_t *post_nodes;  // internal rows, grouped by height 1..H (root last)
    const int32_t *post_off;    // [H + 1]
    const int32_t *pre_nodes;   // non-root node ids grouped by depth 1..D, internal first
    const int32_t *pre_off;     // [D + 1]
    const int32_t *pre_nint;    // [D] internal nodes per depth
    const int32_t *tip_nodes;   // [n_tips]
    int32_t H, D;
};

// One work item of the batched launch: 64 patterns (tile) of one tree.
struct BatchItem {
    int32_t tree, tile;
};

namespace ttk {
hipError_t batch_marginal(int n, const BatchTreeDev *trees, const BatchItem *items, int n_items, int *queue,
                          double *item_out, int tips, hipStream_t s);
}  // namespace ttk

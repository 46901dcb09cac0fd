// C-ABI of the pruning engine (declared in include/ttgpu.h).
//
// Host side of a tree: flattens the topology into a level schedule
// (postorder levels = node heights, preorder levels = node depths), uploads
// the pattern-major tip codes, owns every device buffer of the tree, and
// records each pass (level kernels + reductions) once into a hipGraph that is
// replayed on later calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ttgpu.h"
#include "tt_common.h"

namespace {

thread_local std::string g_err;

int32_t fail(int32_t code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define TT_HIP(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return fail(TT_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

bool supported_states(int n) { return n == 2 || n == 3 || n == 4 || n == 5 || n == 20 || n == 21 || n == 22; }

// Waves a level launch aims for before it starts giving each wave several
// nodes; 256 CUs x 4 SIMDs x 8 waves.
int g_target_waves() {
    static int v = [] {
        const char *e = getenv("TT_TARGET_WAVES");
        int x = e ? atoi(e) : 8192;
        return x > 0 ? x : 8192;
    }();
    return v;
}

bool g_use_graphs() {
    static bool v = [] {
        const char *e = getenv("TT_NO_GRAPH");
        return !(e && e[0] == '1');
    }();
    return v;
}

}  // namespace

// Error text for entry points defined in other translation units.
int32_t ttk::set_error(int32_t code, const char *msg) { return fail(code, "%s", msg); }

struct BatchPlan;

struct tt_engine {
    int device = 0;
    uint64_t budget = 0;
    std::atomic<uint64_t> used{0};
    std::mutex mu;
    std::vector<BatchPlan *> plans;
    std::atomic<int> live_trees{0};
    // grow-only device buffers of tt_compress_alignment: scratch, and the
    // staging of host inputs and outputs; guarded by cmu
    std::mutex cmu;
    void *cscratch = nullptr, *cstage = nullptr;
    size_t cscratch_bytes = 0, cstage_bytes = 0;
};

// A recorded batched launch over a fixed list of trees.
struct BatchPlan {
    tt_engine *eng = nullptr;
    std::vector<tt_tree *> trees;
    uint32_t flags = 0;
    std::vector<BatchItem> items1, items2;  // items of one and of two 64-pattern tiles
    std::vector<BatchTreeDev> host_trees;   // host copy of d_trees
    BatchTreeDev *d_trees = nullptr;
    BatchItem *d_items1 = nullptr, *d_items2 = nullptr;
    int *d_queue = nullptr;
    double *d_item_out = nullptr;
    size_t bytes = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    BatchShape shape{0, 1, 1, 0};
    double tuned_ms = 0.0;  // time per launch of the chosen shape while tuning (0: not timed)
    bool joint = false;     // joint reconstruction plan (k_batch_joint)
    // trees longest first, the die each tree is dealt to, and each tree's
    // items (contiguous) in items1 / items2
    std::vector<int> order, die_of, first[2], count[2];
    int32_t *d_tot = nullptr;  // [2][k][3]: (tree, first item, item count) for items1, items2
    // member runs (tt_batch_run_members): the members' items, queues and
    // outputs on the device, staged through a ring of pinned host slots
    BatchItem *d_sub_items = nullptr;
    int *d_sub_queue = nullptr;
    double *d_sub_out = nullptr;
    int32_t *d_sub_tot = nullptr;
    static constexpr int kSubSlots = 4;
    char *h_sub[kSubSlots] = {};
    hipEvent_t ev_sub[kSubSlots] = {};
    size_t sub_slot_bytes = 0;
    int sub_next = 0;
    hipEvent_t ev_done = nullptr;
    const std::vector<BatchItem> &items_for(const BatchShape &c) const { return c.tiles == 2 ? items2 : items1; }
    const BatchItem *d_items_for(const BatchShape &c) const { return c.tiles == 2 ? d_items2 : d_items1; }
    const int *d_xoff_for(const BatchShape &c) const { return d_queue + (c.tiles == 2 ? kQueueXoff2 : kQueueXoff1); }
    hipError_t launch(int n, int mode, const BatchShape &c, hipStream_t s) const {
        hipError_t e;
        if (joint)
            e = ttk::batch_joint(n, d_trees, d_items_for(c), (int)items_for(c).size(), d_queue, d_xoff_for(c),
                                 d_item_out, mode, c, s);
        else
            e = ttk::batch_marginal(n, d_trees, d_items_for(c), (int)items_for(c).size(), d_queue, d_xoff_for(c),
                                    d_item_out, mode, c, s);
        const int k = (int)trees.size();
        if (e == hipSuccess) e = ttk::batch_totals(d_trees, d_tot + (c.tiles == 2 ? 3 * k : 0), k, d_item_out, s);
        return e;
    }
    ~BatchPlan() {
        if (graph) (void)hipGraphExecDestroy(graph);
        // d_sub_queue is the base of the member-run block (items, totals and outputs live inside it)
        for (void *q : {(void *)d_trees, (void *)d_items1, (void *)d_items2, (void *)d_queue, (void *)d_item_out,
                        (void *)d_tot, (void *)d_sub_queue})
            if (q) (void)hipFree(q);
        for (int i = 0; i < kSubSlots; ++i) {
            if (ev_sub[i]) (void)hipEventDestroy(ev_sub[i]);
            if (h_sub[i]) (void)hipHostFree(h_sub[i]);
        }
        if (ev_done) (void)hipEventDestroy(ev_done);
        if (stream) (void)hipStreamDestroy(stream);
        eng->used -= bytes;
    }
};

struct PostLv {
    int off, cnt, chunk, nchunks, row_off, rows;
};
struct PreLv {
    int off, n_int, n_all;
};
// consecutive narrow levels launched as one run kernel: levels [l0, l1)
struct LvRun {
    int l0, l1;
};

struct tt_tree {
    tt_engine *eng = nullptr;
    int n_nodes = 0, n_tips = 0, n_int = 0, n = 0, P = 0, n_codes = 0, root = -1, root_row = -1;
    std::vector<int32_t> parent, int_row, tip_row, int_node, child_ptr, child_ref, child_node, tip_nodes;
    std::vector<double> t;
    std::vector<int32_t> post_nodes, pre_nodes;
    std::vector<PostLv> post_lv;  // heights 1..H; the last level is the root
    std::vector<PreLv> pre_lv;    // depths 1..D
    std::vector<LvRun> post_runs, pre_runs[2];  // pre_runs[tips]
    int R_all = 0, R_joint = 0;
    bool have_model = false;
    double mu = 1.0;
    std::vector<double> eig, v, vinv;
    // site-specific model (tt_tree_set_site_model): interpolation grid (empty:
    // exact P(t) for every branch), rate scale, per-branch (t_lo, t_hi, w)
    bool site = false;
    std::vector<double> site_grid, site_br;
    double site_rate_scale = 0.0;
    void *d_site = nullptr;  // one allocation holding every per-pattern site array

    TreeDev d{};
    int32_t *d_post_nodes = nullptr, *d_pre_nodes = nullptr, *d_tip_nodes = nullptr, *d_all_branches = nullptr;
    int32_t *d_post_off = nullptr, *d_pre_off = nullptr, *d_pre_nint = nullptr;
    int32_t *d_post_desc = nullptr, *d_pre_desc = nullptr;
    double *d_t = nullptr, *d_eig = nullptr, *d_v = nullptr, *d_vinv = nullptr;
    double *d_tmp_t = nullptr;
    int32_t *d_tmp_br = nullptr;
    int tmp_cap = 0;
    bool joint_ready = false;
    bool joint_site = false;  // the joint buffers hold site-specific (FP64 log) messages
    bool update_mode = false;  // postorder also stores per-node normalisers (d.offs)
    bool offs_valid = false;   // d.offs and the subtree rows match the current lengths
    hipStream_t stream = nullptr;
    std::map<int, hipGraphExec_t> graphs;
    std::vector<std::pair<void *, size_t>> allocs;
    int last_pass = 0;  // 1 marginal, 2 joint, 3 loglik
    hipStream_t last_stream = nullptr;  // stream of the last pass (result reads wait for it)
    hipEvent_t ev_order = nullptr;      // orders a member run of a batch plan after the tree's own work
    // one-launch marginal / loglik pass with LDS-resident rows (tt_tile.hip):
    // patterns per workgroup, 0 when the level kernels run instead
    int tile_tp = 0;
    int tile_Rp = 0, tile_Ri = 0, tile_Ra = 0;  // rounds: postorder, preorder internal nodes, preorder all
    int32_t *d_tile_rounds = nullptr, *d_tile_post_rec = nullptr, *d_tile_pre_rec = nullptr;
    double *d_tile_partial = nullptr;
    unsigned *d_tile_counter = nullptr;
    bool tile_stale = true;  // the records do not hold the current P(t) yet
    bool tile_staged = false;  // the kernel stages each phase's records in LDS
    bool tile_pre_sep = false;  // staged: the preorder's records in their own LDS region (staged at the start)
    bool tile_subg = false;    // LDS slots are reused; the preorder re-reads subtree rows from memory
    bool tile_joint = false;   // the joint pass runs in one launch over the same tiles (k_tile_joint)
    bool tile_joint_staged = false;  // ... with both phases' descriptors staged in LDS
    int tile_slots = 0, tile_root_post = 0, tile_root_pre = 0;
    int32_t *d_tile_post_slot = nullptr, *d_tile_pre_slot = nullptr;  // slot of every internal row / node
    bool have_states = false;  // d.states holds a marginal or joint reconstruction
    bool tips_dirty = false;   // tip rows of d.states hold reconstructed states (else 255)
    bool profiles_fresh = false;  // d.prof and d.sub belong to the current lengths (last marginal pass)
    double *d_grid_t = nullptr, *d_grid_out = nullptr;
    double *d_mut_counts = nullptr;
    int grid_cap = 0;
    int16_t *d_code_state = nullptr;  // branch optimisation buffers, allocated on first use
    double *d_pair_counts = nullptr, *d_t_opt = nullptr;
    int32_t *d_opt_status = nullptr;
    uint64_t bytes = 0;

    int32_t alloc(void **p, size_t nbytes) {
        nbytes = std::max<size_t>(nbytes, 16);
        uint64_t cur = eng->used.load();
        if (eng->budget && cur + nbytes > eng->budget)
            return fail(TT_ENOMEM_BUDGET, "HBM budget exceeded: %llu in use + %llu requested > budget %llu",
                        (unsigned long long)cur, (unsigned long long)nbytes, (unsigned long long)eng->budget);
        hipError_t e = hipMalloc(p, nbytes);
        if (e != hipSuccess) return fail(TT_EDEVICE, "hipMalloc(%zu): %s", nbytes, hipGetErrorString(e));
        eng->used += nbytes;
        bytes += nbytes;
        allocs.emplace_back(*p, nbytes);
        return TT_OK;
    }
    // free one allocation made by alloc() (returns its bytes to the budget)
    void release(void *p) {
        if (!p) return;
        for (size_t i = 0; i < allocs.size(); ++i)
            if (allocs[i].first == p) {
                (void)hipFree(p);
                eng->used -= allocs[i].second;
                bytes -= allocs[i].second;
                allocs.erase(allocs.begin() + i);
                return;
            }
    }
    template <class T>
    int32_t upload(T **dst, const T *src, size_t count) {
        int32_t rc = alloc((void **)dst, count * sizeof(T));
        if (rc) return rc;
        if (count) TT_HIP(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
        return TT_OK;
    }
    ~tt_tree() {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
        for (auto &a : allocs) {
            (void)hipFree(a.first);
            eng->used -= a.second;
        }
        if (ev_order) (void)hipEventDestroy(ev_order);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

const char *tt_last_error(void) { return g_err.c_str(); }
int32_t tt_version(void) { return (1 << 16) | 0; }

int32_t tt_engine_create(int32_t device, uint64_t hbm_budget_bytes, tt_engine **out) {
    if (!out) return fail(TT_EINVAL, "out is NULL");
    int count = 0;
    TT_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(TT_EINVAL, "device %d out of range (%d devices)", device, count);
    TT_HIP(hipSetDevice(device));
    auto *e = new tt_engine;
    e->device = device;
    e->budget = hbm_budget_bytes;
    *out = e;
    return TT_OK;
}

int32_t tt_engine_destroy(tt_engine *eng) {
    if (eng && eng->live_trees.load() > 0)
        return fail(TT_ESTATE, "engine still has %d trees; destroy them first", eng->live_trees.load());
    if (eng) {
        (void)hipDeviceSynchronize();
        for (BatchPlan *p : eng->plans) delete p;
        if (eng->cscratch) (void)hipFree(eng->cscratch);
        if (eng->cstage) (void)hipFree(eng->cstage);
    }
    delete eng;
    return TT_OK;
}

int32_t tt_engine_usage(const tt_engine *eng, uint64_t *bytes_in_use, uint64_t *budget) {
    if (!eng) return fail(TT_EINVAL, "engine is NULL");
    if (bytes_in_use) *bytes_in_use = eng->used.load();
    if (budget) *budget = eng->budget;
    return TT_OK;
}

}  // extern "C"

namespace {

// Topology -> children lists, heights, depths and the level schedule.
int32_t build_schedule(tt_tree *T, const tt_tree_desc *D) {
    const int nn = D->n_nodes;
    T->parent.assign(D->parent, D->parent + nn);
    std::vector<std::vector<int>> kids(nn);
    int root = -1;
    for (int i = 0; i < nn; ++i) {
        int p = T->parent[i];
        if (p < 0) {
            if (root >= 0) return fail(TT_EINVAL, "more than one root (nodes %d and %d)", root, i);
            root = i;
        } else {
            if (p >= nn) return fail(TT_EINVAL, "parent[%d]=%d out of range", i, p);
            kids[p].push_back(i);
        }
    }
    if (root < 0) return fail(TT_EINVAL, "no root (parent == -1) found");
    // preorder from the root: detects cycles / disconnected nodes
    std::vector<int> order;
    order.reserve(nn);
    std::vector<int> depth(nn, -1);
    order.push_back(root);
    depth[root] = 0;
    for (size_t k = 0; k < order.size(); ++k) {
        for (int c : kids[order[k]]) {
            depth[c] = depth[order[k]] + 1;
            order.push_back(c);
        }
    }
    if ((int)order.size() != nn) return fail(TT_EINVAL, "topology is not a single rooted tree");
    if (kids[root].empty()) return fail(TT_EINVAL, "root has no children");
    T->root = root;
    T->int_row.assign(nn, -1);
    T->tip_row.assign(nn, -1);
    int ni = 0, nt = 0;
    for (int i = 0; i < nn; ++i) {
        if (kids[i].empty())
            T->tip_row[i] = nt++;
        else
            T->int_row[i] = ni++;
    }
    T->n_tips = nt;
    T->n_int = ni;
    T->root_row = T->int_row[root];
    T->int_node.assign(ni, 0);
    T->tip_nodes.clear();
    for (int i = 0; i < nn; ++i) {
        if (T->int_row[i] >= 0) T->int_node[T->int_row[i]] = i;
        if (T->tip_row[i] >= 0) T->tip_nodes.push_back(i);
    }
    T->child_ptr.assign(ni + 1, 0);
    T->child_ref.clear();
    T->child_node.clear();
    for (int r = 0; r < ni; ++r) {
        int v = T->int_node[r];
        for (int c : kids[v]) {
            T->child_ref.push_back(T->int_row[c] >= 0 ? T->int_row[c] : -(T->tip_row[c] + 1));
            T->child_node.push_back(c);
        }
        T->child_ptr[r + 1] = (int)T->child_ref.size();
    }
    // heights (postorder levels)
    std::vector<int> height(nn, 0);
    for (int k = nn - 1; k >= 0; --k) {
        int v = order[k];
        for (int c : kids[v]) height[v] = std::max(height[v], height[c] + 1);
    }
    const int H = height[root];
    std::vector<std::vector<int>> byh(H + 1);
    for (int r = 0; r < ni; ++r) byh[height[T->int_node[r]]].push_back(r);
    const int npt = (D->n_patterns + TT_WAVE - 1) / TT_WAVE;
    T->post_nodes.clear();
    T->post_lv.clear();
    int row = 0;
    for (int h = 1; h <= H; ++h) {
        PostLv L;
        L.off = (int)T->post_nodes.size();
        L.cnt = (int)byh[h].size();
        long waves = (long)L.cnt * npt;
        L.chunk = (int)std::max<long>(1, (waves + g_target_waves() - 1) / g_target_waves());
        if (D->n_states == 20 && ttk::mfma_enabled()) L.chunk = 1;  // the MFMA kernel takes one node per wave
        L.nchunks = (L.cnt + L.chunk - 1) / L.chunk;
        L.rows = (L.nchunks + 3) / 4;
        L.row_off = row;
        row += L.rows;
        for (int r : byh[h]) T->post_nodes.push_back(r);
        T->post_lv.push_back(L);
    }
    T->R_all = row;
    T->R_joint = T->post_lv.back().row_off;  // rows before the root level
    // depths (preorder levels): internal nodes first, then tips
    const int Dm = *std::max_element(depth.begin(), depth.end());
    std::vector<std::vector<int>> byd_i(Dm + 1), byd_t(Dm + 1);
    for (int i = 0; i < nn; ++i) {
        if (i == root) continue;
        (T->int_row[i] >= 0 ? byd_i : byd_t)[depth[i]].push_back(i);
    }
    T->pre_nodes.clear();
    T->pre_lv.clear();
    for (int dd = 1; dd <= Dm; ++dd) {
        PreLv L;
        L.off = (int)T->pre_nodes.size();
        L.n_int = (int)byd_i[dd].size();
        L.n_all = L.n_int + (int)byd_t[dd].size();
        for (int x : byd_i[dd]) T->pre_nodes.push_back(x);
        for (int x : byd_t[dd]) T->pre_nodes.push_back(x);
        T->pre_lv.push_back(L);
    }
    // runs of narrow levels: a level kernel costs a launch (several us), a
    // level inside a run one or two node steps of a 4-wave block
    const char *nr = getenv("TT_NO_RUNS");
    if (!(nr && nr[0] == '1')) {
        const int narrow = (D->n_states == 20 && ttk::mfma_enabled()) ? 8 : 16;
        auto runs = [&](int n_levels, auto width, std::vector<LvRun> &out) {
            out.clear();
            for (int l = 0; l < n_levels;) {
                int e = l;
                while (e < n_levels && width(e) <= narrow) ++e;
                if (e - l >= 2) out.push_back({l, e});
                l = e > l ? e : l + 1;
            }
        };
        runs((int)T->post_lv.size(), [&](int l) { return T->post_lv[l].cnt; }, T->post_runs);
        runs((int)T->pre_lv.size(), [&](int l) { return T->pre_lv[l].n_int; }, T->pre_runs[0]);
        runs((int)T->pre_lv.size(), [&](int l) { return T->pre_lv[l].n_all; }, T->pre_runs[1]);
    }
    return TT_OK;
}

// Two-stage sum of the R partial rows: k_rowsum adds G = ceil(R / S) rows
// into each of S slices, the root kernel adds the S slices; S ~ sqrt(R)
// keeps both chains short.
int rowsum_slices(int R) {
    int s = 1;
    while (s * s < R) ++s;
    return std::min(R, s);
}

// Joint buffers: FP32 product-space messages for a shared model, FP64 log
// messages for a site-specific one (reallocated when the model kind changed).
int32_t ensure_joint(tt_tree *T) {
    int32_t rc;
    if (T->joint_ready && T->joint_site == T->site) return TT_OK;
    if (T->joint_ready && (rc = tt_tree_trim(T))) return rc;
    size_t rows = (size_t)T->n_int * T->P * T->n;
    if (T->site)
        rc = T->alloc((void **)(void *)&T->d.Lx64, rows * sizeof(double));
    else
        rc = T->alloc((void **)(void *)&T->d.Lx, rows * sizeof(float));
    if (rc) return rc;
    if ((rc = T->alloc((void **)(void *)&T->d.Cx, (size_t)T->n_nodes * T->P * cx_row_bytes(T->n)))) {
        // give the message buffer back, so a retry under budget pressure
        // starts from nothing instead of stacking message buffers
        T->release(T->site ? (void *)T->d.Lx64 : (void *)T->d.Lx);
        T->d.Lx = nullptr;
        T->d.Lx64 = nullptr;
        T->d.Cx = nullptr;
        return rc;
    }
    if (!T->site && T->d.n_cc > 0) {  // the tip tables of wide alphabets
        const size_t cells = (size_t)T->n_tips * T->d.n_cc * T->n;
        if ((rc = T->alloc((void **)(void *)&T->d.jtab, cells * sizeof(float))) ||
            (rc = T->alloc((void **)(void *)&T->d.jcx, cells))) {
            for (void *q : {(void *)T->d.jtab, (void *)T->d.Lx, (void *)T->d.Cx})
                if (q) T->release(q);
            T->d.jtab = nullptr;
            T->d.jcx = nullptr;
            T->d.Lx = nullptr;
            T->d.Cx = nullptr;
            return rc;
        }
    }
    T->joint_site = T->site;
    T->joint_ready = true;
    return TT_OK;
}

// Enqueue one pass on stream s.  mode: 1 marginal, 2 joint, 3 loglik,
// 4 the joint pass's backtrack and score from caller-set root states.
// flag of enqueue_pass (not part of the ABI): reset the tip rows of the
// states to 255 because an earlier pass reconstructed them
constexpr uint32_t kFillTips = 0x100u;
// plan flag of a loglik-only (postorder + root) batched pass
constexpr uint32_t kPlanLoglik = 2u;
// plan flag of a joint batched pass
constexpr uint32_t kPlanJoint = 4u;

// Joint pass after the root states are set: backtrack (treeanc.py:1048-1063)
// and the log-likelihood of the decoded states.  given_root: the root states
// came from the caller (sampled, treeanc.py:1022-1037) rather than from the
// root kernel; a site-specific model then takes sequence_LH from the root's
// stored log message at those states.
int32_t enqueue_joint_tail(tt_tree *T, bool tips, bool given_root, hipStream_t s) {
    const TreeDev &D = T->d;
    const int nb = ttk::root_blocks(T->P);
    const std::vector<LvRun> &runs = T->pre_runs[tips ? 1 : 0];
    size_t ri = 0;
    for (size_t d = 0; d < T->pre_lv.size();) {
        if (ri < runs.size() && runs[ri].l0 == (int)d) {  // narrow levels in one launch
            const LvRun &r = runs[ri++];
            RunDev rd{(decltype(RunDev::nodes))T->d_pre_nodes, (decltype(RunDev::desc))T->d_pre_desc,
                      (decltype(RunDev::off))T->d_pre_off, (decltype(RunDev::nint))T->d_pre_nint, r.l0, r.l1, 0, 0,
                      tips ? 1 : 0};
            TT_HIP(ttk::backtrack_run(D, rd, s));
            d = r.l1;
            continue;
        }
        const PreLv &L = T->pre_lv[d++];
        LevelDev ld{(decltype(LevelDev::nodes))(T->d_pre_nodes + L.off),
                    (decltype(LevelDev::desc))(T->d_pre_desc + kPreDesc * L.off), tips ? L.n_all : L.n_int, 1, 0, 0};
        TT_HIP(ttk::backtrack(D, ld, s));
    }
    if (!T->site) {  // logL: FP64 score of the decoded states, slices of G node ids (S <= R_all rows)
        const int G = std::max(16, (T->n_nodes + std::max(T->R_all, 1) - 1) / std::max(T->R_all, 1));
        const int S = (T->n_nodes + G - 1) / G, S2 = rowsum_slices(S);
        TT_HIP(ttk::joint_score(D, S, G, s));
        TT_HIP(ttk::rowsum(D, S, S2, s));
        TT_HIP(ttk::joint_final(D, S2, s));
        TT_HIP(ttk::final_sum(D, nb, s));
    } else if (given_root) {  // part2 still holds the node offsets of the joint pass
        TT_HIP(ttk::root_given_site(D, rowsum_slices(T->R_joint), s));
        TT_HIP(ttk::final_sum(D, nb, s));
    }
    return TT_OK;
}

int32_t enqueue_pass(tt_tree *T, int mode, uint32_t flags, hipStream_t s) {
    const TreeDev &D = T->d;
    const bool tips = (flags & TT_RECONSTRUCT_TIPS) != 0;
    const bool fill = (flags & kFillTips) != 0;
    const int nb = ttk::root_blocks(T->P);
    if ((mode == 1 || mode == 3) && T->tile_tp && !T->site) {
        if (mode == 1 && fill) TT_HIP(ttk::fill_states(D, T->d_tip_nodes, T->n_tips, 255, s));
        ttk::TileArgs A{T->d_tile_post_rec, T->d_tile_pre_rec, T->d_tile_rounds, T->tile_Rp, T->tile_Ri,
                        T->tile_Ra, T->tile_tp, (int)T->post_nodes.size(), (int)T->pre_nodes.size(),
                        T->tile_slots, T->tile_root_post, T->tile_root_pre, T->d_tile_post_slot, T->tile_staged,
                        T->tile_subg, T->d_tile_partial, T->d_tile_counter};
        A.pre_sep = T->tile_pre_sep;
        TT_HIP(ttk::tile_marginal(D, A, tips ? 1 : 0, mode == 1 ? 1 : 0, s));
    } else if (mode == 1 || mode == 3) {
        size_t ri = 0;
        for (size_t h = 0; h < T->post_lv.size();) {
            if (!T->site && ri < T->post_runs.size() && T->post_runs[ri].l0 == (int)h) {
                const LvRun &r = T->post_runs[ri++];
                const PostLv &last = T->post_lv[r.l1 - 1];
                RunDev rd{(decltype(RunDev::nodes))T->d_post_nodes, (decltype(RunDev::desc))T->d_post_desc,
                          (decltype(RunDev::off))T->d_post_off, nullptr,
                          r.l0, r.l1, T->post_lv[r.l0].row_off, last.row_off + last.rows, 0};
                TT_HIP(ttk::post_run(D, rd, s));
                h = r.l1;
                continue;
            }
            const PostLv &L = T->post_lv[h++];
            LevelDev ld{(decltype(LevelDev::nodes))(T->d_post_nodes + L.off),
                        (decltype(LevelDev::desc))(T->d_post_desc + kPostDesc * L.off), L.cnt, L.chunk, L.nchunks,
                        L.row_off};
            TT_HIP(T->site ? ttk::post_site(D, ld, s) : ttk::post_marginal(D, ld, s));
        }
        const int S = rowsum_slices(T->R_all);
        TT_HIP(ttk::rowsum(D, T->R_all, S, s));
        TT_HIP(T->site ? ttk::root_site(D, S, s) : ttk::root_marginal(D, S, s));
        TT_HIP(ttk::final_sum(D, nb, s));
        if (mode == 1) {
            if (fill) TT_HIP(ttk::fill_states(D, T->d_tip_nodes, T->n_tips, 255, s));
            const std::vector<LvRun> &runs = T->pre_runs[tips ? 1 : 0];
            size_t ri = 0;
            for (size_t d = 0; d < T->pre_lv.size();) {
                if (!T->site && ri < runs.size() && runs[ri].l0 == (int)d) {
                    const LvRun &r = runs[ri++];
                    RunDev rd{(decltype(RunDev::nodes))T->d_pre_nodes, (decltype(RunDev::desc))T->d_pre_desc,
                              (decltype(RunDev::off))T->d_pre_off,
                              (decltype(RunDev::nint))T->d_pre_nint, r.l0, r.l1, 0, 0, tips ? 1 : 0};
                    TT_HIP(ttk::pre_run(D, rd, s));
                    d = r.l1;
                    continue;
                }
                const PreLv &L = T->pre_lv[d++];
                LevelDev ld{(decltype(LevelDev::nodes))(T->d_pre_nodes + L.off),
                            (decltype(LevelDev::desc))(T->d_pre_desc + kPreDesc * L.off), tips ? L.n_all : L.n_int, 1,
                            0, 0};
                TT_HIP(T->site ? ttk::pre_site(D, ld, s) : ttk::pre_marginal(D, ld, s));
            }
        }
    } else if (mode == 4) {
        if (int32_t rc = enqueue_joint_tail(T, tips, true, s)) return rc;
    } else if (T->tile_joint && !T->site) {  // one launch: postorder, root, backtrack and score
        if (fill) TT_HIP(ttk::fill_states(D, T->d_tip_nodes, T->n_tips, 255, s));
        ttk::TileArgs A{T->d_tile_post_rec, T->d_tile_pre_rec, T->d_tile_rounds, T->tile_Rp, T->tile_Ri,
                        T->tile_Ra, T->tile_tp, (int)T->post_nodes.size(), (int)T->pre_nodes.size(),
                        T->tile_slots, T->tile_root_post, T->tile_root_pre, T->d_tile_post_slot, T->tile_staged,
                        T->tile_subg, T->d_tile_partial, T->d_tile_counter};
        TT_HIP(ttk::tile_joint(D, A, T->d_pre_desc, tips ? 1 : 0, T->tile_joint_staged, s));
    } else {
        if (!T->site) TT_HIP(ttk::joint_tiptab(D, T->d_tip_nodes, s));
        const int H = (int)T->post_lv.size() - 1;  // the root level runs on its own
        size_t ri = 0;
        for (int h = 0; h < H;) {
            // narrow levels below the root in one launch (shared model, n <= 8;
            // the site-specific pass keeps per-level normaliser rows)
            while (ri < T->post_runs.size() && T->post_runs[ri].l0 < h) ++ri;
            if (!T->site && T->n <= 8 && ri < T->post_runs.size() && T->post_runs[ri].l0 == h &&
                std::min(T->post_runs[ri].l1, H) - h >= 2) {
                const int l1 = std::min(T->post_runs[ri].l1, H);
                RunDev rd{(decltype(RunDev::nodes))T->d_post_nodes, (decltype(RunDev::desc))T->d_post_desc,
                          (decltype(RunDev::off))T->d_post_off, nullptr, h, l1, 0, 0, 0};
                TT_HIP(ttk::post_joint_run(D, rd, tips ? 1 : 0, s));
                h = l1;
                ++ri;
                continue;
            }
            auto &L = T->post_lv[h++];
            LevelDev ld{(decltype(LevelDev::nodes))(T->d_post_nodes + L.off),
                        (decltype(LevelDev::desc))(T->d_post_desc + kPostDesc * L.off), L.cnt, L.chunk, L.nchunks,
                        L.row_off};
            TT_HIP(T->site ? ttk::post_joint_site(D, ld, tips ? 1 : 0, s) : ttk::post_joint(D, ld, tips ? 1 : 0, s));
        }
        if (T->site) {  // logL from the root and the node normalisers
            const int S = rowsum_slices(T->R_joint);
            TT_HIP(ttk::rowsum(D, T->R_joint, S, s));
            TT_HIP(ttk::root_joint_site(D, S, tips ? 1 : 0, s));
            TT_HIP(ttk::final_sum(D, nb, s));
        } else {
            TT_HIP(ttk::root_joint(D, tips ? 1 : 0, s));
        }
        if (fill) TT_HIP(ttk::fill_states(D, T->d_tip_nodes, T->n_tips, 255, s));
        if (int32_t rc = enqueue_joint_tail(T, tips, false, s)) return rc;
    }
    return TT_OK;
}

// Result reads wait for the tree's own stream and the stream of its last
// pass, not for the whole device.
int32_t wait_results(tt_tree *T) {
    TT_HIP(hipSetDevice(T->eng->device));
    if (T->last_stream && T->last_stream != T->stream) TT_HIP(hipStreamSynchronize(T->last_stream));
    TT_HIP(hipStreamSynchronize(T->stream));
    return TT_OK;
}

int32_t run_pass(tt_tree *T, int mode, uint32_t flags, double *logL, void *stream) {
    if (!T) return fail(TT_EINVAL, "tree is NULL");
    if (!T->have_model) return fail(TT_ESTATE, "tt_tree_set_model has not been called");
    TT_HIP(hipSetDevice(T->eng->device));
    if (mode == 2) {
        int32_t rc = ensure_joint(T);
        if (rc) return rc;
    }
    hipStream_t s = stream ? (hipStream_t)stream : T->stream;
    const bool tips = (flags & TT_RECONSTRUCT_TIPS) != 0;
    const bool fill = !tips && T->tips_dirty && (mode == 1 || mode == 2);
    flags = (flags & TT_RECONSTRUCT_TIPS) | (fill ? kFillTips : 0u);
    const int key = mode * 8 + (tips ? 1 : 0) + (fill ? 2 : 0);
    if (T->tile_tp && !T->site && (mode == 1 || mode == 3 || (mode == 2 && T->tile_joint)) && T->tile_stale) {
        TT_HIP(ttk::tile_pack(T->d, T->d_post_desc, (int)T->post_nodes.size(), T->d_pre_desc,
                              (int)T->pre_nodes.size(), T->d_tile_post_slot, T->d_tile_pre_slot, T->tile_subg,
                              T->d_tile_post_rec, T->d_tile_pre_rec, s));
        T->tile_stale = false;
    }
    // a one-launch pass is one kernel (two when tip states are cleared): it
    // starts sooner launched directly than as a graph replay (~4-5 us less
    // per pass on the 200- and 5000-tip trees); passes of level kernels
    // replay their captured graphs
    const bool one_launch = T->tile_tp && !T->site && (mode == 1 || mode == 3 || (mode == 2 && T->tile_joint));
    if (g_use_graphs() && !one_launch) {
        auto it = T->graphs.find(key);
        if (it == T->graphs.end()) {
            hipGraph_t g = nullptr;
            TT_HIP(hipStreamBeginCapture(T->stream, hipStreamCaptureModeThreadLocal));
            int32_t rc = enqueue_pass(T, mode, flags, T->stream);
            hipError_t e = hipStreamEndCapture(T->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(TT_EDEVICE, "graph capture: %s", hipGetErrorString(e));
            hipGraphExec_t ex = nullptr;
            e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return fail(TT_EDEVICE, "graph instantiate: %s", hipGetErrorString(e));
            it = T->graphs.emplace(key, ex).first;
        }
        TT_HIP(hipGraphLaunch(it->second, s));
    } else {
        int32_t rc = enqueue_pass(T, mode, flags, s);
        if (rc) return rc;
    }
    T->last_pass = mode == 4 ? 2 : mode;
    T->last_stream = s;
    if (mode == 1 || mode == 2) {
        T->have_states = true;
        T->tips_dirty = tips;
    }
    if (mode == 1) T->profiles_fresh = true;
    if (mode == 3) T->profiles_fresh = false;  // subtree rows rewritten, profiles not
    if (T->update_mode && (mode == 1 || mode == 3)) T->offs_valid = true;
    if (logL) {
        TT_HIP(hipMemcpyAsync(logL, T->d.total, sizeof(double), hipMemcpyDeviceToHost, s));
        TT_HIP(hipStreamSynchronize(s));
    }
    return TT_OK;
}

// Switch a tree to update mode: per-node normalisers are kept so that a
// branch change can be re-evaluated along its root path only.
int32_t enable_update_mode(tt_tree *T) {
    if (T->update_mode) return TT_OK;
    int32_t rc;
    if ((rc = T->alloc((void **)(void *)&T->d.offs, (size_t)T->n_int * T->P * sizeof(double)))) return rc;
    for (auto &g : T->graphs) (void)hipGraphExecDestroy(g.second);  // recorded without d.offs
    T->graphs.clear();
    T->update_mode = true;
    T->offs_valid = false;
    return TT_OK;
}

int32_t refresh_expqt(tt_tree *T, const int32_t *d_branches, const double *d_t, int k, hipStream_t s) {
    TT_HIP(ttk::expqt(T->n, d_t, d_branches, k, T->d_eig, T->d_v, T->d_vinv, T->mu, (double *)T->d.P64,
                      (double *)T->d.Pj64, (double *)T->d.logP64, (float *)T->d.Pj32, (float *)T->d.P32, (float *)T->d.D32, s));
    T->tile_stale = true;
    return TT_OK;
}

int32_t check_range(tt_tree *T, const double *t, int k) {
    double lmax = -INFINITY;
    for (double x : T->eig) lmax = std::max(lmax, x);
    for (int i = 0; i < k; ++i) {
        if (!(t[i] >= 0.0) && !(t[i] < 0.0)) return fail(TT_EINVAL, "branch length is NaN");
        if (T->mu * t[i] * lmax > 10.0)
            return fail(TT_ERANGE,
                        "Error in computing exp(Q * t): Q has positive eigenvalues or the branch length t is too "
                        "large (mu*t*lambda = %g)",
                        T->mu * t[i] * lmax);
    }
    return TT_OK;
}

int32_t refuse_site(const tt_tree *T, const char *what) {
    return T->site ? fail(TT_EINVAL, "%s is not available for site-specific models", what) : TT_OK;
}

void destroy_graphs(tt_tree *T) {
    for (auto &g : T->graphs) (void)hipGraphExecDestroy(g.second);
    T->graphs.clear();
}

// (t_lo, t_hi, w) of every branch for the site-specific model: the
// reference's interp1d over the grid (gtr_site_specific.py:331-371) when
// t * rate_scale < 10 -- lo/hi around t as searchsorted (left side) clipped
// to [1, G - 1] picks them -- else the exact P(t) as (t, t, 0).
int32_t upload_site_branches(tt_tree *T) {
    const int nn = T->n_nodes;
    T->site_br.assign((size_t)3 * nn, 0.0);
    const std::vector<double> &g = T->site_grid;
    const int G = (int)g.size();
    for (int c = 0; c < nn; ++c) {
        const double t = T->t[c];
        double *b = &T->site_br[3 * (size_t)c];
        if (G >= 2 && t * T->site_rate_scale < 10.0) {
            int hi = (int)(std::lower_bound(g.begin(), g.end(), t) - g.begin());
            hi = std::min(std::max(hi, 1), G - 1);
            b[0] = g[hi - 1];
            b[1] = g[hi];
            b[2] = (t - g[hi - 1]) / (g[hi] - g[hi - 1]);
        } else {
            b[0] = b[1] = t;
            b[2] = 0.0;
        }
    }
    TT_HIP(hipMemcpyAsync((void *)T->d.s_br, T->site_br.data(), T->site_br.size() * sizeof(double),
                          hipMemcpyHostToDevice, T->stream));
    return TT_OK;
}

// Per-branch outputs of the branch-length sweeps: [n_nodes + 1] (the extra
// entry is the bifurcating root's pair of children).
int32_t ensure_opt_out(tt_tree *T) {
    if (T->d_t_opt) return TT_OK;
    int32_t rc;
    if ((rc = T->alloc((void **)&T->d_t_opt, (size_t)(T->n_nodes + 1) * sizeof(double)))) return rc;
    if ((rc = T->alloc((void **)&T->d_opt_status, (size_t)(T->n_nodes + 1) * sizeof(int32_t)))) return rc;
    return TT_OK;
}

// wave_items value of the shape that runs every tree's own level-kernel
// graph in turn instead of the persistent kernel (best for very few, very
// large trees, whose items would each walk thousands of levels alone)
constexpr int kPerTreeGraphs = 2;

// mode: bit 0 reconstruct tips, bit 1 loglik only (postorder and root)
int32_t launch_plan(BatchPlan *P, int n, int mode, const BatchShape &c, hipStream_t s) {
    if (c.wave_items == kPerTreeGraphs) {
        const int pass = P->joint ? 2 : ((mode & 2) ? 3 : 1);
        for (tt_tree *T : P->trees)
            if (int32_t rc = run_pass(T, pass, (mode & 1) ? TT_RECONSTRUCT_TIPS : 0u, nullptr, s)) return rc;
        return TT_OK;
    }
    TT_HIP(P->launch(n, mode, c, s));
    return TT_OK;
}

// Autotuner of the batched launch: every candidate shape (item = one 4-wave
// block over two 64-pattern tiles with the compiler's register budget or one
// fitting 5 waves per SIMD, one 4-wave block over one tile, or one wave over
// one tile; resident blocks per CU at the occupancy limit or half of it; or
// each tree's own level-kernel graph in turn) runs
// once to warm up and twice timed with HIP events on the plan's own trees;
// the fastest is kept in the plan.  TT_BATCH_MODE=block|wave|plain and
// TT_BATCH_BLOCKS_PER_CU=k restrict the candidates; TT_AUTOTUNE=0 skips the
// timing and takes the first candidate (one block per two-tile item with the
// compiler's register budget, the fastest on the 256-tree nucleotide mix).  Shapes differ only in how items are spread over the chip and in
// the order the per-node normalisers of an item are multiplied, so states
// and profiles are identical for every shape and logL agrees to rounding.
int32_t tune_batch(BatchPlan *P, int n, int tips, hipStream_t s) {
    std::vector<BatchShape> cand;
    const char *mode = getenv("TT_BATCH_MODE");
    const char *bpc = getenv("TT_BATCH_BLOCKS_PER_CU");
    const char *tune = getenv("TT_AUTOTUNE");
    const char *jw = getenv("TT_JOINT_WAVES_PER_SIMD");  // joint plans: keep only this register budget
    const char *tl = getenv("TT_BATCH_TILES");           // keep only block items of this many tiles
    const bool log = getenv("TT_AUTOTUNE_LOG") != nullptr;  // print every candidate's time to stderr
    auto allowed = [&](int wave_items) {
        if (!mode || !mode[0]) return true;
        const int want = mode[0] == 'b' ? 0 : mode[0] == 'w' ? 1 : mode[0] == 'p' ? kPerTreeGraphs : -1;
        return wave_items == want;
    };
    // the joint kernel has block items of one or two tiles, with the
    // compiler's register budget or one that fits 6 or 8 waves per SIMD
    const std::vector<BatchShape> kinds =
        P->joint ? std::vector<BatchShape>{{0, 0, 1, 8}, {0, 0, 2, 8}, {0, 0, 1, 6}, {0, 0, 2, 6}, {0, 0, 1, 0},
                                           {0, 0, 2, 0}}
                 : std::vector<BatchShape>{{0, 0, 2, 0}, {0, 0, 2, 5}, {0, 0, 1, 0}, {1, 0, 1, 0}};
    // a pinned joint register budget must exist for this alphabet: refuse it
    // rather than silently falling back to another launch path
    if (P->joint && jw && jw[0] && !ttk::batch_joint_budget(n, atoi(jw)))
        return fail(TT_EINVAL,
                    "TT_JOINT_WAVES_PER_SIMD=%s: the batched joint kernel has no register budget of %s waves per "
                    "SIMD for an alphabet of %d states",
                    jw, jw, n);
    for (BatchShape kd : kinds) {
        if (!allowed(kd.wave_items)) continue;
        if (P->joint && jw && jw[0] && kd.waves_per_simd != atoi(jw)) continue;
        if (tl && tl[0] && kd.tiles != atoi(tl)) continue;
        if (P->joint ? !ttk::batch_joint_budget(n, kd.waves_per_simd)
                     : !ttk::batch_all_shapes(n) && !(kd.wave_items == 0 && kd.tiles == 1))
            continue;
        const int occ = P->joint ? ttk::batch_joint_max_blocks_per_cu(n, kd.waves_per_simd, kd.tiles)
                                 : ttk::batch_max_blocks_per_cu(n, kd);
        std::vector<int> per_cu;
        if (bpc && atoi(bpc) > 0) {
            per_cu.push_back(atoi(bpc));
        } else {
            per_cu.push_back(occ);
            if (occ > 1) per_cu.push_back(occ / 2);
        }
        for (int x : per_cu) {
            kd.blocks_per_cu = x;
            cand.push_back(kd);
        }
    }
    if (allowed(kPerTreeGraphs) && !(bpc && atoi(bpc) > 0)) cand.push_back({kPerTreeGraphs, 0, 1, 0});
    if (cand.empty()) return fail(TT_EINVAL, "TT_BATCH_MODE=%s selects no launch shape", mode);
    P->shape = cand[0];
    P->tuned_ms = 0.0;
    if (cand.size() == 1 || (tune && tune[0] == '0')) return TT_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TT_HIP(hipEventCreate(&e0));
    TT_HIP(hipEventCreate(&e1));
    double best = INFINITY;
    hipError_t e = hipSuccess;
    int32_t rc = TT_OK;
    for (const BatchShape &c : cand) {
        rc = launch_plan(P, n, tips, c, s);
        if (rc) break;
        e = hipEventRecord(e0, s);
        for (int r = 0; r < 2 && e == hipSuccess && !rc; ++r) rc = launch_plan(P, n, tips, c, s);
        if (rc) break;
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) break;
        if (log)
            fprintf(stderr, "tt autotune: wave_items=%d tiles=%d waves_per_simd=%d blocks_per_cu=%d %.3f ms\n",
                    c.wave_items, c.tiles, c.waves_per_simd, c.blocks_per_cu, ms / 2.0);
        if (ms / 2.0 < best) {
            best = ms / 2.0;
            P->shape = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (e != hipSuccess) return fail(TT_EDEVICE, "batch autotune: %s", hipGetErrorString(e));
    P->tuned_ms = best;
    return TT_OK;
}

}  // namespace

extern "C" {

// Rounds of the one-launch pass: each level's entries in chunks of at most
// cap, as [first, end) pairs.
void tile_rounds(const tt_tree *T, int cap, std::vector<int32_t> &out, int &Rp, int &Ri, int &Ra) {
    out.clear();
    auto add = [&](int first, int end) {
        for (int b = first; b < end; b += cap) {
            out.push_back(b);
            out.push_back(std::min(end, b + cap));
        }
    };
    for (const PostLv &L : T->post_lv) add(L.off, L.off + L.cnt);
    Rp = (int)out.size() / 2;
    for (const PreLv &L : T->pre_lv) add(L.off, L.off + L.n_int);
    Ri = (int)out.size() / 2 - Rp;
    for (const PreLv &L : T->pre_lv) add(L.off, L.off + L.n_all);
    Ra = (int)out.size() / 2 - Rp - Ri;
}

// LDS slots of the one-launch pass with reuse (level-based lifetimes): an
// internal row occupies its slot from its postorder level to its parent's
// (the root's to the end); an internal node's profile from its preorder
// depth to the next depth (its children read it).  A slot is handed to a new
// row only at a level after the last one that reads the old row.  Returns
// the number of slots of each phase; post_slot is indexed by row, pre_slot
// by node (-1 for tips).
void tile_slots(const tt_tree *T, std::vector<int32_t> &post_slot, std::vector<int32_t> &pre_slot, int &n_post,
                int &n_pre) {
    auto assign = [](int n_levels, const std::vector<std::vector<int>> &by_level, const std::vector<int> &last,
                     std::vector<int32_t> &slot) {
        // active: (last level that reads it, slot)
        std::vector<std::pair<int, int>> active;
        std::vector<int> free_slots;
        int count = 0;
        for (int l = 0; l < n_levels; ++l) {
            std::vector<std::pair<int, int>> keep;
            for (auto &a : active) (a.first < l ? free_slots.push_back(a.second) : keep.push_back(a));
            active.swap(keep);
            for (int x : by_level[l]) {
                int sl;
                if (!free_slots.empty()) {
                    sl = free_slots.back();
                    free_slots.pop_back();
                } else {
                    sl = count++;
                }
                slot[x] = sl;
                active.push_back({last[x], sl});
            }
        }
        return count;
    };
    const int H = (int)T->post_lv.size();
    std::vector<int> lvl(T->n_int, 0), last(T->n_int, H);
    std::vector<std::vector<int>> by_level(H);
    for (int h = 0; h < H; ++h)
        for (int k = T->post_lv[h].off; k < T->post_lv[h].off + T->post_lv[h].cnt; ++k) {
            lvl[T->post_nodes[k]] = h;
            by_level[h].push_back(T->post_nodes[k]);
        }
    for (int r = 0; r < T->n_int; ++r) {
        const int par = T->parent[T->int_node[r]];
        last[r] = par < 0 ? H : lvl[T->int_row[par]];
    }
    post_slot.assign(T->n_int, -1);
    n_post = assign(H, by_level, last, post_slot);
    // preorder: depth 0 is the root, depth d + 1 the nodes of pre_lv[d]
    const int D = (int)T->pre_lv.size();
    std::vector<std::vector<int>> by_depth(D + 1);
    std::vector<int> lastp(T->n_nodes, 0);
    by_depth[0].push_back(T->root);
    lastp[T->root] = 1;
    for (int d = 0; d < D; ++d)
        for (int k = T->pre_lv[d].off; k < T->pre_lv[d].off + T->pre_lv[d].n_int; ++k) {
            const int c = T->pre_nodes[k];
            by_depth[d + 1].push_back(c);
            lastp[c] = d + 2;
        }
    pre_slot.assign(T->n_nodes, -1);
    n_pre = assign(D + 1, by_depth, lastp, pre_slot);
}

// Choose the one-launch LDS-resident pass for a tree (nucleotide-sized
// alphabets, shared model) and its power-of-two pattern tile: rows of the
// tile in a workgroup's LDS, and every workgroup resident at once (beyond
// that the level kernels stream the tree at full bandwidth).
int32_t setup_tile(tt_tree *T) {
    T->tile_tp = 0;
    const char *e = getenv("TT_TILE");
    if (e && e[0] == '0') return TT_OK;
    const bool forced = e && e[0] == '1';
    if (T->n < 2 || T->n > 5 || T->n_int < 1) return TT_OK;
    int dev = 0, lds_max = 0, n_cu = 0;
    TT_HIP(hipGetDevice(&dev));
    TT_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    TT_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    // dynamic LDS up to the device's per-workgroup limit needs no opt-in here
    const size_t budget = (size_t)std::max(0, lds_max - 2048);  // the kernel's static LDS
    // Tile choice by an estimate of the pass time: every round (a level, or
    // a chunk of a wide one) costs about a microsecond of one wave's latency
    // chain, a quarter more when its records stream from memory; staging the
    // records costs every workgroup a copy of them through L2 (about 2 TB/s
    // for the whole device), streaming them costs every workgroup one read
    // of them (about 4 TB/s); re-reading subtree rows from memory (slots
    // reused) adds about half a microsecond per preorder round; a second wave
    // of workgroups runs the rounds twice.  Measured on the 200-tip tree: 8
    // patterns per workgroup with staged records beat 2 (staging traffic) and
    // 32 (rounds).  Slots: one per internal row (the preorder reads subtree
    // rows from LDS) or reused by lifetime (fewer rows per pattern, so larger
    // trees fit; tile_slots).
    // TT_TILE=1 also allows two waves of workgroups; TT_TILE=0 turns the path
    // off; TT_TILE_STAGE=0 never stages the records; TT_TILE_TP fixes the
    // tile; TT_TILE_REUSE=0 / 1 never / always reuses slots.
    const char *we = getenv("TT_TILE_WAVES");  // waves of workgroups allowed (default 1, 2 with TT_TILE=1)
    const int kTileMaxWaves = we && atoi(we) > 0 ? atoi(we) : (forced ? 2 : 1);
    const char *se = getenv("TT_TILE_STAGE");
    const bool may_stage = !(se && se[0] == '0');
    const char *te = getenv("TT_TILE_TP");
    const int only_c = te ? atoi(te) : 0;
    const char *re = getenv("TT_TILE_REUSE");
    const int reuse_pin = re && re[0] ? (re[0] == '1' ? 1 : 0) : -1;
    const size_t stage_words = ttk::tile_stage_words(T->n, (int)T->post_nodes.size(), (int)T->pre_nodes.size());
    // records every workgroup streams through L2 when they are not staged
    const double stream_bytes = 4.0 * ((double)T->post_nodes.size() * ttk::tile_post_rec_words(T->n) +
                                       (double)T->pre_nodes.size() * ttk::tile_pre_rec_words(T->n));
    std::vector<int32_t> post_live, pre_live;
    int n_post_live = 0, n_pre_live = 0;
    tile_slots(T, post_live, pre_live, n_post_live, n_pre_live);
    const int live_slots = std::max(n_post_live, n_pre_live);
    std::vector<int32_t> rounds;
    int Rp = 0, Ri = 0, Ra = 0, tp = 0;
    bool staged = false, subg = false;
    double best_us = 1e300;
    for (int c = 1; c <= 64; c *= 2) {
        if (only_c > 0 && c != only_c) continue;
        tile_rounds(T, ttk::tile_threads() / c, rounds, Rp, Ri, Ra);
        bool fits = false;
        for (int rs = 0; rs < 2; ++rs) {
            if (reuse_pin >= 0 && rs != reuse_pin) continue;
            const int n_slots = rs ? live_slots : T->n_int;
            for (int st = may_stage ? 1 : 0; st >= 0; --st) {
                const size_t lds = ttk::tile_lds_bytes(T->d, c, n_slots, (int)rounds.size(), st ? stage_words : 0);
                if (lds > budget) continue;
                const int bpc = ttk::tile_blocks_per_cu(T->n, lds, st != 0, rs != 0);
                if (bpc < 1) continue;
                fits = true;
                const long blocks = (T->P + c - 1) / c;
                const long waves = (blocks + (long)bpc * n_cu - 1) / ((long)bpc * n_cu);
                if (waves > kTileMaxWaves) continue;
                const double us = (double)waves * ((Rp + Ri) * (st ? 1.0 : 1.25) + (rs ? 0.5 * Ri : 0.0)) +
                                  (double)blocks * (st ? stage_words * 4.0 / 2.0e6 : stream_bytes / 4.0e6);
                if (us < best_us) {
                    best_us = us;
                    tp = c;
                    staged = st != 0;
                    subg = rs != 0;
                }
            }
        }
        if (!fits) break;  // no larger tile fits either
    }
    if (!tp) return TT_OK;
    const int blocks = (T->P + tp - 1) / tp;
    tile_rounds(T, ttk::tile_threads() / tp, rounds, Rp, Ri, Ra);
    // a staged pass keeps the preorder's records in an LDS region of their
    // own, staged with everything else at the start (no memory read at the
    // root step), when that fits with no fewer workgroups resident
    // (TT_TILE_PRE_SEP=0: never)
    T->tile_pre_sep = false;
    if (staged) {
        const char *pe = getenv("TT_TILE_PRE_SEP");
        const int n_slots = subg ? live_slots : T->n_int;
        const size_t l0 = ttk::tile_lds_bytes(T->d, tp, n_slots, (int)rounds.size(), stage_words);
        const size_t l1 = ttk::tile_lds_bytes(
            T->d, tp, n_slots, (int)rounds.size(),
            ttk::tile_stage_words(T->n, (int)T->post_nodes.size(), (int)T->pre_nodes.size(), true));
        if (!(pe && pe[0] == '0') && l1 <= budget) {
            const int b0 = ttk::tile_blocks_per_cu(T->n, l0, true, subg);
            const int b1 = ttk::tile_blocks_per_cu(T->n, l1, true, subg);
            T->tile_pre_sep = b1 >= 1 && (b1 >= b0 || (long)b1 * n_cu >= blocks);
        }
    }
    // slot maps: reused by lifetime, or one slot per internal row (the
    // preorder then finds each node's subtree row in its own slot)
    std::vector<int32_t> post_slot, pre_slot;
    if (subg) {
        post_slot = post_live;
        pre_slot = pre_live;
        T->tile_slots = live_slots;
    } else {
        post_slot.resize(T->n_int);
        for (int r = 0; r < T->n_int; ++r) post_slot[r] = r;
        pre_slot.assign(T->int_row.begin(), T->int_row.end());
        T->tile_slots = T->n_int;
    }
    T->tile_root_post = post_slot[T->root_row];
    T->tile_root_pre = pre_slot[T->root];
    int32_t rc;
    if ((rc = T->upload(&T->d_tile_rounds, rounds.data(), rounds.size()))) return rc;
    if ((rc = T->upload(&T->d_tile_post_slot, post_slot.data(), post_slot.size()))) return rc;
    if ((rc = T->upload(&T->d_tile_pre_slot, pre_slot.data(), pre_slot.size()))) return rc;
    if ((rc = T->alloc((void **)&T->d_tile_post_rec,
                       (size_t)T->post_nodes.size() * ttk::tile_post_rec_words(T->n) * sizeof(int32_t))))
        return rc;
    if ((rc = T->alloc((void **)&T->d_tile_pre_rec,
                       (size_t)T->pre_nodes.size() * ttk::tile_pre_rec_words(T->n) * sizeof(int32_t))))
        return rc;
    if ((rc = T->alloc((void **)&T->d_tile_partial, (size_t)blocks * sizeof(double)))) return rc;
    if ((rc = T->alloc((void **)&T->d_tile_counter, sizeof(unsigned)))) return rc;
    T->tile_Rp = Rp;
    T->tile_Ri = Ri;
    T->tile_Ra = Ra;
    T->tile_tp = tp;
    T->tile_staged = staged;
    T->tile_subg = subg;
    T->tile_stale = true;
    // the joint pass takes the same tiles in one launch when its LDS (message
    // rows, both phases' descriptors, argmax rows and states) fits
    // (TT_TILE_JOINT=0: level kernels)
    // (descriptors staged in LDS when they fit too, else streamed)
    T->tile_joint = false;
    const char *ej = getenv("TT_TILE_JOINT");
    const char *js = getenv("TT_TILE_JOINT_STAGE");  // 1 / 0: only the staged / streamed descriptors
    if (!(ej && ej[0] == '0')) {
        for (bool st : {true, false}) {
            if (js && js[0] && (js[0] == '1') != st) continue;
            const size_t jl = ttk::tile_joint_lds_bytes(T->d, tp, T->tile_slots, 2 * (Rp + Ri + Ra),
                                                        (int)T->post_nodes.size(), (int)T->pre_nodes.size(), st);
            if (jl <= budget && ttk::tile_joint_blocks_per_cu(T->n, jl, st) >= 1) {
                T->tile_joint = true;
                T->tile_joint_staged = st;
                break;
            }
        }
    }
    return TT_OK;
}

int32_t tt_tree_create(tt_engine *eng, const tt_tree_desc *D, tt_tree **out) {
    if (!eng || !D || !out) return fail(TT_EINVAL, "NULL argument");
    if (D->n_nodes < 2) return fail(TT_EINVAL, "n_nodes=%d: need at least 2 nodes", D->n_nodes);
    if (!supported_states(D->n_states))
        return fail(TT_EINVAL, "n_states=%d not supported (2,3,4,5,20,21,22)", D->n_states);
    if (D->n_patterns < 1) return fail(TT_EINVAL, "n_patterns=%d", D->n_patterns);
    if (D->n_codes < 1 || D->n_codes > 256) return fail(TT_EINVAL, "n_codes=%d out of 1..256", D->n_codes);
    if (!D->parent || !D->branch_len || !D->tip_codes || !D->code_lut || !D->multiplicity || !D->pi)
        return fail(TT_EINVAL, "NULL array in tree descriptor");
    TT_HIP(hipSetDevice(eng->device));
    auto *T = new tt_tree;
    T->eng = eng;
    T->n_nodes = D->n_nodes;
    T->n = D->n_states;
    T->P = D->n_patterns;
    T->n_codes = D->n_codes;
    int32_t rc = build_schedule(T, D);
    if (rc) {
        delete T;
        return rc;
    }
    // validate codes on the host: every code must index code_lut
    const size_t ncodes_total = (size_t)T->n_tips * T->P;
    for (size_t i = 0; i < ncodes_total; ++i)
        if (D->tip_codes[i] >= D->n_codes) {
            delete T;
            return fail(TT_EINVAL, "tip code %d at %zu >= n_codes %d", D->tip_codes[i], i, D->n_codes);
        }
    T->t.assign(D->branch_len, D->branch_len + D->n_nodes);
    hipError_t he = hipStreamCreateWithFlags(&T->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete T;
        return fail(TT_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(he));
    }
    const int n = T->n, P = T->P, nn = T->n_nodes;
    std::vector<double> lutj((size_t)D->n_codes * n), loglut(lutj.size());
    std::vector<float> lutj32(lutj.size());
    for (size_t i = 0; i < lutj.size(); ++i) {
        lutj[i] = std::max((double)D->code_lut[i], TT_TINY);
        loglut[i] = std::log(lutj[i]);
        lutj32[i] = (float)lutj[i];
    }
    std::vector<int32_t> all(nn);
    for (int i = 0; i < nn; ++i) all[i] = i;
    TreeDev &d = T->d;
#define UP(dst, src, cnt)                                   \
    if ((rc = T->upload(&dst, src, (size_t)(cnt)))) {       \
        delete T;                                           \
        return rc;                                          \
    }
    int32_t *p_child_ptr, *p_child_ref, *p_child_node, *p_parent, *p_int_row, *p_tip_row, *p_int_node;
    uint8_t *p_codes;
    float *p_lut;
    double *p_lutj, *p_loglut, *p_mult, *p_pi;
    float *p_lutj32;
    UP(p_child_ptr, T->child_ptr.data(), T->child_ptr.size());
    UP(p_child_ref, T->child_ref.data(), T->child_ref.size());
    UP(p_child_node, T->child_node.data(), T->child_node.size());
    UP(p_parent, T->parent.data(), nn);
    UP(p_int_row, T->int_row.data(), nn);
    UP(p_tip_row, T->tip_row.data(), nn);
    UP(p_int_node, T->int_node.data(), T->n_int);
    UP(p_codes, D->tip_codes, ncodes_total);
    UP(p_lut, D->code_lut, (size_t)D->n_codes * n);
    UP(p_lutj, lutj.data(), lutj.size());
    UP(p_loglut, loglut.data(), loglut.size());
    UP(p_lutj32, lutj32.data(), lutj32.size());
    UP(p_mult, D->multiplicity, P);
    UP(p_pi, D->pi, n);
    UP(T->d_post_nodes, T->post_nodes.data(), T->post_nodes.size());
    UP(T->d_pre_nodes, T->pre_nodes.data(), T->pre_nodes.size());
    UP(T->d_tip_nodes, T->tip_nodes.data(), T->tip_nodes.size());
    UP(T->d_all_branches, all.data(), nn);
    UP(T->d_t, T->t.data(), nn);
    {
        std::vector<int32_t> post_off, pre_off, pre_nint;
        for (auto &L : T->post_lv) post_off.push_back(L.off);
        post_off.push_back((int32_t)T->post_nodes.size());
        for (auto &L : T->pre_lv) {
            pre_off.push_back(L.off);
            pre_nint.push_back(L.n_int);
        }
        pre_off.push_back((int32_t)T->pre_nodes.size());
        if (pre_nint.empty()) pre_nint.push_back(0);
        // node descriptors of the level lists (kPostDesc / kPreDesc ints per entry)
        std::vector<int32_t> post_desc, pre_desc;
        for (int r : T->post_nodes) {
            const int e0 = T->child_ptr[r], nc = T->child_ptr[r + 1] - e0;
            const bool bin = nc == 2;
            post_desc.insert(post_desc.end(), {r, e0, nc, bin ? T->child_ref[e0] : 0, bin ? T->child_ref[e0 + 1] : 0,
                                               bin ? T->child_node[e0] : 0, bin ? T->child_node[e0 + 1] : 0,
                                               T->int_node[r]});
        }
        for (int c : T->pre_nodes) pre_desc.insert(pre_desc.end(), {c, T->parent[c], T->int_row[c], T->tip_row[c]});
        if (post_desc.empty()) post_desc.assign(kPostDesc, 0);
        if (pre_desc.empty()) pre_desc.assign(kPreDesc, 0);
        UP(T->d_post_desc, post_desc.data(), post_desc.size());
        UP(T->d_pre_desc, pre_desc.data(), pre_desc.size());
        UP(T->d_post_off, post_off.data(), post_off.size());
        UP(T->d_pre_off, pre_off.data(), pre_off.size());
        UP(T->d_pre_nint, pre_nint.data(), pre_nint.size());
    }
#undef UP
    d.n_nodes = nn;
    d.n_tips = T->n_tips;
    d.n_int = T->n_int;
    d.n = n;
    d.P = P;
    d.n_codes = D->n_codes;
    d.root = T->root;
    d.root_row = T->root_row;
    d.child_ptr = (decltype(d.child_ptr))p_child_ptr;
    d.child_ref = (decltype(d.child_ref))p_child_ref;
    d.child_node = (decltype(d.child_node))p_child_node;
    d.parent = (decltype(d.parent))p_parent;
    d.int_row = (decltype(d.int_row))p_int_row;
    d.tip_row = (decltype(d.tip_row))p_tip_row;
    d.int_node = (decltype(d.int_node))p_int_node;
    d.tip_codes = (decltype(d.tip_codes))p_codes;
    d.lut = (decltype(d.lut))p_lut;
    d.lutj = (decltype(d.lutj))p_lutj;
    d.loglut = (decltype(d.loglut))p_loglut;
    d.lutj32 = (decltype(d.lutj32))p_lutj32;
    d.mult = (decltype(d.mult))p_mult;
    d.pi = (decltype(d.pi))p_pi;
    d.tip_cc = nullptr;
    d.lutj_c = nullptr;
    d.jtab = nullptr;
    d.jcx = nullptr;
    d.n_cc = 0;
    if (n > 8 && T->n_tips > 0) {
        // the codes the tips use, numbered in code order: the joint pass takes
        // a tip child's message from its branch's table of these codes
        std::vector<int> idx(256, -1);
        for (size_t i = 0; i < ncodes_total; ++i) idx[D->tip_codes[i]] = 0;
        int m = 0;
        std::vector<float> lutj_c;
        for (int c = 0; c < 256; ++c)
            if (idx[c] == 0) {
                idx[c] = m++;
                lutj_c.insert(lutj_c.end(), lutj32.begin() + (size_t)c * n, lutj32.begin() + (size_t)(c + 1) * n);
            }
        std::vector<uint8_t> cc(ncodes_total);
        for (size_t i = 0; i < ncodes_total; ++i) cc[i] = (uint8_t)idx[D->tip_codes[i]];
        uint8_t *p_cc;
        float *p_lutj_c;
        if ((rc = T->upload(&p_cc, cc.data(), cc.size())) || (rc = T->upload(&p_lutj_c, lutj_c.data(), lutj_c.size()))) {
            delete T;
            return rc;
        }
        d.tip_cc = (decltype(d.tip_cc))p_cc;
        d.lutj_c = (decltype(d.lutj_c))p_lutj_c;
        d.n_cc = m;
    }
    struct {
        void **p;
        size_t bytes;
    } bufs[] = {
        {(void **)(void *)&d.P64, (size_t)nn * n * n * sizeof(double)},
        {(void **)(void *)&d.Pj64, (size_t)nn * n * n * sizeof(double)},
        {(void **)(void *)&d.logP64, (size_t)nn * n * n * sizeof(double)},
        {(void **)(void *)&d.Pj32, (size_t)nn * n * n * sizeof(float)},
        {(void **)(void *)&d.P32, (size_t)nn * n * n * sizeof(float)},
        {(void **)(void *)&d.D32, (size_t)nn * n * n * sizeof(float)},
        {(void **)(void *)&d.sub, (size_t)T->n_int * P * n * sizeof(float)},
        {(void **)(void *)&d.prof, (size_t)nn * P * n * sizeof(float)},
        {(void **)(void *)&d.states, (size_t)nn * P},
        {(void **)(void *)&d.part, (size_t)std::max(T->R_all, 1) * P * sizeof(double)},
        {(void **)(void *)&d.part2, (size_t)std::max(rowsum_slices(T->R_all), 1) * P * sizeof(double)},
        {(void **)(void *)&d.seqlh, (size_t)P * sizeof(double)},
        {(void **)(void *)&d.jseqlh, (size_t)P * sizeof(double)},
        {(void **)(void *)&d.bpart, (size_t)ttk::root_blocks(P) * sizeof(double)},
        {(void **)(void *)&d.total, sizeof(double)},
        {(void **)&T->d_eig, TT_MAX_STATES * sizeof(double)},
        {(void **)&T->d_v, TT_MAX_STATES * TT_MAX_STATES * sizeof(double)},
        {(void **)&T->d_vinv, TT_MAX_STATES * TT_MAX_STATES * sizeof(double)},
    };
    for (auto &b : bufs) {
        if ((rc = T->alloc(b.p, b.bytes))) {
            delete T;
            return rc;
        }
    }
    if ((rc = setup_tile(T))) {
        delete T;
        return rc;
    }
    he = hipMemset(d.states, 255, (size_t)nn * P);
    if (he == hipSuccess) he = hipMemset(d.prof, 0, (size_t)nn * P * n * sizeof(float));
    if (he == hipSuccess && T->d_tile_counter) he = hipMemset(T->d_tile_counter, 0, sizeof(unsigned));
    // the one-launch passes' per-workgroup partials start empty (all-ones bits)
    if (he == hipSuccess && T->d_tile_partial)
        he = hipMemset(T->d_tile_partial, 0xFF, (size_t)((T->P + T->tile_tp - 1) / T->tile_tp) * sizeof(double));
    // the passes run on a non-blocking stream that does not order after the
    // null-stream memsets
    if (he == hipSuccess) he = hipDeviceSynchronize();
    if (he != hipSuccess) {
        delete T;
        return fail(TT_EDEVICE, "hipMemset: %s", hipGetErrorString(he));
    }
    eng->live_trees += 1;
    *out = T;
    return TT_OK;
}

int32_t tt_tree_destroy(tt_tree *tree) {
    if (tree) {
        (void)hipSetDevice(tree->eng->device);
        (void)hipDeviceSynchronize();
        tt_engine *eng = tree->eng;
        eng->live_trees -= 1;
        std::lock_guard<std::mutex> g(eng->mu);
        auto &pl = eng->plans;
        for (size_t i = 0; i < pl.size();) {
            if (std::find(pl[i]->trees.begin(), pl[i]->trees.end(), tree) != pl[i]->trees.end()) {
                delete pl[i];
                pl.erase(pl.begin() + i);
            } else {
                ++i;
            }
        }
    }
    delete tree;
    return TT_OK;
}

int32_t tt_tree_info(const tt_tree *T, int32_t *n_levels_post, int32_t *n_levels_pre, uint64_t *device_bytes) {
    if (!T) return fail(TT_EINVAL, "tree is NULL");
    if (n_levels_post) *n_levels_post = (int32_t)T->post_lv.size();
    if (n_levels_pre) *n_levels_pre = (int32_t)T->pre_lv.size();
    if (device_bytes) *device_bytes = T->bytes;
    return TT_OK;
}

int32_t tt_tree_pass_shape(const tt_tree *T, int32_t *tile_patterns, int32_t *records_staged,
                           int32_t *slots_reused, int32_t *joint_one_launch) {
    if (!T) return fail(TT_EINVAL, "tree is NULL");
    const bool tile = T->tile_tp && !T->site;
    if (tile_patterns) *tile_patterns = tile ? T->tile_tp : 0;
    if (records_staged) *records_staged = tile && T->tile_staged ? 1 : 0;
    if (slots_reused) *slots_reused = tile && T->tile_subg ? 1 : 0;
    if (joint_one_launch) *joint_one_launch = tile && T->tile_joint ? 1 : 0;
    return TT_OK;
}

int32_t tt_tree_set_model(tt_tree *T, const double *eigvals, const double *v, const double *v_inv, double mu) {
    if (!T || !eigvals || !v || !v_inv) return fail(TT_EINVAL, "NULL argument");
    if (!(mu >= 0.0)) return fail(TT_EINVAL, "mu must be >= 0");
    const int n = T->n;
    T->eig.assign(eigvals, eigvals + n);
    T->v.assign(v, v + n * n);
    T->vinv.assign(v_inv, v_inv + n * n);
    int32_t rc = check_range(T, T->t.data(), T->n_nodes);
    if (rc) return rc;
    T->mu = mu;
    T->offs_valid = false;
    T->profiles_fresh = false;
    if (T->site) {  // back to one model for all patterns: other kernels
        T->site = false;
        T->d.s_V32 = T->d.s_Vi32 = T->d.s_rho32 = nullptr;
        T->d.s_V64 = T->d.s_Vi64 = T->d.s_rho64 = T->d.s_pi = T->d.s_br = T->d.s_grid = nullptr;
        T->d.s_G = 0;
        destroy_graphs(T);
    }
    TT_HIP(hipSetDevice(T->eng->device));
    TT_HIP(hipMemcpyAsync(T->d_eig, eigvals, n * sizeof(double), hipMemcpyHostToDevice, T->stream));
    TT_HIP(hipMemcpyAsync(T->d_v, v, n * n * sizeof(double), hipMemcpyHostToDevice, T->stream));
    TT_HIP(hipMemcpyAsync(T->d_vinv, v_inv, n * n * sizeof(double), hipMemcpyHostToDevice, T->stream));
    if ((rc = refresh_expqt(T, T->d_all_branches, T->d_t, T->n_nodes, T->stream))) return rc;
    TT_HIP(hipStreamSynchronize(T->stream));
    T->have_model = true;
    return TT_OK;
}

namespace {
int32_t set_lengths(tt_tree *T, const int32_t *nodes, const double *t, int32_t k);
}

int32_t tt_tree_set_site_model(tt_tree *T, const double *eigvals, const double *v, const double *v_inv,
                               const double *mu, const double *pi, const double *t_grid, int32_t G,
                               double rate_scale) {
    if (!T || !eigvals || !v || !v_inv || !mu || !pi || (G > 0 && !t_grid)) return fail(TT_EINVAL, "NULL argument");
    const int n = T->n, P = T->P;
    if (n > 5) return fail(TT_EINVAL, "site-specific models are built for alphabets of up to 5 states (n = %d)", n);
    if (G < 0 || G == 1) return fail(TT_EINVAL, "G = %d: the grid needs 0 or at least 2 points", G);
    for (int i = 1; i < G; ++i)
        if (!(t_grid[i] > t_grid[i - 1])) return fail(TT_EINVAL, "t_grid must increase strictly");
    if (G > 0 && !(rate_scale > 0.0)) return fail(TT_EINVAL, "rate_scale must be > 0");
    for (int p = 0; p < P; ++p) {
        if (!(mu[p] >= 0.0)) return fail(TT_EINVAL, "mu[%d] must be >= 0", p);
        for (int j = 0; j < n; ++j)
            if (!(pi[(size_t)p * n + j] > 0.0)) return fail(TT_EINVAL, "pi[%d][%d] must be > 0", p, j);
    }
    TT_HIP(hipSetDevice(T->eng->device));
    // one allocation: V32, Vi32, rho32 (float), then V64, Vi64, rho64, pi, branch params (double)
    const size_t nn2 = (size_t)P * n * n, nn1 = (size_t)P * n;
    const size_t f_bytes = ((2 * nn2 + nn1) * sizeof(float) + 255) & ~(size_t)255;
    const size_t bytes = f_bytes + (2 * nn2 + 2 * nn1 + 3 * (size_t)T->n_nodes + 64) * sizeof(double);
    if (G > 64) return fail(TT_EINVAL, "G = %d: at most 64 grid points", G);
    int32_t rc;
    if (!T->d_site && (rc = T->alloc(&T->d_site, bytes))) return rc;
    std::vector<float> f(2 * nn2 + nn1);
    std::vector<double> dd(2 * nn2 + 2 * nn1);
    for (size_t i = 0; i < nn2; ++i) {
        f[i] = (float)v[i];
        f[nn2 + i] = (float)v_inv[i];
        dd[i] = v[i];
        dd[nn2 + i] = v_inv[i];
    }
    for (int p = 0; p < P; ++p)
        for (int k = 0; k < n; ++k) {
            const size_t i = (size_t)p * n + k;
            const double r = mu[p] * eigvals[i];
            f[2 * nn2 + i] = (float)r;
            dd[2 * nn2 + i] = r;
            dd[2 * nn2 + nn1 + i] = pi[i];
        }
    char *base = static_cast<char *>(T->d_site);
    TT_HIP(hipMemcpyAsync(base, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, T->stream));
    TT_HIP(hipMemcpyAsync(base + f_bytes, dd.data(), dd.size() * sizeof(double), hipMemcpyHostToDevice, T->stream));
    const float *fb = reinterpret_cast<const float *>(base);
    const double *db = reinterpret_cast<const double *>(base + f_bytes);
    T->d.s_V32 = (decltype(T->d.s_V32))fb;
    T->d.s_Vi32 = (decltype(T->d.s_Vi32))(fb + nn2);
    T->d.s_rho32 = (decltype(T->d.s_rho32))(fb + 2 * nn2);
    T->d.s_V64 = (decltype(T->d.s_V64))db;
    T->d.s_Vi64 = (decltype(T->d.s_Vi64))(db + nn2);
    T->d.s_rho64 = (decltype(T->d.s_rho64))(db + 2 * nn2);
    T->d.s_pi = (decltype(T->d.s_pi))(db + 2 * nn2 + nn1);
    T->d.s_br = (decltype(T->d.s_br))(db + 2 * nn2 + 2 * nn1);
    double *d_grid = const_cast<double *>(db + 2 * nn2 + 2 * nn1 + 3 * (size_t)T->n_nodes);
    if (G > 0) TT_HIP(hipMemcpyAsync(d_grid, t_grid, G * sizeof(double), hipMemcpyHostToDevice, T->stream));
    T->d.s_grid = (decltype(T->d.s_grid))d_grid;
    T->d.s_G = std::max(G, 0);
    T->d.s_rate = rate_scale;
    T->site_grid.assign(t_grid, t_grid + std::max(G, 0));
    T->site_rate_scale = rate_scale;
    if (!T->site) destroy_graphs(T);
    T->site = true;
    if ((rc = upload_site_branches(T))) return rc;
    TT_HIP(hipStreamSynchronize(T->stream));
    T->have_model = true;
    T->offs_valid = false;
    T->profiles_fresh = false;
    return TT_OK;
}

int32_t tt_tree_set_branch_lengths(tt_tree *T, const int32_t *nodes, const double *t, int32_t k) {
    int32_t rc = set_lengths(T, nodes, t, k);
    if (rc == TT_OK && k > 0) {
        T->offs_valid = false;
        T->profiles_fresh = false;
    }
    return rc;
}

namespace {
int32_t set_lengths(tt_tree *T, const int32_t *nodes, const double *t, int32_t k) {
    if (!T || (k > 0 && (!nodes || !t))) return fail(TT_EINVAL, "NULL argument");
    if (k <= 0) return TT_OK;
    for (int i = 0; i < k; ++i)
        if (nodes[i] < 0 || nodes[i] >= T->n_nodes) return fail(TT_EINVAL, "node %d out of range", nodes[i]);
    if (T->have_model && !T->site) {
        int32_t rc = check_range(T, t, k);
        if (rc) return rc;
    }
    TT_HIP(hipSetDevice(T->eng->device));
    for (int i = 0; i < k; ++i) T->t[nodes[i]] = t[i];
    TT_HIP(hipMemcpyAsync(T->d_t, T->t.data(), T->n_nodes * sizeof(double), hipMemcpyHostToDevice, T->stream));
    if (T->site) {
        int32_t rc = upload_site_branches(T);
        if (rc) return rc;
    } else if (T->have_model) {
        if (k > T->tmp_cap) {
            TT_HIP(hipStreamSynchronize(T->stream));  // earlier copies may still read the old buffers
            T->release(T->d_tmp_t);
            T->release(T->d_tmp_br);
            T->d_tmp_t = nullptr;
            T->d_tmp_br = nullptr;
            T->tmp_cap = 0;
            int32_t rc;
            if ((rc = T->alloc((void **)&T->d_tmp_t, (size_t)k * sizeof(double)))) return rc;
            if ((rc = T->alloc((void **)&T->d_tmp_br, (size_t)k * sizeof(int32_t)))) return rc;
            T->tmp_cap = k;
        }
        TT_HIP(hipMemcpyAsync(T->d_tmp_t, t, k * sizeof(double), hipMemcpyHostToDevice, T->stream));
        TT_HIP(hipMemcpyAsync(T->d_tmp_br, nodes, k * sizeof(int32_t), hipMemcpyHostToDevice, T->stream));
        int32_t rc = refresh_expqt(T, T->d_tmp_br, T->d_tmp_t, k, T->stream);
        if (rc) return rc;
    }
    TT_HIP(hipStreamSynchronize(T->stream));
    return TT_OK;
}
}  // namespace

int32_t tt_marginal(tt_tree *tree, uint32_t flags, double *logL, void *stream) {
    if (tree && (flags & TT_KEEP_NORMALISERS)) {
        if (tree->site) return fail(TT_EINVAL, "TT_KEEP_NORMALISERS: not available for site-specific models");
        TT_HIP(hipSetDevice(tree->eng->device));
        if (int32_t rc = enable_update_mode(tree)) return rc;
    }
    return run_pass(tree, 1, flags & TT_RECONSTRUCT_TIPS, logL, stream);
}

int32_t tt_loglik(tt_tree *tree, double *logL, void *stream) { return run_pass(tree, 3, 0, logL, stream); }

int32_t tt_joint(tt_tree *tree, uint32_t flags, double *logL, void *stream) {
    return run_pass(tree, 2, flags, logL, stream);
}

int32_t tt_update_branch(tt_tree *T, int32_t node, double t, double *logL, void *stream) {
    if (!T) return fail(TT_EINVAL, "tree is NULL");
    if (node < 0 || node >= T->n_nodes) return fail(TT_EINVAL, "node %d out of range", node);
    if (node == T->root) return fail(TT_EINVAL, "the root has no branch");
    if (!T->have_model) return fail(TT_ESTATE, "tt_tree_set_model has not been called");
    if (int32_t r = refuse_site(T, "tt_update_branch")) return r;
    TT_HIP(hipSetDevice(T->eng->device));
    int32_t rc = enable_update_mode(T);
    if (rc) return rc;
    if (!T->offs_valid) {  // first call: set the length, full pass that keeps normalisers
        if ((rc = set_lengths(T, &node, &t, 1))) return rc;
        return run_pass(T, 3, 0, logL, stream);
    }
    if ((rc = check_range(T, &t, 1))) return rc;
    // asynchronous on the stream: new length, its P(t), the root path, the total
    hipStream_t s = stream ? (hipStream_t)stream : T->stream;
    T->t[node] = t;
    TT_HIP(hipMemcpyAsync(T->d_t + node, &T->t[node], sizeof(double), hipMemcpyHostToDevice, s));
    TT_HIP(ttk::expqt(T->n, T->d_t + node, T->d_all_branches + node, 1, T->d_eig, T->d_v, T->d_vinv, T->mu,
                      (double *)T->d.P64, (double *)T->d.Pj64, (double *)T->d.logP64, (float *)T->d.Pj32, (float *)T->d.P32,
                      (float *)T->d.D32, s));
    TT_HIP(ttk::path_update(T->d, node, s));
    TT_HIP(ttk::final_sum(T->d, ttk::root_blocks(T->P), s));
    T->tile_stale = true;  // P(t) of the branch changed
    T->last_pass = 3;
    T->last_stream = s;
    T->profiles_fresh = false;
    if (logL) {
        TT_HIP(hipMemcpyAsync(logL, T->d.total, sizeof(double), hipMemcpyDeviceToHost, s));
        TT_HIP(hipStreamSynchronize(s));
    }
    return TT_OK;
}

int32_t tt_optimize_branch_lengths(tt_tree *T, const int16_t *code_state, int32_t gap_state, double max_t,
                                   double tol, double *t_opt, int32_t *status, double *pair_counts, void *stream) {
    if (!T || !code_state || !t_opt || !status) return fail(TT_EINVAL, "NULL argument");
    if (!T->have_model) return fail(TT_ESTATE, "tt_tree_set_model has not been called");
    if (int32_t r = refuse_site(T, "tt_optimize_branch_lengths")) return r;
    if (!T->have_states) return fail(TT_ESTATE, "no marginal or joint reconstruction has run on this tree");
    if (gap_state < -1 || gap_state >= T->n) return fail(TT_EINVAL, "gap_state %d outside [-1, %d)", gap_state, T->n);
    if (!(max_t > 0.0) || !(tol >= 0.0)) return fail(TT_EINVAL, "need max_t > 0 and tol >= 0");
    for (int c = 0; c < 256; ++c)
        if (code_state[c] < -1 || code_state[c] >= T->n)
            return fail(TT_EINVAL, "code_state[%d] = %d outside [-1, %d)", c, code_state[c], T->n);
    TT_HIP(hipSetDevice(T->eng->device));
    int32_t rc;
    const size_t nn = (size_t)T->n * T->n;
    if ((rc = ensure_opt_out(T))) return rc;
    if (!T->d_code_state) {
        if ((rc = T->alloc((void **)&T->d_code_state, 256 * sizeof(int16_t)))) return rc;
        if ((rc = T->alloc((void **)&T->d_pair_counts, (size_t)T->n_nodes * nn * sizeof(double)))) return rc;
    }
    hipStream_t s = stream ? (hipStream_t)stream : T->stream;
    TT_HIP(hipMemcpyAsync(T->d_code_state, code_state, 256 * sizeof(int16_t), hipMemcpyHostToDevice, s));
    TT_HIP(ttk::branch_opt(T->d, T->d_code_state, gap_state, T->d_eig, T->d_v, T->d_vinv, T->mu, max_t, tol,
                           T->d_pair_counts, T->d_t_opt, T->d_opt_status, 1, s));
    TT_HIP(hipMemcpyAsync(t_opt, T->d_t_opt, (size_t)T->n_nodes * sizeof(double), hipMemcpyDeviceToHost, s));
    TT_HIP(hipMemcpyAsync(status, T->d_opt_status, (size_t)T->n_nodes * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (pair_counts)
        TT_HIP(hipMemcpyAsync(pair_counts, T->d_pair_counts, (size_t)T->n_nodes * nn * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    TT_HIP(hipStreamSynchronize(s));
    return TT_OK;
}

int32_t tt_optimize_branch_lengths_marginal(tt_tree *T, int32_t gap_state, double max_t, double tol, double *t_opt,
                                            int32_t *status, double *root_pair, void *stream) {
    if (!T || !t_opt || !status) return fail(TT_EINVAL, "NULL argument");
    if (!T->have_model) return fail(TT_ESTATE, "tt_tree_set_model has not been called");
    if (!T->profiles_fresh)
        return fail(TT_ESTATE, "marginal ancestral inference needs to be performed first (profiles are stale)");
    if (gap_state < -1 || gap_state >= T->n) return fail(TT_EINVAL, "gap_state %d outside [-1, %d)", gap_state, T->n);
    if (!(max_t > 0.0) || !(tol >= 0.0)) return fail(TT_EINVAL, "need max_t > 0 and tol >= 0");
    TT_HIP(hipSetDevice(T->eng->device));
    int32_t rc;
    if ((rc = ensure_opt_out(T))) return rc;
    // children of a bifurcating root, in child order (treeanc.py:1318-1320)
    int n1 = -1, n2 = -1;
    const int rr = T->int_row[T->root];
    if (rr >= 0 && T->child_ptr[rr + 1] - T->child_ptr[rr] == 2) {
        n1 = T->child_node[T->child_ptr[rr]];
        n2 = T->child_node[T->child_ptr[rr] + 1];
    }
    hipStream_t s = stream ? (hipStream_t)stream : T->stream;
    TT_HIP(ttk::branch_opt_marginal(T->d, gap_state, T->d_eig, T->d_v, T->d_vinv, T->mu, max_t, tol, n1, n2,
                                    T->d_t_opt, T->d_opt_status, s));
    TT_HIP(hipMemcpyAsync(t_opt, T->d_t_opt, (size_t)T->n_nodes * sizeof(double), hipMemcpyDeviceToHost, s));
    TT_HIP(hipMemcpyAsync(status, T->d_opt_status, (size_t)T->n_nodes * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    double pair_t = NAN;
    int32_t pair_st = -1;
    if (n1 >= 0) {
        TT_HIP(hipMemcpyAsync(&pair_t, T->d_t_opt + T->n_nodes, sizeof(double), hipMemcpyDeviceToHost, s));
        TT_HIP(hipMemcpyAsync(&pair_st, T->d_opt_status + T->n_nodes, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    }
    TT_HIP(hipStreamSynchronize(s));
    if (root_pair) {
        root_pair[0] = pair_t;
        root_pair[1] = (double)pair_st;
    }
    return TT_OK;
}

int32_t tt_branch_profile_lh(tt_tree *T, int32_t gap_state, const double *t, int32_t K, double *out, void *stream) {
    if (!T || (K > 0 && (!t || !out))) return fail(TT_EINVAL, "NULL argument");
    if (K <= 0) return TT_OK;
    if (!T->have_model) return fail(TT_ESTATE, "tt_tree_set_model has not been called");
    if (!T->profiles_fresh)
        return fail(TT_ESTATE, "marginal ancestral inference needs to be performed first (profiles are stale)");
    if (gap_state < -1 || gap_state >= T->n) return fail(TT_EINVAL, "gap_state %d outside [-1, %d)", gap_state, T->n);
    TT_HIP(hipSetDevice(T->eng->device));
    int32_t rc;
    if (K > T->grid_cap) {
        TT_HIP(hipStreamSynchronize(T->stream));
        T->release(T->d_grid_t);
        T->release(T->d_grid_out);
        T->d_grid_t = T->d_grid_out = nullptr;
        T->grid_cap = 0;
        if ((rc = T->alloc((void **)&T->d_grid_t, (size_t)K * sizeof(double)))) return rc;
        if ((rc = T->alloc((void **)&T->d_grid_out, (size_t)K * T->n_nodes * sizeof(double)))) return rc;
        T->grid_cap = K;
    }
    hipStream_t s = stream ? (hipStream_t)stream : T->stream;
    TT_HIP(hipMemcpyAsync(T->d_grid_t, t, (size_t)K * sizeof(double), hipMemcpyHostToDevice, s));
    TT_HIP(ttk::branch_profile_lh(T->d, gap_state, T->d_eig, T->d_v, T->d_vinv, T->mu, T->d_grid_t, K,
                                  T->d_grid_out, s));
    TT_HIP(hipMemcpyAsync(out, T->d_grid_out, (size_t)K * T->n_nodes * sizeof(double), hipMemcpyDeviceToHost, s));
    TT_HIP(hipStreamSynchronize(s));
    return TT_OK;
}

int32_t tt_branch_state_counts(tt_tree *T, const int16_t *code_state, int32_t gap_state, double *pair_counts,
                               void *stream) {
    if (!T || !code_state || !pair_counts) return fail(TT_EINVAL, "NULL argument");
    if (!T->have_states) return fail(TT_ESTATE, "no marginal or joint reconstruction has run on this tree");
    if (gap_state < -1 || gap_state >= T->n) return fail(TT_EINVAL, "gap_state %d outside [-1, %d)", gap_state, T->n);
    for (int c = 0; c < 256; ++c)
        if (code_state[c] < -1 || code_state[c] >= T->n)
            return fail(TT_EINVAL, "code_state[%d] = %d outside [-1, %d)", c, code_state[c], T->n);
    TT_HIP(hipSetDevice(T->eng->device));
    int32_t rc;
    const size_t nn = (size_t)T->n * T->n;
    if ((rc = ensure_opt_out(T))) return rc;
    if (!T->d_code_state) {
        if ((rc = T->alloc((void **)&T->d_code_state, 256 * sizeof(int16_t)))) return rc;
        if ((rc = T->alloc((void **)&T->d_pair_counts, (size_t)T->n_nodes * nn * sizeof(double)))) return rc;
    }
    hipStream_t s = stream ? (hipStream_t)stream : T->stream;
    TT_HIP(hipMemcpyAsync(T->d_code_state, code_state, 256 * sizeof(int16_t), hipMemcpyHostToDevice, s));
    TT_HIP(ttk::branch_opt(T->d, T->d_code_state, gap_state, T->d_eig, T->d_v, T->d_vinv, T->mu, 1.0, 0.0,
                           T->d_pair_counts, T->d_t_opt, T->d_opt_status, 0, s));
    TT_HIP(hipMemcpyAsync(pair_counts, T->d_pair_counts, (size_t)T->n_nodes * nn * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    TT_HIP(hipStreamSynchronize(s));
    return TT_OK;
}

int32_t tt_mutation_counts(tt_tree *T, double *nij, double *Ti, double *per_branch, void *stream) {
    if (!T || !nij || !Ti) return fail(TT_EINVAL, "NULL argument");
    if (int32_t r = refuse_site(T, "tt_mutation_counts")) return r;
    if (!T->profiles_fresh)
        return fail(TT_ESTATE, "marginal ancestral inference needs to be performed first (profiles are stale)");
    TT_HIP(hipSetDevice(T->eng->device));
    const int n = T->n, w = n * n + n;
    if (!T->d_mut_counts) {
        int32_t rc = T->alloc((void **)&T->d_mut_counts, (size_t)T->n_nodes * w * sizeof(double));
        if (rc) return rc;
    }
    hipStream_t s = stream ? (hipStream_t)stream : T->stream;
    TT_HIP(ttk::mutation_counts(T->d, T->d_t, T->d_mut_counts, s));
    std::vector<double> h((size_t)T->n_nodes * w);
    TT_HIP(hipMemcpyAsync(h.data(), T->d_mut_counts, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    TT_HIP(hipStreamSynchronize(s));
    // sums over branches in node order
    for (int k = 0; k < n * n; ++k) nij[k] = 0.0;
    for (int k = 0; k < n; ++k) Ti[k] = 0.0;
    for (int c = 0; c < T->n_nodes; ++c) {
        const double *o = h.data() + (size_t)c * w;
        for (int k = 0; k < n * n; ++k) nij[k] += o[k];
        for (int k = 0; k < n; ++k) Ti[k] += o[n * n + k];
    }
    if (per_branch) std::memcpy(per_branch, h.data(), h.size() * sizeof(double));
    return TT_OK;
}

int32_t tt_site_counts(tt_tree *T, int32_t marginal, const int16_t *code_state, double *n_out, double *t_out,
                       void *stream) {
    if (!T || !n_out || !t_out || (!marginal && !code_state)) return fail(TT_EINVAL, "NULL argument");
    if (T->n > 5) return fail(TT_EINVAL, "site statistics are built for alphabets of up to 5 states (n = %d)", T->n);
    if (marginal && !T->profiles_fresh)
        return fail(TT_ESTATE, "marginal ancestral inference needs to be performed first (profiles are stale)");
    if (!marginal && !T->have_states)
        return fail(TT_ESTATE, "no marginal or joint reconstruction has run on this tree");
    TT_HIP(hipSetDevice(T->eng->device));
    hipStream_t s = stream ? (hipStream_t)stream : T->stream;
    int32_t rc;
    if (!marginal) {
        if (!T->d_code_state && (rc = T->alloc((void **)&T->d_code_state, 256 * sizeof(int16_t)))) return rc;
        TT_HIP(hipMemcpyAsync(T->d_code_state, code_state, 256 * sizeof(int16_t), hipMemcpyHostToDevice, s));
    }
    const size_t nn = (size_t)T->P * T->n * T->n, n1 = (size_t)T->P * T->n;
    double *d_out = nullptr;
    if ((rc = T->alloc((void **)&d_out, (nn + n1) * sizeof(double)))) return rc;
    hipError_t e = ttk::site_counts(T->d, T->d_t, marginal ? 0 : 1, T->d_code_state, d_out, d_out + nn, s);
    if (e == hipSuccess) e = hipMemcpyAsync(n_out, d_out, nn * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(t_out, d_out + nn, n1 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    T->release(d_out);
    if (e != hipSuccess) return fail(TT_EDEVICE, "tt_site_counts: %s", hipGetErrorString(e));
    return TT_OK;
}

int32_t tt_tree_trim(tt_tree *T) {
    if (!T) return fail(TT_EINVAL, "tree is NULL");
    if (!T->joint_ready) return TT_OK;
    TT_HIP(hipSetDevice(T->eng->device));
    TT_HIP(hipStreamSynchronize(T->stream));
    TT_HIP(hipDeviceSynchronize());
    {  // joint batch plans hold the freed pointers too
        std::lock_guard<std::mutex> g(T->eng->mu);
        auto &pl = T->eng->plans;
        for (size_t i = 0; i < pl.size();) {
            if ((pl[i]->flags & kPlanJoint) && std::find(pl[i]->trees.begin(), pl[i]->trees.end(), T) != pl[i]->trees.end()) {
                delete pl[i];
                pl.erase(pl.begin() + i);
            } else {
                ++i;
            }
        }
    }
    for (auto it = T->graphs.begin(); it != T->graphs.end();) {  // joint graphs hold the freed pointers
        if (it->first / 8 == 2) {
            (void)hipGraphExecDestroy(it->second);
            it = T->graphs.erase(it);
        } else {
            ++it;
        }
    }
    for (void *q : {(void *)T->d.Lx, (void *)T->d.Lx64, (void *)T->d.Cx, (void *)T->d.jtab, (void *)T->d.jcx}) {
        if (!q) continue;
        for (size_t i = 0; i < T->allocs.size(); ++i)
            if (T->allocs[i].first == q) {
                (void)hipFree(q);
                T->eng->used -= T->allocs[i].second;
                T->bytes -= T->allocs[i].second;
                T->allocs.erase(T->allocs.begin() + i);
                break;
            }
    }
    T->d.Lx = nullptr;
    T->d.Lx64 = nullptr;
    T->d.Cx = nullptr;
    T->d.jtab = nullptr;
    T->d.jcx = nullptr;
    T->joint_ready = false;
    if (T->last_pass == 2) T->last_pass = 0;
    return TT_OK;
}

int32_t tt_tree_query(tt_tree *T) {
    if (!T) return fail(TT_EINVAL, "tree is NULL");
    hipError_t e = hipStreamQuery(T->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(TT_EDEVICE, "hipStreamQuery: %s", hipGetErrorString(e));
}

int32_t tt_get_total(tt_tree *T, double *logL) {
    if (!T || !logL) return fail(TT_EINVAL, "NULL argument");
    if (!T->last_pass) return fail(TT_ESTATE, "no pass has run on this tree");
    if (int32_t rc = wait_results(T)) return rc;
    TT_HIP(hipMemcpy(logL, T->d.total, sizeof(double), hipMemcpyDeviceToHost));
    return TT_OK;
}

int32_t tt_get_sequence_lh(tt_tree *T, double *out) {
    if (!T || !out) return fail(TT_EINVAL, "NULL argument");
    if (!T->last_pass) return fail(TT_ESTATE, "no pass has run on this tree");
    if (int32_t rc = wait_results(T)) return rc;
    TT_HIP(hipMemcpy(out, T->last_pass == 2 ? T->d.jseqlh : T->d.seqlh, (size_t)T->P * sizeof(double),
                     hipMemcpyDeviceToHost));
    return TT_OK;
}

int32_t tt_get_states(tt_tree *T, uint8_t *out) {
    if (!T || !out) return fail(TT_EINVAL, "NULL argument");
    if (!T->last_pass) return fail(TT_ESTATE, "no pass has run on this tree");
    if (int32_t rc = wait_results(T)) return rc;
    TT_HIP(hipMemcpy(out, T->d.states, (size_t)T->n_nodes * T->P, hipMemcpyDeviceToHost));
    return TT_OK;
}

int32_t tt_get_profiles(tt_tree *T, int32_t which, float *out) {
    if (!T || !out) return fail(TT_EINVAL, "NULL argument");
    if (!T->last_pass) return fail(TT_ESTATE, "no pass has run on this tree");
    if (which < TT_PROFILE_MARGINAL || which > TT_PROFILE_LOG_MSG)
        return fail(TT_EINVAL, "unknown profile kind %d", which);
    if (which != TT_PROFILE_MARGINAL && !T->profiles_fresh)
        return fail(TT_ESTATE, "marginal ancestral inference needs to be performed first (profiles are stale)");
    if (int32_t rc = wait_results(T)) return rc;
    const size_t bytes = (size_t)T->n_nodes * T->P * T->n * sizeof(float);
    if (which == TT_PROFILE_MARGINAL) {
        TT_HIP(hipMemcpy(out, T->d.prof, bytes, hipMemcpyDeviceToHost));
        return TT_OK;
    }
    // derived rows are formed on the device into a staging buffer
    hipStream_t s = T->stream;
    float *d_rows = nullptr;
    TT_HIP(hipMallocAsync((void **)&d_rows, std::max<size_t>(bytes, 16), s));
    hipError_t e = ttk::node_rows(T->d, which, d_rows, s);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_rows, bytes, hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipFreeAsync(d_rows, s);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(TT_EDEVICE, "tt_get_profiles: %s", hipGetErrorString(e));
    return TT_OK;
}

int32_t tt_branch_mutation_matrix(tt_tree *T, int32_t node, double *out) {
    if (!T || !out) return fail(TT_EINVAL, "NULL argument");
    if (node < 0 || node >= T->n_nodes) return fail(TT_EINVAL, "node %d out of range", node);
    if (T->parent[node] < 0) return fail(TT_EINVAL, "Branch profiles can't be calculated for the root!");
    if (!T->profiles_fresh)
        return fail(TT_ESTATE, "marginal ancestral inference needs to be performed first!");
    if (int32_t rc = wait_results(T)) return rc;
    hipStream_t s = T->stream;
    const size_t bytes = (size_t)T->P * T->n * T->n * sizeof(double);
    double *d_m = nullptr;
    TT_HIP(hipMallocAsync((void **)&d_m, std::max<size_t>(bytes, 16), s));
    hipError_t e = ttk::branch_mutation_matrix(T->d, node, d_m, s);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_m, bytes, hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipFreeAsync(d_m, s);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(TT_EDEVICE, "tt_branch_mutation_matrix: %s", hipGetErrorString(e));
    return TT_OK;
}

int32_t tt_get_joint_root(tt_tree *T, double *out) {
    if (!T || !out) return fail(TT_EINVAL, "NULL argument");
    if (T->last_pass != 2 || !T->joint_ready) return fail(TT_ESTATE, "the last pass on this tree is not a joint pass");
    if (int32_t rc = wait_results(T)) return rc;
    const int n = T->n, P = T->P;
    const size_t cnt = (size_t)P * n, off = (size_t)T->root_row * cnt;
    if (T->site) {  // log message: subtract the per-pattern maximum
        TT_HIP(hipMemcpy(out, T->d.Lx64 + off, cnt * sizeof(double), hipMemcpyDeviceToHost));
        for (int p = 0; p < P; ++p) {
            double *r = out + (size_t)p * n, mx = r[0];
            for (int j = 1; j < n; ++j) mx = std::max(mx, r[j]);
            for (int j = 0; j < n; ++j) r[j] -= mx;
        }
    } else {  // product space, largest entry 1
        std::vector<float> h(cnt);
        TT_HIP(hipMemcpy(h.data(), T->d.Lx + off, cnt * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < cnt; ++i) out[i] = std::log((double)h[i]);
    }
    return TT_OK;
}

int32_t tt_get_joint_debug(tt_tree *T, double *lx, uint8_t *cx) {
    if (!T || !lx) return fail(TT_EINVAL, "NULL argument");
    if (T->last_pass != 2 || !T->joint_ready || T->joint_site != T->site)
        return fail(TT_ESTATE, "the last pass on this tree is not a joint pass");
    if (int32_t rc = wait_results(T)) return rc;
    hipStream_t s = T->stream;
    const size_t rows = (size_t)T->n_nodes * T->P, nrow = rows * T->n;
    double *d_rel = nullptr;
    TT_HIP(hipMallocAsync((void **)&d_rel, (rows + nrow) * sizeof(double), s));
    double *d_own = d_rel + nrow;
    std::vector<double> own(rows);
    hipError_t e = T->site ? ttk::joint_rows_site(T->d, d_rel, d_own, s) : ttk::joint_rows(T->d, d_rel, d_own, s);
    if (e == hipSuccess) e = hipMemcpyAsync(lx, d_rel, nrow * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(own.data(), d_own, rows * sizeof(double), hipMemcpyDeviceToHost, s);
    // argmax rows: packed rows are unpacked below, in place from the end
    if (e == hipSuccess && cx) e = hipMemcpyAsync(cx, T->d.Cx, rows * cx_row_bytes(T->n), hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipFreeAsync(d_rel, s);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(TT_EDEVICE, "tt_get_joint_debug: %s", hipGetErrorString(e));
    if (cx && cx_packed(T->n)) {
        const int nn = T->n;
        for (size_t r = rows; r-- > 0;) {
            const uint8_t b = cx[r];
            for (int j = nn - 1; j >= 0; --j) cx[r * nn + j] = (uint8_t)((b >> (2 * j)) & 3);
        }
    }
    // absolute log messages: add, per internal node, the sum of the log
    // maxima over its subtree's internal nodes (children before parents)
    const int P = T->P, n = T->n;
    for (int idx : T->post_nodes) {
        const int v = T->int_node[idx];
        double *ov = own.data() + (size_t)v * P;
        for (int e0 = T->child_ptr[idx]; e0 < T->child_ptr[idx + 1]; ++e0) {
            if (T->child_ref[e0] < 0) continue;
            const double *oc = own.data() + (size_t)T->child_node[e0] * P;
            for (int p = 0; p < P; ++p) ov[p] += oc[p];
        }
        double *row = lx + (size_t)v * P * n;
        for (int p = 0; p < P; ++p)
            for (int j = 0; j < n; ++j) row[(size_t)p * n + j] += ov[p];
    }
    return TT_OK;
}

int32_t tt_get_subtree_prefactor(tt_tree *T, double *out) {
    if (!T || !out) return fail(TT_EINVAL, "NULL argument");
    if (!T->update_mode || !T->offs_valid || !T->profiles_fresh)
        return fail(TT_ESTATE, "the last marginal pass did not keep its node normalisers (TT_KEEP_NORMALISERS)");
    if (int32_t rc = wait_results(T)) return rc;
    const int P = T->P, n = T->n;
    std::vector<double> offs((size_t)T->n_int * P);
    std::vector<float> sub((size_t)T->n_int * P * n);
    TT_HIP(hipMemcpy(offs.data(), T->d.offs, offs.size() * sizeof(double), hipMemcpyDeviceToHost));
    TT_HIP(hipMemcpy(sub.data(), T->d.sub, sub.size() * sizeof(float), hipMemcpyDeviceToHost));
    // prefactor = sum of the node normalisers over the subtree's internal
    // nodes (children before parents) + log of the row sum of the stored row
    std::memset(out, 0, (size_t)T->n_nodes * P * sizeof(double));
    for (int idx : T->post_nodes) {
        const int v = T->int_node[idx];
        double *o = out + (size_t)v * P;
        const double *own = offs.data() + (size_t)idx * P;
        for (int p = 0; p < P; ++p) o[p] = own[p];
        for (int e0 = T->child_ptr[idx]; e0 < T->child_ptr[idx + 1]; ++e0) {
            const int cr = T->child_ref[e0];
            if (cr < 0) continue;
            const double *oc = offs.data() + (size_t)cr * P;  // child's own subtree sum, kept in offs below
            for (int p = 0; p < P; ++p) o[p] += oc[p];
        }
        // keep the subtree sum for the parent
        std::memcpy(offs.data() + (size_t)idx * P, o, (size_t)P * sizeof(double));
    }
    for (int idx = 0; idx < T->n_int; ++idx) {
        double *o = out + (size_t)T->int_node[idx] * P;
        const float *r = sub.data() + (size_t)idx * P * n;
        for (int p = 0; p < P; ++p) {
            double z = 0.0;
            for (int j = 0; j < n; ++j) z += r[(size_t)p * n + j];
            o[p] += std::log(z);
        }
    }
    return TT_OK;
}

int32_t tt_joint_from_root(tt_tree *T, uint32_t flags, const uint8_t *root_states, double *logL, void *stream) {
    if (!T || !root_states) return fail(TT_EINVAL, "NULL argument");
    if (T->last_pass != 2 || !T->joint_ready || T->joint_site != T->site)
        return fail(TT_ESTATE, "tt_joint_from_root needs the messages of a joint pass (the last pass on this tree)");
    if (((flags & TT_RECONSTRUCT_TIPS) != 0) != T->tips_dirty)
        return fail(TT_EINVAL, "reconstruct-tips flag differs from the last joint pass");
    for (int p = 0; p < T->P; ++p)
        if (root_states[p] >= T->n) return fail(TT_EINVAL, "root state %d at pattern %d >= n %d", root_states[p], p, T->n);
    TT_HIP(hipSetDevice(T->eng->device));
    hipStream_t s = stream ? (hipStream_t)stream : T->stream;
    if (T->last_stream && T->last_stream != s) TT_HIP(hipStreamSynchronize(T->last_stream));
    TT_HIP(hipMemcpyAsync(T->d.states + (size_t)T->root * T->P, root_states, T->P, hipMemcpyHostToDevice, s));
    TT_HIP(hipStreamSynchronize(s));  // the caller's array may go away after the call
    return run_pass(T, 4, flags & TT_RECONSTRUCT_TIPS, logL, stream);
}

int32_t tt_get_device_buffers(tt_tree *T, float **profiles, uint8_t **states, double **sequence_lh) {
    if (!T) return fail(TT_EINVAL, "tree is NULL");
    if (profiles) *profiles = T->d.prof;
    if (states) *states = T->d.states;
    if (sequence_lh) *sequence_lh = T->d.seqlh;
    return TT_OK;
}

int32_t tt_propagate_profile(const double *in_dev, int32_t P, int32_t n, const double *Q, int32_t transpose,
                             int32_t return_log, double *out_dev, void *stream) {
    if (!in_dev || !out_dev || !Q) return fail(TT_EINVAL, "NULL argument");
    if (P < 0 || n < 1 || n > TT_MAX_STATES) return fail(TT_EINVAL, "bad shape P=%d n=%d", P, n);
    if (P == 0) return TT_OK;
    hipStream_t s = (hipStream_t)stream;
    double *dQ = nullptr;
    TT_HIP(hipMallocAsync((void **)&dQ, (size_t)n * n * sizeof(double), s));
    TT_HIP(hipMemcpyAsync(dQ, Q, (size_t)n * n * sizeof(double), hipMemcpyHostToDevice, s));
    TT_HIP(ttk::propagate(in_dev, P, n, dQ, transpose, return_log, out_dev, s));
    TT_HIP(hipFreeAsync(dQ, s));
    TT_HIP(hipStreamSynchronize(s));
    return TT_OK;
}

int32_t tt_propagate_profile_site(const double *in_dev, int32_t P, int32_t n, const double *Q_dev,
                                  int32_t transpose, int32_t return_log, double *out_dev, void *stream) {
    if (!in_dev || !out_dev || !Q_dev) return fail(TT_EINVAL, "NULL argument");
    if (P < 0 || n < 1 || n > TT_MAX_STATES) return fail(TT_EINVAL, "bad shape P=%d n=%d", P, n);
    if (P == 0) return TT_OK;
    hipStream_t s = (hipStream_t)stream;
    TT_HIP(ttk::propagate_site(in_dev, P, n, Q_dev, transpose, return_log, out_dev, s));
    TT_HIP(hipStreamSynchronize(s));
    return TT_OK;
}

int32_t tt_prob_t_profiles(const double *pp_dev, const double *pc_dev, const double *mult_dev, int32_t P, int32_t n,
                           const double *Q, int32_t K, int32_t gap_index, double *out, void *stream) {
    if (!pp_dev || !pc_dev || !mult_dev || !Q || !out) return fail(TT_EINVAL, "NULL argument");
    if (P < 1 || n < 1 || n > TT_MAX_STATES || K < 1) return fail(TT_EINVAL, "bad shape P=%d n=%d K=%d", P, n, K);
    if (gap_index >= n) return fail(TT_EINVAL, "gap_index %d >= n %d", gap_index, n);
    hipStream_t s = (hipStream_t)stream;
    const int nblk = ttk::root_blocks(P);
    double *dQ = nullptr, *dB = nullptr;
    TT_HIP(hipMallocAsync((void **)&dQ, (size_t)K * n * n * sizeof(double), s));
    TT_HIP(hipMallocAsync((void **)&dB, (size_t)K * nblk * sizeof(double), s));
    TT_HIP(hipMemcpyAsync(dQ, Q, (size_t)K * n * n * sizeof(double), hipMemcpyHostToDevice, s));
    TT_HIP(ttk::prob_t_profiles(pp_dev, pc_dev, mult_dev, P, n, dQ, K, gap_index, dB, nblk, s));
    std::vector<double> hb((size_t)K * nblk);
    TT_HIP(hipMemcpyAsync(hb.data(), dB, hb.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    TT_HIP(hipFreeAsync(dQ, s));
    TT_HIP(hipFreeAsync(dB, s));
    TT_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < K; ++k) {
        double acc = 0.0;
        for (int b = 0; b < nblk; ++b) acc += hb[(size_t)k * nblk + b];
        out[k] = acc;
    }
    return TT_OK;
}

int32_t tt_batch_info(tt_engine *eng, tt_tree *const *trees, int32_t k, uint32_t flags, int32_t *wave_items,
                      int32_t *tiles, int32_t *blocks_per_cu, int32_t *waves_per_simd, double *tuned_ms) {
    if (!eng || (k > 0 && !trees)) return fail(TT_EINVAL, "NULL argument");
    std::vector<tt_tree *> key(trees, trees + std::max(k, 0));
    const uint32_t kind = flags & 0x30u;
    const uint32_t want = kind == TT_BATCH_LOGLIK ? kPlanLoglik
                          : kind == TT_BATCH_JOINT ? (kPlanJoint | (flags & TT_RECONSTRUCT_TIPS))
                                                   : (flags & TT_RECONSTRUCT_TIPS);
    std::lock_guard<std::mutex> g(eng->mu);
    for (BatchPlan *p : eng->plans)
        if (p->flags == want && p->trees == key) {
            if (wave_items) *wave_items = p->shape.wave_items;
            if (tiles) *tiles = p->shape.tiles;
            if (blocks_per_cu) *blocks_per_cu = p->shape.blocks_per_cu;
            if (waves_per_simd) *waves_per_simd = p->shape.waves_per_simd;
            if (tuned_ms) *tuned_ms = p->tuned_ms;
            return TT_OK;
        }
    return fail(TT_ESTATE, "no batch plan for these trees and flags");
}

}  // extern "C"

namespace {

// kind: 0 marginal, 1 loglik (postorder + root), 2 joint
// A batched pass over the members of a plan's trees: the members' items
// (per die, in the plan's order) go through the plan's tuned launch shape,
// followed by the members' totals.  The run waits for work already queued
// on each member's own stream, and each member's stream waits for the run,
// so per-tree calls before and after it stay ordered.
int32_t plan_run_members(BatchPlan *plan, tt_tree *const *trees, int k, int n, int mode, uint32_t fl, int pass,
                         const int32_t *members, int m, double *logL, hipStream_t s) {
    std::vector<char> is(k, 0);
    for (int j = 0; j < m; ++j) {
        if (members[j] < 0 || members[j] >= k)
            return fail(TT_EINVAL, "members[%d]=%d out of range [0, %d)", j, members[j], k);
        if (is[members[j]]) return fail(TT_EINVAL, "members[%d]=%d listed twice", j, members[j]);
        is[members[j]] = 1;
    }
    if (m == 0) return TT_OK;
    const BatchShape &c = plan->shape;
    if (c.wave_items == kPerTreeGraphs) {
        for (int j = 0; j < m; ++j)
            if (int32_t rc = run_pass(trees[members[j]], pass, fl ? TT_RECONSTRUCT_TIPS : 0u,
                                      logL ? &logL[j] : nullptr, nullptr))
                return rc;
        return TT_OK;
    }
    const int ci = c.tiles == 2 ? 1 : 0;
    const std::vector<BatchItem> &all = plan->items_for(c);
    const size_t ni_max = std::max(plan->items1.size(), plan->items2.size());
    // device block: [queue counters | xoff (16) | items | totals (3 per tree) | item outputs]
    const size_t off_items = (size_t)(kQueueXoff1 + 16) * sizeof(int32_t);
    const size_t off_tot = off_items + ni_max * sizeof(BatchItem);
    const size_t off_out = (off_tot + (size_t)3 * k * sizeof(int32_t) + 255) & ~(size_t)255;
    const size_t dev_bytes = off_out + ni_max * sizeof(double);
    if (!plan->d_sub_queue) {
        tt_engine *eng = plan->eng;
        if (eng->budget && eng->used.load() + dev_bytes > eng->budget)
            return fail(TT_ENOMEM_BUDGET, "HBM budget exceeded by the member-run buffers of a batch plan");
        char *blk = nullptr;
        TT_HIP(hipMalloc((void **)&blk, dev_bytes));
        eng->used += dev_bytes;
        plan->bytes += dev_bytes;
        plan->d_sub_queue = reinterpret_cast<int *>(blk);
        plan->d_sub_items = reinterpret_cast<BatchItem *>(blk + off_items);
        plan->d_sub_tot = reinterpret_cast<int32_t *>(blk + off_tot);
        plan->d_sub_out = reinterpret_cast<double *>(blk + off_out);
        TT_HIP(hipMemset(blk, 0, (size_t)kQueueXoff1 * sizeof(int32_t)));
        plan->sub_slot_bytes = off_tot - (size_t)kQueueXoff1 * sizeof(int32_t) + (size_t)3 * k * sizeof(int32_t);
        for (int i = 0; i < BatchPlan::kSubSlots; ++i) {
            TT_HIP(hipHostMalloc((void **)&plan->h_sub[i], plan->sub_slot_bytes, hipHostMallocDefault));
            TT_HIP(hipEventCreateWithFlags(&plan->ev_sub[i], hipEventDisableTiming));
        }
        TT_HIP(hipEventCreateWithFlags(&plan->ev_done, hipEventDisableTiming));
    }
    // host slot: [xoff (16) | items | totals], the same layout as the device block after the counters
    const int slot = plan->sub_next;
    plan->sub_next = (slot + 1) % BatchPlan::kSubSlots;
    TT_HIP(hipEventSynchronize(plan->ev_sub[slot]));  // its previous copy has been read
    int32_t *hx = reinterpret_cast<int32_t *>(plan->h_sub[slot]);
    BatchItem *hi = reinterpret_cast<BatchItem *>(plan->h_sub[slot] + 16 * sizeof(int32_t));
    int32_t *ht = reinterpret_cast<int32_t *>(plan->h_sub[slot] + (off_tot - (size_t)kQueueXoff1 * sizeof(int32_t)));
    std::fill(hx, hx + 16, 0);
    int ni = 0, nt = 0;
    for (int x = 0; x < kXcds; ++x) {
        hx[x] = ni;
        for (int i : plan->order) {
            if (!is[i] || plan->die_of[i] != x) continue;
            const int f = plan->first[ci][i], cnt = plan->count[ci][i];
            ht[3 * nt] = i;
            ht[3 * nt + 1] = ni;
            ht[3 * nt + 2] = cnt;
            ++nt;
            std::copy(all.begin() + f, all.begin() + f + cnt, hi + ni);
            ni += cnt;
        }
    }
    hx[kXcds] = ni;
    for (int j = 0; j < m; ++j) {  // the run follows each member's own queued work
        tt_tree *T = trees[members[j]];
        if (!T->ev_order) TT_HIP(hipEventCreateWithFlags(&T->ev_order, hipEventDisableTiming));
        TT_HIP(hipEventRecord(T->ev_order, T->stream));
        TT_HIP(hipStreamWaitEvent(s, T->ev_order, 0));
    }
    TT_HIP(hipMemcpyAsync(plan->d_sub_queue + kQueueXoff1, plan->h_sub[slot], plan->sub_slot_bytes,
                          hipMemcpyHostToDevice, s));
    TT_HIP(hipEventRecord(plan->ev_sub[slot], s));
    const int *xoff = plan->d_sub_queue + kQueueXoff1;
    if (plan->joint)
        TT_HIP(ttk::batch_joint(n, plan->d_trees, plan->d_sub_items, ni, plan->d_sub_queue, xoff, plan->d_sub_out, mode,
                                c, s));
    else
        TT_HIP(ttk::batch_marginal(n, plan->d_trees, plan->d_sub_items, ni, plan->d_sub_queue, xoff,
                                   plan->d_sub_out, mode, c, s));
    TT_HIP(ttk::batch_totals(plan->d_trees, plan->d_sub_tot, nt, plan->d_sub_out, s));
    TT_HIP(hipEventRecord(plan->ev_done, s));
    for (int j = 0; j < m; ++j) {
        tt_tree *T = trees[members[j]];
        TT_HIP(hipStreamWaitEvent(T->stream, plan->ev_done, 0));
        T->last_stream = T->stream;
        if (T->update_mode && pass != 2) T->offs_valid = true;  // the plan's rows carry d.offs
        if (pass == 3) {
            T->last_pass = 3;
            T->profiles_fresh = false;
            continue;
        }
        T->last_pass = pass == 2 ? 2 : 1;
        T->have_states = true;
        T->tips_dirty = fl != 0;
        if (pass != 2) T->profiles_fresh = true;
    }
    if (logL) {
        TT_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < m; ++j)
            TT_HIP(hipMemcpy(&logL[j], trees[members[j]]->d.total, sizeof(double), hipMemcpyDeviceToHost));
    }
    return TT_OK;
}

int32_t batch_plan_run(tt_engine *eng, tt_tree *const *trees, int32_t k, uint32_t flags, int kind, double *logL,
                       void *stream, const int32_t *members = nullptr, int32_t m = -1) {
    const bool loglik = kind == 1, joint = kind == 2;
    if (!eng || (k > 0 && !trees)) return fail(TT_EINVAL, "NULL argument");
    if (k <= 0) return TT_OK;
    const int n = trees[0] ? trees[0]->n : 0;
    for (int i = 0; i < k; ++i) {
        tt_tree *T = trees[i];
        if (!T) return fail(TT_EINVAL, "trees[%d] is NULL", i);
        if (T->eng != eng) return fail(TT_EINVAL, "trees[%d] belongs to another engine", i);
        if (T->n != n) return fail(TT_EINVAL, "trees[%d] has %d states, trees[0] has %d", i, T->n, n);
        if (!T->have_model) return fail(TT_ESTATE, "trees[%d]: tt_tree_set_model has not been called", i);
        if (T->site) return fail(TT_EINVAL, "trees[%d] has a site-specific model: run it with tt_marginal", i);
    }
    TT_HIP(hipSetDevice(eng->device));
    if (joint)
        for (int i = 0; i < k; ++i)
            if (int32_t rc = ensure_joint(trees[i])) return rc;
    const uint32_t fl = loglik ? 0u : (flags & TT_RECONSTRUCT_TIPS);
    const uint32_t pflags = fl | (loglik ? kPlanLoglik : 0u) | (joint ? kPlanJoint : 0u);
    const int mode = (fl ? 1 : 0) | (loglik ? 2 : 0);
    std::vector<tt_tree *> key(trees, trees + k);
    BatchPlan *plan = nullptr;
    bool created = false;
    {
        std::lock_guard<std::mutex> g(eng->mu);
        for (BatchPlan *p : eng->plans)
            if (p->flags == pflags && p->trees == key) plan = p;
    }
    if (!plan) {
        auto *P = new BatchPlan;
        P->eng = eng;
        P->trees = key;
        P->flags = pflags;
        P->joint = joint;
        // items, longest tree first (LPT), tiles in ascending order within a tree
        std::vector<std::pair<long, int>> order;
        for (int i = 0; i < k; ++i) {
            long pre = loglik ? 0L : (long)trees[i]->pre_nodes.size();
            if (!loglik && !fl) pre -= trees[i]->n_tips;
            order.push_back({(long)trees[i]->n_int + pre, i});
        }
        std::stable_sort(order.begin(), order.end(), [](auto &a, auto &b) { return a.first > b.first; });
        // trees are dealt to the dies longest first, each to the least loaded
        // die (work: tiles x node steps), so that a tree's P(t) tables and
        // rows stay in one die's L2; each die's items keep the global order
        std::vector<int> die_of(k, 0);
        {
            std::vector<double> load(kXcds, 0.0);
            for (auto &o : order) {
                const int tiles = (trees[o.second]->P + TT_WAVE - 1) / TT_WAVE;
                const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                die_of[o.second] = x;
                load[x] += (double)tiles * (double)std::max(1L, o.first);
            }
        }
        std::vector<int32_t> xoff(2 * 16, 0);
        for (int c = 0; c < 2; ++c) {
            P->first[c].assign(k, 0);
            P->count[c].assign(k, 0);
        }
        for (int x = 0; x < kXcds; ++x) {
            xoff[x] = (int32_t)P->items1.size();
            xoff[16 + x] = (int32_t)P->items2.size();
            for (auto &o : order) {
                if (die_of[o.second] != x) continue;
                const int tiles = (trees[o.second]->P + TT_WAVE - 1) / TT_WAVE;
                P->first[0][o.second] = (int)P->items1.size();
                P->first[1][o.second] = (int)P->items2.size();
                for (int t = 0; t < tiles; ++t) P->items1.push_back({o.second, t});
                for (int t = 0; t < tiles; t += 2) P->items2.push_back({o.second, t});
                P->count[0][o.second] = (int)P->items1.size() - P->first[0][o.second];
                P->count[1][o.second] = (int)P->items2.size() - P->first[1][o.second];
            }
        }
        for (auto &o : order) P->order.push_back(o.second);
        P->die_of = die_of;
        std::vector<int32_t> tot;
        for (int c = 0; c < 2; ++c)
            for (int i = 0; i < k; ++i) tot.insert(tot.end(), {i, P->first[c][i], P->count[c][i]});
        xoff[kXcds] = (int32_t)P->items1.size();
        xoff[16 + kXcds] = (int32_t)P->items2.size();
        std::vector<BatchTreeDev> bt(k);
        for (int i = 0; i < k; ++i) {
            tt_tree *T = trees[i];
            bt[i].T = T->d;
            bt[i].T.jtab = nullptr;  // the tip tables belong to the level-kernel joint pass
            bt[i].post_nodes = (decltype(bt[i].post_nodes))T->d_post_nodes;
            bt[i].post_desc = (decltype(bt[i].post_desc))T->d_post_desc;
            bt[i].pre_desc = (decltype(bt[i].pre_desc))T->d_pre_desc;
            bt[i].post_off = (decltype(bt[i].post_off))T->d_post_off;
            bt[i].pre_nodes = (decltype(bt[i].pre_nodes))T->d_pre_nodes;
            bt[i].pre_off = (decltype(bt[i].pre_off))T->d_pre_off;
            bt[i].pre_nint = (decltype(bt[i].pre_nint))T->d_pre_nint;
            bt[i].tip_nodes = (decltype(bt[i].tip_nodes))T->d_tip_nodes;
            bt[i].H = (int32_t)T->post_lv.size();
            bt[i].D = (int32_t)T->pre_lv.size();
            bt[i].fill_tips = 0;
        }
        const size_t nb_trees = bt.size() * sizeof(BatchTreeDev), nb_items1 = P->items1.size() * sizeof(BatchItem),
                     nb_items2 = P->items2.size() * sizeof(BatchItem), nb_out = P->items1.size() * sizeof(double);
        P->bytes = nb_trees + nb_items1 + nb_items2 + nb_out + kQueueInts * sizeof(int) + tot.size() * sizeof(int32_t);
        if (eng->budget && eng->used.load() + P->bytes > eng->budget) {
            P->bytes = 0;
            delete P;
            return fail(TT_ENOMEM_BUDGET, "HBM budget exceeded by the batch plan");
        }
        eng->used += P->bytes;
        hipError_t e = hipMalloc((void **)&P->d_trees, nb_trees);
        if (e == hipSuccess) e = hipMalloc((void **)&P->d_items1, nb_items1);
        if (e == hipSuccess) e = hipMalloc((void **)&P->d_items2, nb_items2);
        if (e == hipSuccess) e = hipMalloc((void **)&P->d_queue, kQueueInts * sizeof(int));
        if (e == hipSuccess) e = hipMemset(P->d_queue, 0, kQueueInts * sizeof(int));
        if (e == hipSuccess)
            e = hipMemcpy(P->d_queue + kQueueXoff1, xoff.data(), xoff.size() * sizeof(int32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc((void **)&P->d_item_out, nb_out);
        if (e == hipSuccess) e = hipMalloc((void **)&P->d_tot, tot.size() * sizeof(int32_t));
        if (e == hipSuccess)
            e = hipMemcpy(P->d_tot, tot.data(), tot.size() * sizeof(int32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(P->d_trees, bt.data(), nb_trees, hipMemcpyHostToDevice);
        P->host_trees = bt;
        if (e == hipSuccess) e = hipMemcpy(P->d_items1, P->items1.data(), nb_items1, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(P->d_items2, P->items2.data(), nb_items2, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete P;
            return fail(TT_EDEVICE, "batch plan set-up: %s", hipGetErrorString(e));
        }
        if (int32_t rc = tune_batch(P, n, mode, P->stream)) {
            delete P;
            return rc;
        }
        if (g_use_graphs() && P->shape.wave_items != kPerTreeGraphs) {
            hipGraph_t g = nullptr;
            e = hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal);
            hipError_t e2 = hipSuccess;
            if (e == hipSuccess)
                e2 = P->launch(n, mode, P->shape, P->stream);
            if (e == hipSuccess) e = hipStreamEndCapture(P->stream, &g);
            if (e == hipSuccess) e = e2;
            if (e == hipSuccess) e = hipGraphInstantiate(&P->graph, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                delete P;
                return fail(TT_EDEVICE, "batch graph capture: %s", hipGetErrorString(e));
            }
        }
        std::lock_guard<std::mutex> g(eng->mu);
        eng->plans.push_back(P);
        plan = P;
        created = true;
    }
    hipStream_t s = stream ? (hipStream_t)stream : plan->stream;
    // tip rows that an earlier pass reconstructed are reset to 255 (read by the
    // kernel at run time); a tree switched to update mode after the plan was
    // built gets its normaliser rows into the plan's copy
    bool changed = false;
    for (int i = 0; i < k; ++i) {
        BatchTreeDev &b = plan->host_trees[i];
        if (b.T.offs != trees[i]->d.offs) {
            b.T.offs = trees[i]->d.offs;
            changed = true;
        }
        if (loglik) continue;
        const int32_t want = (!fl && trees[i]->tips_dirty) ? 1 : 0;
        if (b.fill_tips != want) {
            b.fill_tips = want;
            changed = true;
        }
    }
    if (changed)
        TT_HIP(hipMemcpyAsync(plan->d_trees, plan->host_trees.data(), plan->host_trees.size() * sizeof(BatchTreeDev),
                              hipMemcpyHostToDevice, s));
    if (m >= 0) {
        if (created)  // the autotuner ran full passes over every tree of the plan
            for (int i = 0; i < k; ++i) {
                trees[i]->last_stream = plan->stream;
                trees[i]->last_pass = joint ? 2 : (loglik ? 3 : 1);
                if (!loglik) trees[i]->have_states = true;
                if (!loglik) trees[i]->tips_dirty = fl != 0;
                trees[i]->profiles_fresh = !joint && !loglik;
                if (trees[i]->update_mode && !joint) trees[i]->offs_valid = true;
            }
        return plan_run_members(plan, trees, k, n, mode, fl, joint ? 2 : (loglik ? 3 : 1), members, m, logL, s);
    }
    if (plan->graph) {
        TT_HIP(hipGraphLaunch(plan->graph, s));
    } else if (int32_t rc = launch_plan(plan, n, mode, plan->shape, s)) {
        return rc;
    }
    for (int i = 0; i < k; ++i) {
        trees[i]->last_stream = s;
        if (trees[i]->update_mode && !joint) trees[i]->offs_valid = true;  // the plan's rows carry d.offs
        if (loglik) {  // subtree rows rewritten, profiles and states not
            trees[i]->last_pass = 3;
            trees[i]->profiles_fresh = false;
            continue;
        }
        trees[i]->last_pass = joint ? 2 : 1;
        trees[i]->have_states = true;
        trees[i]->tips_dirty = fl != 0;
        if (!joint) trees[i]->profiles_fresh = true;
    }
    if (logL && plan->shape.wave_items == kPerTreeGraphs) {
        for (int i = 0; i < k; ++i)
            TT_HIP(hipMemcpyAsync(&logL[i], trees[i]->d.total, sizeof(double), hipMemcpyDeviceToHost, s));
        TT_HIP(hipStreamSynchronize(s));
    } else if (logL) {
        const std::vector<BatchItem> &items = plan->items_for(plan->shape);
        std::vector<double> out(items.size());
        TT_HIP(hipMemcpyAsync(out.data(), plan->d_item_out, out.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        TT_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < k; ++i) logL[i] = 0.0;
        for (size_t j = 0; j < out.size(); ++j) logL[items[j].tree] += out[j];
    }
    return TT_OK;
}

}  // namespace

extern "C" {

int32_t tt_batch_marginal(tt_engine *eng, tt_tree *const *trees, int32_t k, uint32_t flags, double *logL,
                          void *stream) {
    return batch_plan_run(eng, trees, k, flags, 0, logL, stream);
}

int32_t tt_batch_run_members(tt_engine *eng, tt_tree *const *pool, int32_t k, uint32_t mode, const int32_t *members,
                             int32_t m, double *logL, void *stream) {
    const uint32_t kind = mode & 0x30u;
    if (mode & ~(0x30u | TT_RECONSTRUCT_TIPS)) return fail(TT_EINVAL, "unknown mode bits 0x%x", mode);
    if (m < 0 || (m > 0 && !members)) return fail(TT_EINVAL, "members: NULL or m=%d < 0", m);
    if (kind == TT_BATCH_MARGINAL) return batch_plan_run(eng, pool, k, mode, 0, logL, stream, members, m);
    if (kind == TT_BATCH_LOGLIK) return batch_plan_run(eng, pool, k, 0u, 1, logL, stream, members, m);
    if (kind == TT_BATCH_JOINT) return batch_plan_run(eng, pool, k, mode, 2, logL, stream, members, m);
    return fail(TT_EINVAL, "unknown pass kind 0x%x", kind);
}

int32_t tt_batch_run(tt_engine *eng, tt_tree *const *trees, int32_t k, uint32_t mode, double *logL, void *stream) {
    const uint32_t kind = mode & 0x30u;
    if (mode & ~(0x30u | TT_RECONSTRUCT_TIPS)) return fail(TT_EINVAL, "unknown mode bits 0x%x", mode);
    if (kind == TT_BATCH_MARGINAL) return batch_plan_run(eng, trees, k, mode, 0, logL, stream);
    if (kind == TT_BATCH_LOGLIK) return batch_plan_run(eng, trees, k, 0u, 1, logL, stream);
    if (kind == TT_BATCH_JOINT) return batch_plan_run(eng, trees, k, mode, 2, logL, stream);
    return fail(TT_EINVAL, "unknown pass kind 0x%x", kind);
}

namespace {

// Grow one of the engine's compression buffers to at least `need` bytes
// (contents are not kept), counting it against the HBM budget.
int32_t grow(tt_engine *eng, void **buf, size_t *have, size_t need) {
    if (need <= *have) return TT_OK;
    need = std::max<size_t>(need, 256);
    const uint64_t cur = eng->used.load() - *have;
    if (eng->budget && cur + need > eng->budget)
        return fail(TT_ENOMEM_BUDGET, "HBM budget exceeded: %llu in use + %llu requested > budget %llu",
                    (unsigned long long)cur, (unsigned long long)need, (unsigned long long)eng->budget);
    if (*buf) TT_HIP(hipFree(*buf));
    eng->used -= *have;
    *buf = nullptr;
    *have = 0;
    TT_HIP(hipMalloc(buf, need));
    *have = need;
    eng->used += need;
    return TT_OK;
}

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

int32_t tt_compress_alignment(tt_engine *eng, const uint8_t *chars, int32_t n_seqs, int64_t n_columns,
                              int64_t row_stride, int32_t ambiguous, uint32_t flags, int64_t capacity,
                              int64_t *n_patterns, int32_t *pattern_of_column, uint8_t *patterns,
                              double *multiplicity, int64_t *char_counts, void *stream) {
    if (!eng || !n_patterns || (n_columns > 0 && (!chars || !pattern_of_column)))
        return fail(TT_EINVAL, "NULL argument");
    if (n_seqs < 1) return fail(TT_EINVAL, "n_seqs=%d: need at least one sequence", n_seqs);
    if (n_columns < 0 || n_columns > INT32_MAX / 2)
        return fail(TT_EINVAL, "n_columns=%lld out of range [0, %d]", (long long)n_columns, INT32_MAX / 2);
    if (row_stride < n_columns) return fail(TT_EINVAL, "row_stride %lld < n_columns %lld", (long long)row_stride,
                                            (long long)n_columns);
    if (ambiguous < -1 || ambiguous > 255) return fail(TT_EINVAL, "ambiguous=%d: not a byte or -1", ambiguous);
    if (flags & ~(TT_FILL_OVERHANGS | TT_COMPRESS_DEVICE)) return fail(TT_EINVAL, "unknown flags 0x%x", flags);
    if (capacity < 0) return fail(TT_EINVAL, "capacity < 0");
    if (capacity > 0 && (!patterns || !multiplicity)) return fail(TT_EINVAL, "NULL patterns or multiplicity");
    const bool dev = flags & TT_COMPRESS_DEVICE;
    const bool fill = (flags & TT_FILL_OVERHANGS) && ambiguous >= 0;
    const int L = (int)n_columns;
    *n_patterns = 0;
    if (char_counts) std::fill(char_counts, char_counts + 256, 0);
    if (L == 0) return TT_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> g(eng->cmu);
    TT_HIP(hipSetDevice(eng->device));

    const CompressLayout l = ttk::compress_layout(n_seqs, L);
    int32_t rc = grow(eng, &eng->cscratch, &eng->cscratch_bytes, l.total);
    if (rc) return rc;
    const uint8_t *A = chars;
    int64_t pitch = row_stride;
    int *pat = pattern_of_column;
    double *mult = multiplicity;
    uint8_t *out = patterns;
    size_t off_pat = 0, off_mult = 0, off_out = 0;
    if (!dev) {
        // staging: input rows at a 16-byte pitch, then pattern map, multiplicities, compressed rows
        const int64_t pitch16 = (n_columns + 15) / 16 * 16;
        size_t o = up256((size_t)n_seqs * pitch16);
        off_pat = o;
        o += up256((size_t)L * sizeof(int));
        off_mult = o;
        o += up256((size_t)L * sizeof(double));
        off_out = o;
        o += up256((size_t)n_seqs * L);
        rc = grow(eng, &eng->cstage, &eng->cstage_bytes, o);
        if (rc) return rc;
        char *st = static_cast<char *>(eng->cstage);
        TT_HIP(hipMemcpy2DAsync(st, pitch16, chars, row_stride, n_columns, n_seqs, hipMemcpyHostToDevice, s));
        A = reinterpret_cast<const uint8_t *>(st);
        pitch = pitch16;
        pat = reinterpret_cast<int *>(st + off_pat);
        mult = reinterpret_cast<double *>(st + off_mult);
        out = reinterpret_cast<uint8_t *>(st + off_out);
    }
    TT_HIP(ttk::compress_columns(A, pitch, n_seqs, L, ambiguous, fill, char_counts != nullptr, eng->cscratch, l, pat,
                                 s));
    // device outputs: the caller's capacity bounds the writes; staged host
    // outputs: the staging holds every possible pattern
    TT_HIP(ttk::compress_finish(A, pitch, n_seqs, L, ambiguous, fill, eng->cscratch, l, dev ? capacity : (int64_t)L,
                                pat, mult, out, s));
    int P = 0;
    TT_HIP(hipMemcpyAsync(&P, ttk::compress_total_ptr(eng->cscratch, l), sizeof(int), hipMemcpyDeviceToHost, s));
    unsigned long long hist[256];
    if (char_counts)
        TT_HIP(hipMemcpyAsync(hist, ttk::compress_hist_ptr(eng->cscratch, l), sizeof hist, hipMemcpyDeviceToHost, s));
    TT_HIP(hipStreamSynchronize(s));
    *n_patterns = P;
    if (P > capacity)
        return fail(TT_ERANGE, "%d patterns exceed the output capacity %lld", P, (long long)capacity);
    if (!dev) {
        TT_HIP(hipMemcpyAsync(pattern_of_column, pat, (size_t)L * sizeof(int), hipMemcpyDeviceToHost, s));
        TT_HIP(hipMemcpyAsync(multiplicity, mult, (size_t)P * sizeof(double), hipMemcpyDeviceToHost, s));
        TT_HIP(hipMemcpyAsync(patterns, out, (size_t)n_seqs * P, hipMemcpyDeviceToHost, s));
        TT_HIP(hipStreamSynchronize(s));
    }
    if (char_counts)
        for (int c = 0; c < 256; ++c) char_counts[c] = (int64_t)hist[c];
    return TT_OK;
}

}  // extern "C"

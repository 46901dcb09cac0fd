"""Python handles over the C ABI: ``Engine`` (device + HBM budget) and
``DeviceTree`` (one resident tree with its compressed alignment)."""
import ctypes as C

import numpy as np

from . import _native

N_CODES = 256
MISSING_CODE = 0  # character code of tips without a sequence (all-ones profile)


def code_lut(profile_map, n_states):
    """(256, n) profile of every byte; characters outside the profile map are
    treated as missing information (seq_utils.extend_profile semantics)."""
    lut = np.ones((N_CODES, n_states), dtype=np.float32)
    for c, v in profile_map.items():
        if isinstance(c, str) and len(c) == 1 and ord(c) < N_CODES:
            lut[ord(c)] = v
    lut[MISSING_CODE] = 1.0
    return lut


class Engine:
    """One device and its HBM admission budget (0 = unlimited)."""

    def __init__(self, device=0, hbm_budget=0):
        self.lib = _native.load()
        h = C.c_void_p()
        _native.check(self.lib.tt_engine_create(int(device), int(hbm_budget), C.byref(h)))
        self.handle = h
        self.device = device

    def usage(self):
        used, budget = C.c_uint64(), C.c_uint64()
        _native.check(self.lib.tt_engine_usage(self.handle, C.byref(used), C.byref(budget)))
        return used.value, budget.value

    def compress_alignment(self, codes, ambiguous, fill_overhangs=False, char_counts=False):
        """Pattern compression of a (n_seqs, L) uint8 alignment on the device
        (tt_compress_alignment; sequence_data.py:325-464, with the overhang
        fill of seq_utils.py:196-202 when fill_overhangs).

        Returns (pattern_of_column (L,) int64, patterns (n_seqs, P) uint8,
        multiplicity (P,) float64, byte counts (256,) int64 or None)."""
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        if codes.ndim != 2:
            raise ValueError("alignment must be a (n_seqs, L) byte matrix")
        N, L = codes.shape
        amb = -1 if ambiguous is None else (ord(ambiguous) if isinstance(ambiguous, str) else int(ambiguous))
        flags = _native.TT_FILL_OVERHANGS if fill_overhangs else 0
        pat = np.empty(L, dtype=np.int32)
        out = np.empty(N * max(L, 1), dtype=np.uint8)
        mult = np.empty(max(L, 1), dtype=np.float64)
        counts = np.zeros(256, dtype=np.int64) if char_counts else None
        P = C.c_int64()
        _native.check(self.lib.tt_compress_alignment(
            self.handle, codes.ctypes.data, N, L, L, amb, flags, L, C.byref(P), pat.ctypes.data, out.ctypes.data,
            mult.ctypes.data, counts.ctypes.data_as(_native._PI64) if counts is not None else None, None))
        P = P.value
        return pat.astype(np.int64), out[:N * P].reshape(N, P).copy(), mult[:P].copy(), counts

    def compress_alignment_device(self, chars, n_columns, ambiguous, pat, patterns, mult, fill_overhangs=False,
                                  stream=None):
        """Device-pointer form: chars is a (n_seqs, pitch) uint8 CUDA tensor
        holding n_columns columns per row; pat (L,) int32, patterns
        (n_seqs * capacity,) uint8 and mult (capacity,) float64 CUDA tensors
        receive the results.  Returns P."""
        N, pitch = chars.shape
        amb = -1 if ambiguous is None else ord(ambiguous)
        flags = _native.TT_COMPRESS_DEVICE | (_native.TT_FILL_OVERHANGS if fill_overhangs else 0)
        cap = mult.numel()
        if patterns.numel() < N * cap or pat.numel() < n_columns:
            raise ValueError("output tensors too small")
        P = C.c_int64()
        _native.check(self.lib.tt_compress_alignment(
            self.handle, chars.data_ptr(), N, int(n_columns), int(chars.stride(0)), amb, flags, cap, C.byref(P),
            pat.data_ptr(), patterns.data_ptr(), mult.data_ptr(), None, stream))
        return P.value

    def close(self):
        """Release the engine; fails (MissingDataError) while trees created
        on it are still open."""
        if self.handle:
            _native.check(self.lib.tt_engine_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_engine = None


def default_engine():
    global _default_engine
    if _default_engine is None:
        _default_engine = Engine(0, 0)
    return _default_engine


class DeviceTree:
    """A tree resident on the device.

    parent : (n_nodes,) int32, -1 at the root; children are visited in
             ascending node index.
    t      : (n_nodes,) effective branch lengths.
    tip_codes : (n_tips, P) uint8 character codes, rows in ascending node
             index of the tips.
    """

    def __init__(self, engine, parent, t, tip_codes, lut, multiplicity, pi, eigvals, v, v_inv, mu, site=None):
        self.engine = engine
        self.lib = engine.lib
        self.parent = _native.as_c(parent, np.int32)
        self.n_nodes = len(self.parent)
        self.P = int(tip_codes.shape[1])
        self.n = int(lut.shape[1])
        t = _native.as_c(t, np.float64)
        codes = _native.as_c(tip_codes, np.uint8)
        lut = _native.as_c(lut, np.float32)
        mult = _native.as_c(multiplicity, np.float64)
        pi = _native.as_c(pi, np.float64)
        d = _native.tt_tree_desc(
            self.n_nodes, self.n, self.P, lut.shape[0],
            _native.ptr(self.parent, C.c_int32), _native.ptr(t, C.c_double), _native.ptr(codes, C.c_uint8),
            _native.ptr(lut, C.c_float), _native.ptr(mult, C.c_double), _native.ptr(pi, C.c_double))
        h = C.c_void_p()
        _native.check(self.lib.tt_tree_create(engine.handle, C.byref(d), C.byref(h)))
        self.handle = h
        self.t_cur = t.copy()        # current effective branch lengths (host mirror)
        self.joint_ready = False     # joint buffers allocated on the device
        self.site = False            # a site-specific model is set
        if site is not None:         # (eig, V, Vi, mu, pi, grid, rate_scale) of a site-specific model
            self.set_site_model(*site)
        else:
            self.set_model(eigvals, v, v_inv, mu)

    def set_model(self, eigvals, v, v_inv, mu):
        e = _native.as_c(eigvals, np.float64)
        vv = _native.as_c(v, np.float64)
        vi = _native.as_c(v_inv, np.float64)
        _native.check(self.lib.tt_tree_set_model(self.handle, _native.ptr(e, C.c_double), _native.ptr(vv, C.c_double),
                                                 _native.ptr(vi, C.c_double), float(mu)))
        if self.site:
            # the next joint pass reallocates the joint buffers at the FP32
            # message size (the library frees the FP64 ones first)
            self.joint_ready = False
        self.site = False

    def set_site_model(self, eig, V, Vi, mu, pi, grid, rate_scale):
        """Per-pattern model (tt_tree_set_site_model): eig (P, n), V and Vi
        (P, n, n), mu (P,), pi (P, n); grid (G,) empty for exact P(t)."""
        arrs = [_native.as_c(x, np.float64) for x in (eig, V, Vi, mu, pi, grid)]
        self._site_arrays = arrs
        g = arrs[5]
        _native.check(self.lib.tt_tree_set_site_model(
            self.handle, *[_native.ptr(x, C.c_double) for x in arrs[:5]],
            _native.ptr(g, C.c_double) if len(g) else None, len(g), float(rate_scale)))
        if not self.site:
            self.joint_ready = False  # joint buffers are reallocated at the FP64 message size
        self.site = True

    def nonroot_nodes(self):
        return np.nonzero(self.parent >= 0)[0]

    def path_len(self, node):
        """Internal nodes from the branch's parent up to the root."""
        k, a = 0, int(self.parent[node])
        while a >= 0:
            k += 1
            a = int(self.parent[a])
        return k

    def set_branch_lengths(self, nodes, t):
        nodes = _native.as_c(nodes, np.int32)
        t = _native.as_c(t, np.float64)
        self.t_cur[nodes] = t
        _native.check(self.lib.tt_tree_set_branch_lengths(self.handle, _native.ptr(nodes, C.c_int32),
                                                          _native.ptr(t, C.c_double), len(nodes)))

    def info(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_uint64()
        _native.check(self.lib.tt_tree_info(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return {"levels_post": a.value, "levels_pre": b.value, "device_bytes": c.value}

    def pass_shape(self):
        """(pattern tile of the one-launch pass or 0 for the level kernels,
        whether its records are staged in LDS, whether its LDS rows are
        reused by lifetime)."""
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        _native.check(self.lib.tt_tree_pass_shape(self.handle, C.byref(a), C.byref(b), C.byref(c), None))
        return a.value, bool(b.value), bool(c.value)

    def joint_one_launch(self):
        """Whether the joint pass runs in one launch over the marginal pass's tiles."""
        j = C.c_int32()
        _native.check(self.lib.tt_tree_pass_shape(self.handle, None, None, None, C.byref(j)))
        return bool(j.value)

    def marginal(self, reconstruct_tips=False, stream=None, want_logL=True, keep_normalisers=False):
        out = C.c_double(np.nan)
        flags = _native.TT_RECONSTRUCT_TIPS if reconstruct_tips else 0
        if keep_normalisers:
            flags |= _native.TT_KEEP_NORMALISERS
        _native.check(self.lib.tt_marginal(self.handle, flags, C.byref(out) if want_logL else None, stream))
        return out.value

    def joint(self, reconstruct_tips=False, stream=None, want_logL=True):
        out = C.c_double(np.nan)
        flags = _native.TT_RECONSTRUCT_TIPS if reconstruct_tips else 0
        _native.check(self.lib.tt_joint(self.handle, flags, C.byref(out) if want_logL else None, stream))
        self.joint_ready = True
        return out.value

    def joint_root(self):
        """normalized_profile of the root after a joint pass (treeanc.py:1019):
        the root's joint log message less its per-pattern maximum, (P, n)."""
        out = np.empty((self.P, self.n), dtype=np.float64)
        _native.check(self.lib.tt_get_joint_root(self.handle, _native.ptr(out, C.c_double)))
        return out

    def joint_from_root(self, root_states, reconstruct_tips=False, stream=None, want_logL=True):
        """Finish the last joint pass from given root states (backtrack and
        sequence_LH at those states); returns sequence_joint_LH."""
        rs = _native.as_c(root_states, np.uint8)
        if rs.shape != (self.P,):
            raise ValueError("root_states must hold one state per pattern")
        out = C.c_double(np.nan)
        flags = _native.TT_RECONSTRUCT_TIPS if reconstruct_tips else 0
        _native.check(self.lib.tt_joint_from_root(self.handle, flags, _native.ptr(rs, C.c_uint8),
                                                  C.byref(out) if want_logL else None, stream))
        return out.value

    def joint_debug(self):
        """joint_Lx (n_nodes, P, n) FP64 and joint_Cx (n_nodes, P, n) uint8 of
        the last joint pass (tt_get_joint_debug)."""
        lx = np.empty((self.n_nodes, self.P, self.n), dtype=np.float64)
        cx = np.empty((self.n_nodes, self.P, self.n), dtype=np.uint8)
        _native.check(self.lib.tt_get_joint_debug(self.handle, _native.ptr(lx, C.c_double),
                                                  _native.ptr(cx, C.c_uint8)))
        return lx, cx

    def subtree_prefactor(self):
        """marginal_subtree_LH_prefactor (n_nodes, P) of the last marginal pass
        run with keep_normalisers (tt_get_subtree_prefactor)."""
        out = np.empty((self.n_nodes, self.P), dtype=np.float64)
        _native.check(self.lib.tt_get_subtree_prefactor(self.handle, _native.ptr(out, C.c_double)))
        return out

    def branch_mutation_matrix(self, node):
        """get_branch_mutation_matrix of the branch above ``node`` from the
        last marginal pass: (P, n, n), [p, child state, parent state]."""
        out = np.empty((self.P, self.n, self.n), dtype=np.float64)
        _native.check(self.lib.tt_branch_mutation_matrix(self.handle, int(node), _native.ptr(out, C.c_double)))
        return out

    def loglik(self, stream=None, want_logL=True):
        out = C.c_double(np.nan)
        _native.check(self.lib.tt_loglik(self.handle, C.byref(out) if want_logL else None, stream))
        return out.value

    def update_branch(self, node, t, stream=None, want_logL=True):
        """Set one branch length and re-evaluate logL along its root path."""
        out = C.c_double(np.nan)
        _native.check(self.lib.tt_update_branch(self.handle, int(node), float(t),
                                                C.byref(out) if want_logL else None, stream))
        self.t_cur[int(node)] = t
        return out.value

    def optimize_branch_lengths(self, code_state, gap_state=-1, max_t=4.0, tol=1e-10, stream=None,
                                with_counts=True):
        """One sweep of joint branch-length optimisation from the states of
        the last reconstruction (tt_optimize_branch_lengths).  Returns
        (t_opt, status, pair_counts or None); the tree's lengths are left as
        they are."""
        cs = _native.as_c(code_state, np.int16)
        if cs.shape != (N_CODES,):
            raise ValueError("code_state must have 256 entries")
        t = np.empty(self.n_nodes, dtype=np.float64)
        st = np.empty(self.n_nodes, dtype=np.int32)
        cnt = np.empty((self.n_nodes, self.n, self.n), dtype=np.float64) if with_counts else None
        _native.check(self.lib.tt_optimize_branch_lengths(
            self.handle, _native.ptr(cs, C.c_int16), int(gap_state), float(max_t), float(tol),
            _native.ptr(t, C.c_double), _native.ptr(st, C.c_int32),
            _native.ptr(cnt, C.c_double) if with_counts else None, stream))
        return t, st, cnt

    def optimize_branch_lengths_marginal(self, gap_state=-1, max_t=4.0, tol=1e-10, stream=None):
        """Optimal length of every branch from the profile pairs of the last
        marginal pass (tt_optimize_branch_lengths_marginal).  Returns
        (t_opt, status, (t_pair, status_pair)) -- the last for the summed
        length of a bifurcating root's two children, (nan, -1) otherwise."""
        t = np.empty(self.n_nodes, dtype=np.float64)
        st = np.empty(self.n_nodes, dtype=np.int32)
        pair = np.empty(2, dtype=np.float64)
        _native.check(self.lib.tt_optimize_branch_lengths_marginal(
            self.handle, int(gap_state), float(max_t), float(tol), _native.ptr(t, C.c_double),
            _native.ptr(st, C.c_int32), _native.ptr(pair, C.c_double), stream))
        return t, st, (float(pair[0]), int(pair[1]))

    def branch_profile_lh(self, t, gap_state=-1, stream=None):
        """prob_t_profiles(return_log=True) of every branch at the lengths t:
        (n_nodes, K)."""
        tt = _native.as_c(np.atleast_1d(t), np.float64)
        out = np.empty((self.n_nodes, len(tt)), dtype=np.float64)
        _native.check(self.lib.tt_branch_profile_lh(self.handle, int(gap_state), _native.ptr(tt, C.c_double),
                                                    len(tt), _native.ptr(out, C.c_double), stream))
        return out

    def branch_state_counts(self, code_state, gap_state=-1, stream=None):
        """(n_nodes, n, n) multiplicity-weighted (parent, child) state counts
        of the last reconstruction (tt_branch_state_counts)."""
        cs = _native.as_c(code_state, np.int16)
        out = np.empty((self.n_nodes, self.n, self.n), dtype=np.float64)
        _native.check(self.lib.tt_branch_state_counts(self.handle, _native.ptr(cs, C.c_int16), int(gap_state),
                                                      _native.ptr(out, C.c_double), stream))
        return out

    def mutation_counts(self, per_branch=False, stream=None):
        """GTR-inference statistics of the last marginal pass
        (tt_mutation_counts): (nij (n, n), Ti (n,)[, per-branch (n_nodes, n*n + n)])."""
        nij = np.empty((self.n, self.n), dtype=np.float64)
        Ti = np.empty(self.n, dtype=np.float64)
        pb = np.empty((self.n_nodes, self.n * self.n + self.n), dtype=np.float64) if per_branch else None
        _native.check(self.lib.tt_mutation_counts(self.handle, _native.ptr(nij, C.c_double), _native.ptr(Ti, C.c_double),
                                                  _native.ptr(pb, C.c_double) if per_branch else None, stream))
        return (nij, Ti, pb) if per_branch else (nij, Ti)

    def site_counts(self, marginal, code_state=None, stream=None):
        """Per-pattern statistics summed over branches (tt_site_counts):
        (n_ija (P, n, n), T_ia (P, n))."""
        n_out = np.empty((self.P, self.n, self.n))
        t_out = np.empty((self.P, self.n))
        cs = _native.as_c(code_state, np.int16) if code_state is not None else None
        _native.check(self.lib.tt_site_counts(self.handle, int(bool(marginal)),
                                              _native.ptr(cs, C.c_int16) if cs is not None else None,
                                              _native.ptr(n_out, C.c_double), _native.ptr(t_out, C.c_double),
                                              stream))
        return n_out, t_out

    def trim(self):
        """Free the joint-pass buffers (returned to the engine budget)."""
        _native.check(self.lib.tt_tree_trim(self.handle))
        self.joint_ready = False

    def joint_bytes(self):
        """Device bytes a joint pass allocates on first use: the messages
        (FP32, FP64 for a site-specific model) and the argmax table."""
        n_int = int(np.count_nonzero(np.bincount(self.parent[self.parent >= 0], minlength=self.n_nodes)))
        return n_int * self.P * self.n * (8 if self.site else 4) + self.n_nodes * self.P * self.n

    def query(self):
        """True when the work queued on the tree's own stream has finished."""
        rc = self.lib.tt_tree_query(self.handle)
        if rc < 0:
            _native.check(rc)
        return rc == 1

    def total(self):
        out = C.c_double(np.nan)
        _native.check(self.lib.tt_get_total(self.handle, C.byref(out)))
        return out.value

    def sequence_lh(self):
        out = np.empty(self.P, dtype=np.float64)
        _native.check(self.lib.tt_get_sequence_lh(self.handle, _native.ptr(out, C.c_double)))
        return out

    def states(self):
        out = np.empty((self.n_nodes, self.P), dtype=np.uint8)
        _native.check(self.lib.tt_get_states(self.handle, _native.ptr(out, C.c_uint8)))
        return out

    def profiles(self, which=_native.TT_PROFILE_MARGINAL):
        out = np.empty((self.n_nodes, self.P, self.n), dtype=np.float32)
        _native.check(self.lib.tt_get_profiles(self.handle, int(which), _native.ptr(out, C.c_float)))
        return out

    def close(self):
        if getattr(self, "handle", None):
            self.lib.tt_tree_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_marginal(engine, trees, reconstruct_tips=False, stream=None, want_logL=True):
    """Marginal reconstruction of several resident trees in one persistent launch."""
    k = len(trees)
    arr = (C.c_void_p * k)(*[t.handle.value for t in trees])
    out = np.full(k, np.nan)
    flags = _native.TT_RECONSTRUCT_TIPS if reconstruct_tips else 0
    _native.check(engine.lib.tt_batch_marginal(engine.handle, arr, k, flags,
                                               _native.ptr(out, C.c_double) if want_logL else None, stream))
    return out


def batch_run(engine, trees, kind="marginal", reconstruct_tips=False, stream=None, want_logL=True):
    """tt_batch_run: kind 'marginal' (one persistent launch), 'loglik'
    (the persistent launch without the preorder) or 'joint' (one persistent
    launch of the joint kernel).  Returns logL per tree."""
    modes = {"marginal": _native.TT_BATCH_MARGINAL, "joint": _native.TT_BATCH_JOINT,
             "loglik": _native.TT_BATCH_LOGLIK}
    if kind not in modes:
        raise ValueError(f"kind must be 'marginal', 'loglik' or 'joint', not {kind!r}")
    k = len(trees)
    arr = (C.c_void_p * k)(*[t.handle.value for t in trees])
    out = np.full(k, np.nan)
    mode = modes[kind]
    if reconstruct_tips:
        mode |= _native.TT_RECONSTRUCT_TIPS
    _native.check(engine.lib.tt_batch_run(engine.handle, arr, k, mode,
                                          _native.ptr(out, C.c_double) if want_logL else None, stream))
    if kind == "joint":
        for t in trees:  # the batched joint pass allocated every member's joint buffers
            t.joint_ready = True
    return out


def batch_run_members(engine, pool, members, kind="marginal", reconstruct_tips=False, stream=None, want_logL=False):
    """tt_batch_run_members: the batched pass of ``kind`` over the trees
    ``pool[i]`` for i in ``members`` only, through the plan (items, tuned
    launch shape) of the whole pool, which the first call builds.  The run is
    ordered after work already queued on each member's stream and before
    later work on it; ``tree.query()`` / ``tree.total()`` of a member see it.
    Returns logL per member (host sync) when ``want_logL``."""
    modes = {"marginal": _native.TT_BATCH_MARGINAL, "joint": _native.TT_BATCH_JOINT,
             "loglik": _native.TT_BATCH_LOGLIK}
    if kind not in modes:
        raise ValueError(f"kind must be 'marginal', 'loglik' or 'joint', not {kind!r}")
    k = len(pool)
    arr = (C.c_void_p * k)(*[t.handle.value for t in pool])
    mem = np.ascontiguousarray(members, dtype=np.int32)
    out = np.full(len(mem), np.nan)
    mode = modes[kind]
    if reconstruct_tips:
        mode |= _native.TT_RECONSTRUCT_TIPS
    _native.check(engine.lib.tt_batch_run_members(engine.handle, arr, k, mode, _native.ptr(mem, C.c_int32),
                                                  len(mem), _native.ptr(out, C.c_double) if want_logL else None,
                                                  stream))
    if kind == "joint":
        for t in pool:  # a joint plan holds every pool tree's joint buffers
            t.joint_ready = True
    return out


def batch_info(engine, trees, reconstruct_tips=False, kind="marginal"):
    """Launch shape chosen by the autotuner for this list of trees:
    {"wave_items", "tiles", "blocks_per_cu", "waves_per_simd", "tuned_ms"}
    (waves_per_simd 0 = the compiler's register budget)."""
    if kind not in ("marginal", "loglik", "joint"):
        raise ValueError(f"kind must be 'marginal', 'loglik' or 'joint', not {kind!r}")
    k = len(trees)
    arr = (C.c_void_p * k)(*[t.handle.value for t in trees])
    w, tl, b, wps, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
    flags = _native.TT_RECONSTRUCT_TIPS if reconstruct_tips else 0
    if kind == "loglik":
        flags = _native.TT_BATCH_LOGLIK
    elif kind == "joint":
        flags |= _native.TT_BATCH_JOINT
    _native.check(engine.lib.tt_batch_info(engine.handle, arr, k, flags, C.byref(w), C.byref(tl), C.byref(b),
                                           C.byref(wps), C.byref(ms)))
    return {"wave_items": w.value, "tiles": tl.value, "blocks_per_cu": b.value, "waves_per_simd": wps.value,
            "tuned_ms": ms.value}

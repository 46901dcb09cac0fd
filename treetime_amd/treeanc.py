"""TreeAnc: the reference's ancestral-reconstruction surface on the HIP engine.

Method names, arguments, return values, node/tree attributes and exceptions
follow treetime/treeanc.py (TreeTime 0.12.1):

* ``infer_ancestral_sequences`` / ``reconstruct_anc`` (treeanc.py:512-570)
  run the marginal (treeanc.py:762-932) or joint (934-1080) pass on the GPU
  through ``libttgpu.so`` and write ``node.marginal_profile``,
  ``node.marginal_subtree_LH``, ``node._cseq``, ``tree.sequence_LH``,
  ``tree.total_sequence_LH`` / ``tree.sequence_joint_LH`` as the reference
  does.
* ``sequence_LH`` (691-716), ``ancestral_likelihood`` (718-750),
  ``optimize_branch_len`` / ``optimize_branch_lengths_joint`` (1168-1270).

* ``infer_gtr`` (1500-1632) with the counts from the device; marginal
  branch lengths ``optimize_tree_marginal`` (1297-1358).

* ``marginal_branch_profile`` / ``get_branch_mutation_matrix``
  (1085-1146) from device rows; ``debug=True`` keeps the reference's debug
  attributes (803-810, 1066-1077).

Fitch parsimony is outside this engine (``NotImplementedError``).
"""
import sys
import time

import numpy as np

from . import config as ttconf
from .engine import MISSING_CODE, DeviceTree, code_lut, default_engine
from .errors import MissingDataError, UnknownMethodError
from .gtr import GTR
from .gtr_site_specific import GTR_site_specific
from .seq_utils import extend_profile, seq2prof
from .sequence_data import SequenceData
from .tree import Clade, Tree


def _compressed_sequence(node):
    """treeanc.py:14-25."""
    tt = node.tt
    if node.name in tt.data.compressed_alignment and not tt.reconstructed_tip_sequences:
        return tt.data.compressed_alignment[node.name]
    if hasattr(node, "_cseq"):
        return node._cseq
    if node.is_terminal():
        return None
    raise ValueError("Ancestral sequences are not yet inferred")


def _mutations(node):
    """treeanc.py:28-42."""
    tt = node.tt
    if node.up is None:
        return []
    if (not tt.reconstructed_tip_sequences) and node.name in tt.data.aln:
        return tt.data.differences(node.up.cseq, tt.data.aln[node.name], seq2_compressed=False)
    if node.is_terminal() and node.name not in tt.data.aln:
        return []
    return tt.data.differences(node.up.cseq, node.cseq)


Clade.cseq = property(_compressed_sequence)
Clade.mutations = property(_mutations)
Clade.sequence = property(lambda x: x.tt.sequence(x, as_string=False))


class TreeAnc:
    """Tree + alignment + GTR model with GPU ancestral reconstruction."""

    def __init__(self, tree=None, aln=None, gtr=None, fill_overhangs=True, ref=None, verbose=ttconf.VERBOSE,
                 ignore_gaps=True, convert_upper=True, seq_multiplicity=None, log=None, compress=True, seq_len=None,
                 ignore_missing_alns=False, keep_node_order=False, rng_seed=None, engine=None, **kwargs):
        if tree is None:
            raise TypeError("TreeAnc requires a tree!")
        self.t_start = time.time()
        self.verbose = verbose
        self.log = log
        self.ok = False
        self.data = None
        self.log_messages = set()
        self.use_mutation_length = False
        self.ignore_gaps = ignore_gaps
        self.reconstructed_tip_sequences = False
        self.sequence_reconstruction = None
        self.ignore_missing_alns = ignore_missing_alns
        self.keep_node_order = keep_node_order
        self.rng = np.random.default_rng(seed=rng_seed)
        self.engine = engine
        self._device = None
        self._tree = None
        self.tree = tree
        self._gtr = None
        self.set_gtr(gtr or "JC69", **kwargs)
        self.data = SequenceData(aln, ref=ref, logger=self.logger, compress=compress, convert_upper=convert_upper,
                                 fill_overhangs=fill_overhangs, ambiguous=self.gtr.ambiguous, sequence_length=seq_len,
                                 seq_multiplicity=seq_multiplicity, engine=engine)
        if self.gtr.is_site_specific and self.data.compress:
            raise TypeError("TreeAnc: sequence compression and site specific gtr models are incompatible!")
        if self.data.aln and self.tree:
            self._check_alignment_tree_gtr_consistency()

    # ------------------------------------------------------------------ setup
    def logger(self, msg, level, warn=False, only_once=False):
        if only_once and msg in self.log_messages:
            return
        self.log_messages.add(msg)
        if level < self.verbose or (warn and level <= self.verbose):
            dt = time.time() - self.t_start
            print(("\n" if level < 2 else "") + format(dt, "4.2f") + "\t" + level * "-" + msg, file=sys.stdout)

    @property
    def gtr(self):
        return self._gtr

    @gtr.setter
    def gtr(self, value):
        if not isinstance(value, GTR):
            raise TypeError("GTR instance expected")
        self._gtr = value
        self._drop_device()

    def set_gtr(self, in_gtr, **kwargs):
        if isinstance(in_gtr, str):
            self._gtr = GTR.standard(model=in_gtr, **kwargs)
        elif isinstance(in_gtr, GTR):
            self._gtr = in_gtr
        else:
            raise TypeError("Cannot set GTR model in TreeAnc class: GTR or string expected")
        self._gtr.logger = self.logger
        self._drop_device()

    @property
    def aln(self):
        return self.data.aln

    @aln.setter
    def aln(self, in_aln):
        self.data.aln = in_aln
        self._drop_device()
        if self.tree:
            self._check_alignment_tree_gtr_consistency()

    @property
    def tree(self):
        return self._tree

    @tree.setter
    def tree(self, in_tree):
        """treeanc.py:316-371: accepts a Tree (or Bio.Phylo tree) or a Newick
        file name / string."""
        import os

        if isinstance(in_tree, str):
            if os.path.isfile(in_tree):
                with open(in_tree) as f:
                    in_tree = f.read()
            try:
                in_tree = Tree.from_newick(in_tree)
            except Exception as e:
                raise MissingDataError("TreeAnc: could not load tree! input was " + str(in_tree)[:80]) from e
        if not hasattr(in_tree, "root"):
            raise MissingDataError("TreeAnc: could not load tree! input was " + str(in_tree))
        self._tree = in_tree
        if self._tree.count_terminals() < 3:
            raise MissingDataError("TreeAnc: tree in %s as only %d tips. Please check your tree!"
                                   % (str(in_tree), self._tree.count_terminals()))
        long_branch = False
        for node in self._tree.find_clades():
            node.branch_length = node.branch_length if node.branch_length else 0.0
            long_branch |= node.branch_length > ttconf.MAX_BRANCH_LENGTH
            if hasattr(node, "_cseq"):
                del node._cseq
            node.original_length = node.branch_length
            node.mutation_length = node.branch_length
        if long_branch:
            self.logger("WARNING: TreeTime has detected branches that are longer than %d." % ttconf.MAX_BRANCH_LENGTH,
                        0, warn=True)
        self.prepare_tree()
        if self.data:
            self._check_alignment_tree_gtr_consistency()

    def prepare_tree(self):
        """treeanc.py:446-459."""
        self.sequence_reconstruction = False
        self.tree.root.branch_length = 0.001
        self.tree.root.mask = None
        self.tree.root.mutation_length = self.tree.root.branch_length
        if not self.keep_node_order:
            self.tree.ladderize()
        self._prepare_nodes()
        self._leaves_lookup = {node.name: node for node in self.tree.get_terminals()}
        self._drop_device()

    def _prepare_nodes(self):
        """treeanc.py:461-493: NODE_%07d names, up links, masks."""
        self.tree.root.up = None
        self.tree.root.tt = self
        names = {n.name for n in self.tree.find_clades() if n.name}
        count = 0
        for clade in self.tree.get_nonterminals(order="preorder"):
            if clade.name is None:
                tmp = "NODE_" + format(count, "07d")
                while tmp in names:
                    count += 1
                    tmp = "NODE_" + format(count, "07d")
                clade.name = tmp
                names.add(tmp)
            count += 1
            for c in clade.clades:
                c.up = clade
                c.tt = self
        for clade in self.tree.find_clades(order="postorder"):
            if not hasattr(clade, "mask"):
                clade.mask = None
        self.tree.root.dist2root = 0.0
        for clade in self.tree.get_nonterminals(order="preorder"):
            for c in clade.clades:
                c.dist2root = clade.dist2root + c.mutation_length

    @property
    def leaves_lookup(self):
        return self._leaves_lookup

    @property
    def one_mutation(self):
        return 1.0 / self.data.full_length if self.data.full_length else np.nan

    @property
    def seq_len(self):
        return self.data.full_length

    @property
    def sequence_length(self):
        return self.data.full_length

    def _check_alignment_tree_gtr_consistency(self):
        """treeanc.py:395-444."""
        failed = 0
        for leaf in self.tree.get_terminals():
            leaf.count = self.data.seq_multiplicity.get(leaf.name, 1.0)
            if leaf.name not in self.data.compressed_alignment:
                self.logger("***WARNING: NO SEQUENCE FOR LEAF: '%s'" % leaf.name, 0, warn=True)
                failed += 1
                if not self.ignore_missing_alns and failed > self.tree.count_terminals() / 3:
                    raise MissingDataError("TreeAnc._check_alignment_tree_gtr_consistency: At least 30\\% terminal "
                                           "nodes cannot be assigned a sequence!")
        extend_profile(self.gtr, [self.data.ref] if self.data.is_sparse else self.data.aln.values(),
                       logger=self.logger)
        self.ok = True
        self._drop_device()

    # --------------------------------------------------------- device mirror
    def _drop_device(self):
        if getattr(self, "_device", None) is not None:
            self._device.close()
        self._device = None
        self._nodes = None

    def _branch_length_to_gtr(self, node):
        """treeanc.py:752-760."""
        bl = node.mutation_length if self.use_mutation_length else node.branch_length
        return max(ttconf.MIN_BRANCH_LENGTH * self.one_mutation, bl)

    def _device_tree(self):
        """Flatten the tree (preorder numbering, children in list order) and
        make it resident on the device, or refresh its branch lengths."""
        nodes = list(self.tree.find_clades(order="preorder"))
        t = np.array([self._branch_length_to_gtr(n) for n in nodes])
        if self._device is not None and self._nodes is not None and len(self._nodes) == len(nodes) and \
                all(a is b for a, b in zip(self._nodes, nodes)):
            changed = np.nonzero(t != self._t)[0]
            if len(changed):
                self._device.set_branch_lengths(changed.astype(np.int32), t[changed])
                self._t = t
            return self._device
        self._drop_device()
        idx = {id(n): i for i, n in enumerate(nodes)}
        parent = np.full(len(nodes), -1, dtype=np.int32)
        for i, n in enumerate(nodes):
            for c in n.clades:
                parent[idx[id(c)]] = i
        tips = [n for n in nodes if not n.clades]
        P = self.data.compressed_length
        row = {k: i for i, k in enumerate(self.data.sequence_names)}
        codes = np.full((len(tips), P), MISSING_CODE, dtype=np.uint8)
        for k, leaf in enumerate(tips):
            if leaf.name in row:
                codes[k] = self.data.compressed_codes[row[leaf.name]]
        g = self.gtr
        eng = self.engine or default_engine()
        if g.is_site_specific:
            if g.seq_len != P:
                raise ValueError("site-specific model has {} sites, the alignment {} columns".format(g.seq_len, P))
            self._device = DeviceTree(eng, parent, t, codes, code_lut(g.profile_map, g.n_states),
                                      self.data.multiplicity(), g.Pi.mean(axis=1), None, None, None, None,
                                      site=g.device_model())
        else:
            self._device = DeviceTree(eng, parent, t, codes, code_lut(g.profile_map, g.n_states),
                                      self.data.multiplicity(), g.Pi, g.eigenvals, g.v, g.v_inv, g.mu)
        self._nodes = nodes
        self._t = t
        self._gtr_state = (np.copy(g.mu), g.Pi.copy(), g.W.copy())
        return self._device

    def _sync_model(self, dev):
        g = self.gtr
        if not (np.array_equal(self._gtr_state[0], g.mu) and np.array_equal(self._gtr_state[1], g.Pi) and
                np.array_equal(self._gtr_state[2], g.W)):
            self._drop_device()
            return self._device_tree()
        return dev

    # ------------------------------------------------------- reconstruction
    def reconstruct_anc(self, *args, **kwargs):
        return self.infer_ancestral_sequences(*args, **kwargs)

    def infer_ancestral_sequences(self, method="probabilistic", infer_gtr=False, marginal=False,
                                  reconstruct_tip_states=False, **kwargs):
        """treeanc.py:516-570; returns the number of changed characters."""
        if not self.ok:
            raise MissingDataError("TreeAnc.infer_ancestral_sequences: ERROR, sequences or tree are missing")
        m = method.lower()
        if m in ["ml", "probabilistic"]:
            run = self._ml_anc_marginal if marginal else self._ml_anc_joint
        elif m in ["fitch", "parsimony"]:
            raise NotImplementedError("Fitch parsimony is not part of the GPU pruning engine")
        else:
            raise UnknownMethodError("Reconstruction method needs to be in ['ml', 'probabilistic', 'fitch', "
                                     "'parsimony'], got '{}'".format(method))
        if infer_gtr:
            self.infer_gtr(marginal=marginal, **kwargs)
        return run(reconstruct_tip_states=reconstruct_tip_states, **kwargs)

    def _assign_sequences(self, dev, reconstruct_tip_states, sample_root=False, sample_other=False, prof=None):
        """Write node._cseq from the device states (argmax, first maximum) in
        preorder; nodes flagged for sampling draw their states from their
        marginal profile with this TreeAnc's generator instead, one
        ``rng.random(P)`` per sampled node in the reference's order
        (seq_utils.py:232-276, treeanc.py:830-838, 918-930)."""
        states = dev.states()
        alpha = self.gtr.alphabet
        L = self.data.compressed_length
        n_diff = 0
        for i, node in enumerate(self._nodes):
            if node.is_terminal() and not reconstruct_tip_states:
                continue
            sample = sample_root if node.up is None else sample_other
            if sample:
                cum = np.cumsum(prof[i].astype(float), axis=1).T
                idx = np.argmax(cum >= self.rng.random(size=cum.shape[1]), axis=0)
            else:
                idx = states[i]
            seq = alpha[idx]
            if node.up is None:
                pass  # the root's characters are not counted in N_diff (treeanc.py:830-838)
            elif self.sequence_reconstruction:
                n_diff += int((seq != node.cseq).sum())
            else:
                n_diff += L
            node._cseq = seq
        return n_diff

    def _ml_anc_marginal(self, sample_from_profile=False, reconstruct_tip_states=False, debug=False, **kwargs):
        """treeanc.py:762-812."""
        if sample_from_profile == "root":
            sample_root, sample_other = True, False
        elif isinstance(sample_from_profile, bool):
            sample_root = sample_other = sample_from_profile
        else:
            raise ValueError("sample_from_profile must be True, False or 'root'")
        dev = self._sync_model(self._device_tree())
        debug = debug and not dev.site  # site-specific passes keep no node normalisers
        total = dev.marginal(reconstruct_tips=reconstruct_tip_states, keep_normalisers=debug)
        prof = dev.profiles()
        sub = dev.profiles(which=1)
        outg = dev.profiles(which=2)
        tree = self.tree
        tree.sequence_LH = dev.sequence_lh()
        tree.total_sequence_LH = total
        tree.sequence_marginal_LH = total
        for i, node in enumerate(self._nodes):
            node.marginal_outgroup_LH = outg[i].astype(float)  # treeanc.py:819, 895-899
            if node.is_terminal():
                if node.name in self.data.compressed_alignment:
                    node.marginal_subtree_LH = seq2prof(self.data.compressed_alignment[node.name],
                                                        self.gtr.profile_map)
                else:
                    node.marginal_subtree_LH = np.ones((self.data.compressed_length, self.gtr.n_states))
                if reconstruct_tip_states:
                    node.marginal_profile = prof[i].astype(float)
            else:
                node.marginal_subtree_LH = sub[i].astype(float)
                node.marginal_profile = prof[i].astype(float)
        if debug:  # treeanc.py:803-810 deletes these unless debug
            logm = dev.profiles(which=3)
            pref = dev.subtree_prefactor()
            for i, node in enumerate(self._nodes):
                node.marginal_subtree_LH_prefactor = pref[i]
                if node.up is not None:
                    node.marginal_log_Lx = logm[i].astype(float)
        n_diff = self._assign_sequences(dev, reconstruct_tip_states, sample_root, sample_other, prof)
        self.reconstructed_tip_sequences = reconstruct_tip_states
        self.sequence_reconstruction = "marginal"
        return n_diff

    def _ml_anc_joint(self, sample_from_profile=False, reconstruct_tip_states=False, debug=False, **kwargs):
        """treeanc.py:934-1080.  With sample_from_profile ('root' or True) the
        root state is drawn from the root's joint profile with this TreeAnc's
        generator (one rng.random(P), treeanc.py:1022-1033) and every other
        node follows the joint argmax given it."""
        if sample_from_profile == "root":
            sample_root = True
        elif isinstance(sample_from_profile, bool):
            sample_root = sample_from_profile
        else:
            raise ValueError("sample_from_profile must be True, False or 'root'")
        dev = self._sync_model(self._device_tree())
        total = dev.joint(reconstruct_tips=reconstruct_tip_states)
        if sample_root:
            # prof2seq(np.exp(normalized_profile), sample_from_prof=True): rows
            # normalised to sum 1, cumulative sums against rng.random(P)
            prof = np.exp(dev.joint_root())
            prof = np.einsum("ai,a->ai", prof, 1.0 / prof.sum(axis=1))
            cum = prof.cumsum(axis=1).T
            idx = np.argmax(cum >= self.rng.random(size=cum.shape[1]), axis=0)
            total = dev.joint_from_root(idx.astype(np.uint8), reconstruct_tips=reconstruct_tip_states)
        self.tree.sequence_LH = dev.sequence_lh()
        self.tree.sequence_joint_LH = total
        if debug:  # treeanc.py:1066-1077 deletes these unless debug
            lx, cx = dev.joint_debug()
            states = dev.states()
            for i, node in enumerate(self._nodes):
                node.joint_Lx = lx[i]
                node.joint_Cx = None if node.up is None else cx[i].astype(np.uint16)
                if node.up is None or not node.is_terminal() or reconstruct_tip_states:
                    node.seq_idx = states[i].astype(int)
        n_diff = self._assign_sequences(dev, reconstruct_tip_states)
        self.reconstructed_tip_sequences = reconstruct_tip_states
        self.sequence_reconstruction = "joint"
        return n_diff

    # ---------------------------------------------------------- likelihood
    def sequence_LH(self, pos=None, full_sequence=False):
        """treeanc.py:691-716."""
        if not hasattr(self.tree, "total_sequence_LH"):
            self.logger("TreeAnc.sequence_LH: you need to run marginal ancestral inference first!", 1)
            self.infer_ancestral_sequences(marginal=True)
        if pos is not None:
            cpos = self.data.full_to_compressed_sequence_map[pos] if full_sequence else pos
            return self.tree.sequence_LH[cpos]
        return self.tree.total_sequence_LH

    def ancestral_likelihood(self):
        """treeanc.py:718-750: log-likelihood of the current sequences."""
        log_lh = np.zeros(self.data.multiplicity().shape[0])
        idx = self.gtr.state_index
        for node in self.tree.find_clades(order="postorder"):
            if node.up is None:
                prof = seq2prof(node.cseq, self.gtr.profile_map) * self.gtr.Pi
                log_lh += np.log(prof.sum(axis=1))
                continue
            pairs = np.array([(idx[a], idx[b]) for a, b in zip(node.up.cseq, node.cseq)])
            lq = np.log(self.gtr.expQt(node.branch_length))
            log_lh += lq[pairs[:, 1], pairs[:, 0]]
        return log_lh

    # ------------------------------------------------------ branch lengths
    def optimize_branch_len(self, **kwargs):
        return self.optimize_branch_lengths_joint(**kwargs)

    def optimize_branch_len_joint(self, **kwargs):
        return self.optimize_branch_lengths_joint(**kwargs)

    def optimize_branch_lengths_joint(self, **kwargs):
        """treeanc.py:1176-1242: one sweep over all branches, each set to the
        maximum-likelihood length given the states of the current
        reconstruction at its ends.  All branches are optimised together on
        the device (tt_optimize_branch_lengths); the results are applied in
        postorder as the reference does, so an error at one branch leaves the
        branches before it updated."""
        if self.tree is None or self.data.aln is None:
            raise MissingDataError("TreeAnc.optimize_branch_length: ERROR, alignment or tree are missing.")
        store_old = kwargs.get("store_old", False)
        res = self._branch_optima()
        max_bl = 0
        for node in self.tree.find_clades(order="postorder"):
            if node.up is None:
                continue
            if store_old:
                node._old_length = node.branch_length
            new_len = max(0, self._take_optimum(node, res))
            node.branch_length = new_len
            node.mutation_length = new_len
            max_bl = max(max_bl, new_len)
        if max_bl > 0.15:
            self.logger("TreeAnc.optimize_branch_lengths_joint: THIS TREE HAS LONG BRANCHES.", 0, warn=True)
        self.tree.root.up = None
        self.tree.root.dist2root = 0.0
        self._prepare_nodes()
        return ttconf.SUCCESS

    def _code_states(self, skip_unknown=False):
        """Alphabet index of every character code for gtr.py:672-712: characters
        outside the alphabet are not counted for n < 10 and count as state 1
        for larger alphabets (the reference's np.ones_like default);
        ``skip_unknown`` never counts them (infer_gtr's state_index lookup,
        treeanc.py:1574-1578)."""
        alpha = [str(a) for a in self.gtr.alphabet]
        other = -1 if (len(alpha) < 10 or skip_unknown) else 1
        out = np.full(256, other, dtype=np.int16)
        for i, a in enumerate(alpha):
            if len(a) == 1 and ord(a) < 256:
                out[ord(a)] = i
        return out

    def _branch_optima(self):
        """Device sweep: optimal length, status and state-pair counts of every
        branch, indexed by node."""
        dev = self._device
        if dev is None or self._nodes is None or not self.sequence_reconstruction:
            raise MissingDataError("TreeAnc.optimize_branch_length: ancestral sequences must be reconstructed "
                                   "first (infer_ancestral_sequences)")
        nodes = list(self.tree.find_clades(order="preorder"))
        if len(nodes) != len(self._nodes) or any(a is not b for a, b in zip(nodes, self._nodes)):
            raise MissingDataError("TreeAnc.optimize_branch_length: the tree changed since the last "
                                   "reconstruction; reconstruct ancestral sequences again")
        g = self.gtr
        gap = g.gap_index if (self.ignore_gaps and g.gap_index is not None) else -1
        t, st, cnt = dev.optimize_branch_lengths(self._code_states(), gap, ttconf.MAX_BRANCH_LENGTH, 1e-10)
        return {id(n): (t[i], st[i], cnt[i]) for i, n in enumerate(nodes)}

    def _take_optimum(self, node, res):
        """treeanc.py:1244-1270 for one node, from the device sweep."""
        if node.cseq is None and node.is_terminal():
            raise MissingDataError("TreeAnc.optimal_branch_length: terminal node alignments required; sequence is "
                                   "missing for leaf: '%s'." % node.name)
        t, st, cnt = res[id(node)]
        nz = np.argwhere(cnt != 0)
        node.branch_state = {"pair": nz.astype(int), "multiplicity": cnt[cnt != 0].astype(int)}
        if st == 1:
            raise ValueError("Bracketing values (xa, xb, xc) do not fulfill this requirement: "
                             "(f(xb) < f(xa)) and (f(xb) < f(xc))")
        if t > 0.9 * ttconf.MAX_BRANCH_LENGTH:
            self.logger("WARNING: GTR.optimal_t_compressed -- The branch length seems to be very long!", 4, warn=True)
        return float(t)

    def optimal_branch_length(self, node):
        """treeanc.py:1244-1270 (one branch; the sweep above does all at once)."""
        if node.up is None:
            return self.one_mutation
        return max(0.0, self._take_optimum(node, self._branch_optima()))

    # ----------------------------------------------------------- GTR inference
    def infer_gtr(self, marginal=False, site_specific=False, normalized_rate=True, fixed_pi=None, pc=5.0, **kwargs):
        """treeanc.py:1500-1632: transition counts n_ij and state times T_i
        of the current reconstruction (computed on the device for every
        branch at once), then GTR.infer on the host.  Marginal: expected
        counts from each branch's profile pair (tt_mutation_counts);
        joint: the (parent, child) state counts of every branch
        (tt_branch_state_counts).  site_specific: per-site statistics summed
        over the branches (tt_site_counts), then GTR_site_specific.infer."""
        if site_specific and self.data.compress:
            raise TypeError("TreeAnc.infer_gtr(): sequence compression and site specific GTR models are "
                            "incompatible!")
        if not self.ok:
            raise MissingDataError("TreeAnc.infer_gtr: ERROR, sequences or tree are missing")
        if marginal and self.sequence_reconstruction != "marginal":
            self._ml_anc_marginal(**kwargs)
        elif not self.sequence_reconstruction:
            self._ml_anc_joint(**kwargs)
        g = self.gtr
        n = g.n_states
        dev = self._device
        mult = self.data.multiplicity()
        self.logger("TreeAnc.infer_gtr: counting mutations...", 2)
        if site_specific:
            root_idx = self._nodes.index(self.tree.root)
            if marginal:
                n_pij, t_pi = dev.site_counts(True)
                root_state = dev.profiles()[root_idx].astype(float).T
            else:
                n_pij, t_pi = dev.site_counts(False, self._code_states(skip_unknown=True))
                root_state = np.eye(n)[:, dev.states()[root_idx]]
            self._gtr = GTR_site_specific.infer(np.transpose(n_pij, (1, 2, 0)), t_pi.T, pc=pc, root_state=root_state,
                                                logger=self.logger, alphabet=g.alphabet, prof_map=g.profile_map)
            if normalized_rate:
                self.logger("TreeAnc.infer_gtr: setting overall rate to 1.0...", 2)
                self._gtr.mu /= self._gtr.average_rate().mean()
            return self._gtr
        if marginal:
            nij, Ti = dev.mutation_counts()
        else:
            t = np.array([self._branch_length_to_gtr(x) if x.up is not None else 0.0 for x in self._nodes])
            C = dev.branch_state_counts(self._code_states(skip_unknown=True), -1)  # C[node][parent j][child i]
            nij = np.zeros((n, n))
            Ti = np.zeros(n)
            off = ~np.eye(n, dtype=bool)
            for k, x in enumerate(self._nodes):
                if x.up is None:
                    continue
                c = C[k]
                nij += (c * off).T                      # i derived (child), j ancestral (parent)
                Ti += t[k] * c.sum(axis=0)              # time in the child's state
                Ti += 0.5 * t[k] * (c * off).sum(axis=1)  # mutations: half the branch in the parent state
                Ti -= 0.5 * t[k] * (c * off).sum(axis=0)  # ... and not in the child state
        root_idx = self._nodes.index(self.tree.root)
        root_states = dev.states()[root_idx]
        root_state = np.bincount(root_states, weights=mult, minlength=n)[:n]
        self._gtr = GTR.infer(nij, Ti, root_state, fixed_pi=fixed_pi, pc=pc, alphabet=g.alphabet,
                              logger=self.logger, prof_map=g.profile_map)
        if normalized_rate:
            self.logger("TreeAnc.infer_gtr: setting overall rate to 1.0...", 2)
            self._gtr.mu = 1.0
        return self._gtr

    # ------------------------------------------------- marginal branch lengths
    def _marginal_optima(self, tol):
        """Device sweep over all branches (tt_optimize_branch_lengths_marginal)
        with the profiles of the last marginal pass: {id(node): (t, status)}
        and (t, status) of a bifurcating root's pair of children."""
        dev = self._device
        if dev is None or self._nodes is None or self.sequence_reconstruction != "marginal":
            raise MissingDataError("marginal ancestral inference needs to be performed first!")
        nodes = list(self.tree.find_clades(order="preorder"))
        if len(nodes) != len(self._nodes) or any(a is not b for a, b in zip(nodes, self._nodes)):
            raise MissingDataError("the tree changed since the last marginal reconstruction")
        g = self.gtr
        gap = g.gap_index if g.gap_index is not None else -1
        t, st, pair = dev.optimize_branch_lengths_marginal(gap, ttconf.MAX_BRANCH_LENGTH, tol)
        return {id(n): (float(t[i]), int(st[i])) for i, n in enumerate(nodes)}, pair

    @staticmethod
    def _checked(t, st):
        if st == 1:
            raise ValueError("Bracketing values (xa, xb, xc) do not fulfill this requirement: "
                             "(f(xb) < f(xa)) and (f(xb) < f(xc))")
        return t

    def optimal_marginal_branch_length(self, node, tol=1e-10):
        """treeanc.py:1272-1295 (the device computes every branch at once)."""
        if node.up is None:
            return self.one_mutation
        res, _ = self._marginal_optima(tol)
        return self._checked(*res[id(node)])

    def optimize_tree_marginal(self, max_iter=10, infer_gtr=False, pc=1.0, damping=0.75, LHtol=0.1,
                               site_specific_gtr=False, **kwargs):
        """treeanc.py:1297-1358: iterate (all branch optima from the current
        marginal profiles, damped update, marginal reconstruction) until the
        likelihood gain drops below LHtol."""
        self.infer_ancestral_sequences(marginal=True, **kwargs)
        oldLH = self.sequence_LH()
        self.logger("TreeAnc.optimize_tree_marginal: initial, LH=%1.2f, total branch_length %1.4f"
                    % (oldLH, self.tree.total_branch_length()), 2)
        for i in range(max_iter):
            if infer_gtr:  # treeanc.py:1308-1310
                self.infer_gtr(site_specific=site_specific_gtr, marginal=True, normalized_rate=True, pc=pc)
                self.infer_ancestral_sequences(marginal=True, **kwargs)
            old_bl = self.tree.total_branch_length()
            tol = 1e-8 + 0.01 ** (i + 1)
            res, pair = self._marginal_optima(tol)
            keep = damping ** (i + 1)
            for n in self.tree.find_clades():
                if n.up is None:
                    continue
                if n.up.up is None and len(n.up.clades) == 2:
                    # children of a bifurcating root (visited twice, as in the reference)
                    n1, n2 = n.up.clades
                    total_bl = n1.branch_length + n2.branch_length
                    bl_ratio = n1.branch_length / total_bl
                    new_bl = self._checked(*pair)
                    update_val = new_bl * (1 - keep) + total_bl * keep
                    n1.branch_length = update_val * bl_ratio
                    n2.branch_length = update_val * (1 - bl_ratio)
                    n1.mutation_length = n1.branch_length
                    n2.mutation_length = n2.branch_length
                else:
                    new_val = self._checked(*res[id(n)])
                    n.branch_length = new_val * (1 - keep) + n.branch_length * keep
                    n.mutation_length = n.branch_length
            self.infer_ancestral_sequences(marginal=True, **kwargs)
            LH = self.sequence_LH()
            deltaLH = LH - oldLH
            oldLH = LH
            dbl = self.tree.total_branch_length() - old_bl
            self.logger("TreeAnc.optimize_tree_marginal: iteration %d, LH=%1.2f (%1.2f), delta branch_length=%1.4f, "
                        "total branch_length %1.4f" % (i, LH, deltaLH, dbl, self.tree.total_branch_length()), 2)
            if deltaLH < LHtol:
                self.logger("TreeAnc.optimize_tree_marginal: deltaLH=%f, stopping iteration." % deltaLH, 1)
                break
        return ttconf.SUCCESS

    def optimize_sequences_and_branch_length(self, *args, **kwargs):
        """Deprecated alias of optimize_tree (treeanc.py:1362-1371)."""
        self.logger("Deprecation warning: 'optimize_sequences_and_branch_length' will be removed and replaced by "
                    "'optimize_tree'!", 1, warn=True)
        return self.optimize_tree(*args, **kwargs)

    def optimize_seq_and_branch_len(self, *args, **kwargs):
        """Deprecated alias of optimize_tree (treeanc.py:1373-1382)."""
        self.logger("Deprecation warning: 'optimize_seq_and_branch_len' will be removed and replaced by "
                    "'optimize_tree'!", 1, warn=True)
        return self.optimize_tree(*args, **kwargs)

    def optimize_tree(self, prune_short=True, marginal_sequences=False, branch_length_mode="joint", max_iter=5,
                      infer_gtr=False, pc=1.0, method_anc="probabilistic", **kwargs):
        """treeanc.py:1384-1473: alternate ancestral reconstruction and
        joint branch-length optimisation (each a device pass) until no
        character changes, pruning zero-length internal branches between
        iterations; 'marginal' runs optimize_tree_marginal, 'input' only
        reconstructs."""
        if branch_length_mode == "marginal":
            self.optimize_tree_marginal(max_iter=max_iter, infer_gtr=infer_gtr, pc=pc, **kwargs)
            if prune_short:
                self.prune_short_branches()
            return ttconf.SUCCESS
        if branch_length_mode == "input":
            self.reconstruct_anc(method=method_anc, infer_gtr=infer_gtr, pc=pc, marginal=marginal_sequences,
                                 **kwargs)
            if prune_short:
                self.prune_short_branches()
            return ttconf.SUCCESS
        if branch_length_mode != "joint":
            raise UnknownMethodError("TreeAnc.optimize_tree: `branch_length_mode` should be in ['marginal', "
                                     "'joint', 'input']")
        self.logger("TreeAnc.optimize_tree: sequences...", 1)
        self.reconstruct_anc(method=method_anc, infer_gtr=infer_gtr, pc=pc, marginal=marginal_sequences, **kwargs)
        self.optimize_branch_lengths_joint(store_old=False)
        n = 0
        while n < max_iter:
            n += 1
            if prune_short:
                self.prune_short_branches()
            n_diff = self.reconstruct_anc(method=method_anc, infer_gtr=False, marginal=marginal_sequences,
                                          **kwargs)
            self.logger("TreeAnc.optimize_tree: Iteration %d. #Nuc changed since prev reconstructions: %d"
                        % (n, n_diff), 2)
            if n_diff < 1:
                break
            self.optimize_branch_lengths_joint(store_old=False)
        self.tree.unconstrained_sequence_LH = float((self.tree.sequence_LH * self.data.multiplicity()).sum())
        self._prepare_nodes()
        self.logger("TreeAnc.optimize_tree: Unconstrained sequence LH:%f" % self.tree.unconstrained_sequence_LH, 2)
        return ttconf.SUCCESS

    def prune_short_branches(self):
        """treeanc.py:1475-1495: an internal branch shorter than a tenth of
        one mutation whose two ends carry the same sequence (probability
        above 0.1 at length zero, gaps ignored) is removed; its children
        move to its parent."""
        self.logger("TreeAnc.prune_short_branches: pruning short branches (max prob at zero)...", 1)
        mult = self.data.multiplicity()
        for node in self.tree.find_clades():
            if node.up is None or node.is_terminal():
                continue
            if node.branch_length < 0.1 * self.one_mutation and \
                    self.gtr.prob_t(node.up._cseq, node._cseq, 0.0, pattern_multiplicity=mult) > 0.1:
                node.up.clades = [k for k in node.up.clades if k is not node] + node.clades
                for clade in node.clades:
                    clade.up = node.up

    def marginal_branch_profile(self, node):
        """treeanc.py:1122-1146: (marginal_outgroup_LH, marginal_subtree_LH)
        of the branch above ``node``."""
        if node.up is None:
            raise Exception("Branch profiles can't be calculated for the root!")
        if not hasattr(node, "marginal_outgroup_LH"):
            raise Exception("marginal ancestral inference needs to be performed first!")
        return node.marginal_outgroup_LH, node.marginal_subtree_LH

    def get_branch_mutation_matrix(self, node, full_sequence=False):
        """treeanc.py:1085-1120: per pattern, the joint distribution of the
        states at both ends of the branch above ``node`` (child state i,
        parent state j), computed on the device from the last marginal pass."""
        self.marginal_branch_profile(node)
        if self._device is None or self.sequence_reconstruction != "marginal":
            raise Exception("marginal ancestral inference needs to be performed first!")
        idx = next(i for i, x in enumerate(self._nodes) if x is node)
        m = self._device.branch_mutation_matrix(idx)
        if full_sequence:
            return m[self.data.full_to_compressed_sequence_map]
        return m

    def branch_profile_lh(self, grid):
        """prob_t_profiles(return_log=True) of every branch on the lengths
        ``grid`` (the marginal-mode BranchLenInterpolator grid,
        branch_len_interpolator.py:104-110), from the last marginal pass:
        {node: array(len(grid))}."""
        if self._device is None or self.sequence_reconstruction != "marginal":
            raise MissingDataError("marginal ancestral inference needs to be performed first!")
        g = self.gtr
        out = self._device.branch_profile_lh(np.asarray(grid, dtype=float), g.gap_index if g.gap_index is not None
                                             else -1)
        return {n: out[i] for i, n in enumerate(self._nodes) if n.up is not None}

    # ---------------------------------------------------------------- output
    def sequence(self, node, as_string=True):
        seq = node.cseq
        if seq is None:
            return None
        return self.data.compressed_to_full_sequence(seq, as_string=as_string)

    def get_reconstructed_alignment(self, reconstruct_tip_states=False):
        """Full-length sequences of every node as a dict name -> str; for a
        sparse alignment the reference's dict of 'sequences' (differences
        from the reference per node), 'reference', 'positions' and
        'inferred_const_sites' (treeanc.py:1716-1760)."""
        if not self.sequence_reconstruction:
            self.infer_ancestral_sequences("probabilistic", marginal=True, reconstruct_tip_states=reconstruct_tip_states)
        if self.data.is_sparse:
            return {"sequences": {n.name: self.data.compressed_to_sparse_sequence(n.cseq)
                                  for n in self.tree.find_clades() if n.cseq is not None},
                    "reference": self.data.ref, "positions": self.data.nonref_positions,
                    "inferred_const_sites": self.data.inferred_const_sites}
        return {n.name: self.sequence(n, as_string=True) for n in self.tree.find_clades()
                if n.cseq is not None}

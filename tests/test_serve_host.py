"""CPU tests of the serving loop's scheduling with a fake device: per-tree
order, slot limit, overlap, memory admission, rejection and failure
isolation."""
import numpy as np

from treetime_amd.serve import Request, Server, make_stream, summarize


class FakeEngine:
    def __init__(self, budget):
        self.used = 0
        self.budget = budget

    def usage(self):
        return self.used, self.budget


class FakeTree:
    """Completes each request after ``delay`` polls; total() is a function of
    the request history so ordering mistakes change results."""

    def __init__(self, eng, tid, n_nodes=9, P=10, delay=2, jbytes=100):
        self.eng, self.tid = eng, tid
        self.n_nodes, self.P = n_nodes, P
        self.parent = np.array([-1] + [0] * (n_nodes - 1))
        self.t_cur = np.linspace(0.01, 0.02, n_nodes)
        self.delay, self.jbytes = delay, jbytes
        self.busy = 0
        self.value = float(tid)
        self.joint_ready = False
        self.log = []
        self.max_concurrent_seen = 0

    def nonroot_nodes(self):
        return np.arange(1, self.n_nodes)

    def path_len(self, node):
        return 1

    def _start(self, tag):
        assert self.busy == 0, "a tree ran two requests at once"
        self.busy = self.delay
        self.log.append(tag)

    def update_branch(self, node, t, want_logL=True):
        self._start(("branch", node))
        self.value = self.value * 1.5 + t

    def set_branch_lengths(self, nodes, t):
        self.t_cur[nodes] = t

    def loglik(self, want_logL=True):
        self._start(("loglik",))
        self.value = self.value * 1.5 + float(self.t_cur.sum())

    def marginal(self, want_logL=True):
        self._start(("marginal",))
        self.value += 1.0

    def joint(self, want_logL=True):
        if not self.joint_ready:
            if self.eng.used + self.jbytes > self.eng.budget:
                raise MemoryError("over budget")
            self.eng.used += self.jbytes
            self.joint_ready = True
        self._start(("joint",))
        self.value += 2.0

    def joint_bytes(self):
        return self.jbytes

    def trim(self):
        if self.joint_ready:
            self.eng.used -= self.jbytes
            self.joint_ready = False

    def query(self):
        if self.busy:
            self.busy -= 1
        return self.busy == 0

    def total(self):
        return self.value


def _run(trees, reqs, **kw):
    return Server(trees, **kw).run([Request(**r.__dict__) for r in reqs], realtime=False)


def test_results_same_alone_and_in_mix():
    eng = FakeEngine(10**9)
    rng = np.random.default_rng(0)
    trees = [FakeTree(eng, i) for i in range(6)]
    reqs = make_stream(trees, 80, 1000.0, rng)
    mixed = _run(trees, reqs, engine=eng, slots=4)
    trees2 = [FakeTree(eng, i) for i in range(6)]
    alone = _run(trees2, reqs, engine=eng, slots=1)
    assert [r.result for r in mixed] == [r.result for r in alone]
    assert all(not r.error for r in mixed)


def test_per_tree_fifo_and_slots():
    eng = FakeEngine(10**9)
    trees = [FakeTree(eng, i, delay=3) for i in range(3)]
    reqs = [Request("marginal", 0, rid=0), Request("branch", 0, node=2, scale=2.0, rid=1),
            Request("joint", 1, rid=2), Request("loglik", 0, rid=3), Request("marginal", 2, rid=4)]
    out = _run(trees, reqs, engine=eng, slots=2)
    assert trees[0].log == [("marginal",), ("branch", 2), ("loglik",)]
    assert all(not r.error for r in out)
    s = summarize(out, 1.0)
    assert s["requests"] == 5 and s["failed"] == 0


def test_memory_admission_and_rejection():
    eng = FakeEngine(250)
    trees = [FakeTree(eng, i, jbytes=100) for i in range(3)] + [FakeTree(eng, 3, jbytes=1000)]
    reqs = [Request("joint", i, rid=i) for i in range(3)] + [Request("joint", 3, rid=3),
                                                               Request("marginal", 3, rid=4)]
    out = _run(trees, reqs, engine=eng, slots=8, hbm_budget=250)
    by = {r.rid: r for r in out}
    assert not by[0].error and not by[1].error and not by[2].error   # admitted as memory frees up
    assert "budget" in by[3].error                                    # can never fit: fails alone
    assert not by[4].error                                            # later request of that tree still runs
    assert eng.used <= 250                                            # idle joint buffers were evicted


def test_failure_isolated():
    eng = FakeEngine(10**9)
    trees = [FakeTree(eng, i) for i in range(2)]
    reqs = [Request("bogus", 0, rid=0), Request("marginal", 1, rid=1), Request("marginal", 0, rid=2)]
    out = _run(trees, reqs, engine=eng, slots=2)
    by = {r.rid: r for r in out}
    assert by[0].error and not by[1].error and not by[2].error

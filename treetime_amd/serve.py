"""Request server over resident trees (continuous batching).

Requests: ``branch`` (scale one branch length, re-evaluate logL along its
root path only -- tt_update_branch), ``loglik`` (postorder + root),
``marginal`` (full marginal reconstruction) and ``joint`` (joint
reconstruction).  Trees, patterns and messages stay resident in HBM.

Scheduling:
* requests of one tree run in arrival order (a tree is busy while one of its
  requests is on the device; later requests of that tree wait);
* requests of different trees overlap on the device: each tree has its own
  HIP stream and passes are launched asynchronously; completion is polled;
* at most ``slots`` requests are in flight; a request that arrives while
  others run is admitted as soon as a slot and its memory are free;
* memory admission: a joint request allocates the tree's joint buffers on
  first use; it is admitted only if the engine's budget has room.  When it
  has not, joint buffers of idle trees are returned (``trim``), least
  recently used first; if that is still not enough the request waits for
  running requests, and a request that can never fit fails alone while the
  others proceed.

``scratch=True`` recomputes a branch request with a full postorder instead
of the root path (the baseline the root-path update is measured against).

``device_batch=B`` (B >= 2) fuses, at each scheduling step, the admitted
``marginal`` / ``joint`` / ``loglik`` requests of one kind into one batched
launch over their trees (``tt_batch_run_members`` through the pool's tuned
plan) when there are at least B of them; fewer run as per-tree passes.
Results are the same either way (the batched kernels equal the per-tree
passes).  The pool's plans are built and tuned when the server starts.
"""
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    kind: str                 # "branch" | "loglik" | "marginal" | "joint"
    tree: int
    arrival: float = 0.0      # seconds after the start of the stream
    node: int = -1            # branch requests: node whose branch changes
    scale: float = 1.0        # branch requests: factor on the current length
    rid: int = 0
    result: float = float("nan")
    error: str = ""
    t_admit: float = float("nan")
    t_done: float = float("nan")
    updates: int = 0          # pattern x branch updates this request performed
    meta: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.t_done - self.arrival


class Server:
    """Continuous-batching loop over ``trees`` (DeviceTree-like objects)."""

    def __init__(self, trees, engine=None, slots=16, hbm_budget=None, scratch=False, clock=time.perf_counter,
                 device_batch=0, batch_kinds=("marginal", "joint")):
        self.trees = list(trees)
        self.engine = engine
        self.slots = slots
        self.hbm_budget = hbm_budget
        self.scratch = scratch
        self.clock = clock
        self.last_use = [0.0] * len(self.trees)
        self.device_batch = device_batch if device_batch and device_batch >= 2 else 0
        self.batches = 0          # batched launches made
        if self.device_batch:
            from .engine import batch_run_members
            self._members = batch_run_members
            for kind in batch_kinds:  # build and tune the pool's plan of each kind now
                batch_run_members(engine, self.trees, [], kind=kind)

    # -- memory ------------------------------------------------------------
    def _used(self):
        return self.engine.usage()[0] if self.engine is not None else 0

    def _need(self, req):
        tree = self.trees[req.tree]
        if req.kind == "joint" and not getattr(tree, "joint_ready", False):
            return tree.joint_bytes()
        return 0

    def _fits(self, req, in_flight_need):
        if self.hbm_budget is None:
            return True
        return self._used() + in_flight_need + self._need(req) <= self.hbm_budget

    def _evict(self, req, in_flight_need, inflight):
        """Trim joint buffers of idle trees (LRU) until ``req`` fits."""
        idle = [k for k, t in enumerate(self.trees)
                if getattr(t, "joint_ready", False) and k not in inflight and k != req.tree]
        for k in sorted(idle, key=lambda k: self.last_use[k]):
            if self._fits(req, in_flight_need):
                break
            self.trees[k].trim()
        return self._fits(req, in_flight_need)

    # -- launch / complete ---------------------------------------------------
    def _launch(self, req):
        tree = self.trees[req.tree]
        n_br = tree.n_nodes - 1
        if req.kind == "branch":
            t_new = float(tree.t_cur[req.node] * req.scale)
            tree.t_cur[req.node] = t_new
            if self.scratch:
                tree.set_branch_lengths(np.array([req.node], np.int32), np.array([t_new]))
                tree.loglik(want_logL=False)
                req.updates = tree.P * n_br
            else:
                tree.update_branch(req.node, t_new, want_logL=False)
                req.updates = tree.P * tree.path_len(req.node)
        elif req.kind == "loglik":
            tree.loglik(want_logL=False)
            req.updates = tree.P * n_br
        elif req.kind == "marginal":
            tree.marginal(want_logL=False)
            req.updates = tree.P * n_br
        elif req.kind == "joint":
            tree.joint(want_logL=False)
            req.updates = tree.P * n_br
        else:
            raise ValueError("unknown request kind %r" % req.kind)

    def _launch_batch(self, kind, reqs):
        """One batched launch for same-kind requests on distinct trees."""
        self._members(self.engine, self.trees, [r.tree for r in reqs], kind=kind)
        self.batches += 1
        for r in reqs:
            tree = self.trees[r.tree]
            r.updates = tree.P * (tree.n_nodes - 1)
            r.meta["batched"] = len(reqs)

    def _complete(self, req):
        req.result = self.trees[req.tree].total()

    # -- main loop -------------------------------------------------------------
    def run(self, requests, realtime=True):
        """Serve ``requests`` (sorted by arrival); returns them completed."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        pending = deque()
        inflight = {}          # tree -> request
        i = 0
        t0 = self.clock()
        while i < len(reqs) or pending or inflight:
            now = self.clock() - t0
            while i < len(reqs) and (reqs[i].arrival <= now or not realtime):
                pending.append(reqs[i])
                i += 1
            for tid in [t for t, r in inflight.items() if self.trees[t].query()]:
                r = inflight.pop(tid)
                try:
                    self._complete(r)
                except Exception as e:  # one failed request does not stop the others
                    r.error = repr(e)
                r.t_done = self.clock() - t0
            blocked = set(inflight)
            need_inflight = sum(self._need(r) for r in inflight.values())
            admitted = []
            for r in list(pending):
                if len(inflight) + len(admitted) >= self.slots:
                    break
                if r.tree in blocked:
                    continue
                if not self._fits(r, need_inflight) and not self._evict(r, need_inflight, inflight):
                    if not inflight and self._need(r) + self._used() > (self.hbm_budget or 0):
                        pending.remove(r)      # can never fit: fail this request alone
                        r.error = "rejected: exceeds the HBM budget"
                        r.t_done = self.clock() - t0
                        continue
                    blocked.add(r.tree)        # keep per-tree order behind it
                    continue
                pending.remove(r)
                r.t_admit = self.clock() - t0
                self.last_use[r.tree] = r.t_admit
                admitted.append(r)
                need_inflight += self._need(r)
                blocked.add(r.tree)
            groups = {}
            if self.device_batch:
                for r in admitted:
                    if r.kind in ("marginal", "joint", "loglik"):
                        groups.setdefault(r.kind, []).append(r)
                groups = {k: g for k, g in groups.items() if len(g) >= self.device_batch}
            fused = {id(r) for g in groups.values() for r in g}
            for kind, g in groups.items():
                try:
                    self._launch_batch(kind, g)
                    for r in g:
                        inflight[r.tree] = r
                except Exception as e:  # the batch fails alone
                    for r in g:
                        r.error = repr(e)
                        r.t_done = self.clock() - t0
            for r in admitted:
                if id(r) in fused:
                    continue
                try:
                    self._launch(r)
                    inflight[r.tree] = r
                except Exception as e:
                    r.error = repr(e)
                    r.t_done = self.clock() - t0
            if not inflight and pending and i >= len(reqs):
                # everything left is blocked by memory with nothing in flight
                for r in list(pending):
                    r.error = "rejected: exceeds the HBM budget"
                    r.t_done = self.clock() - t0
                    pending.remove(r)
        return reqs


def make_stream(trees, n_requests, rate, rng, mix=(0.7, 0.2, 0.1)):
    """Poisson arrivals at ``rate`` requests/s over the tree pool; kinds drawn
    with probabilities (branch, marginal, joint); branch scale LogNormal(0, 0.3)."""
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
    kinds = rng.choice(["branch", "marginal", "joint"], size=n_requests, p=list(mix))
    out = []
    for k in range(n_requests):
        tid = int(rng.integers(len(trees)))
        tree = trees[tid]
        node = int(rng.choice(tree.nonroot_nodes())) if kinds[k] == "branch" else -1
        out.append(Request(kind=str(kinds[k]), tree=tid, arrival=float(t[k]), node=node,
                           scale=float(rng.lognormal(0.0, 0.3)), rid=k))
    return out


def summarize(reqs, wall):
    ok = [r for r in reqs if not r.error]
    lat = np.array([r.latency for r in ok]) if ok else np.array([np.nan])
    return {
        "requests": len(reqs),
        "failed": len(reqs) - len(ok),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "throughput_req_s": len(ok) / wall if wall > 0 else float("nan"),
        "updates_per_s": sum(r.updates for r in ok) / wall if wall > 0 else float("nan"),
    }

#!/usr/bin/env python3
"""Throughput of marginal ancestral reconstruction on one MI355X.

Default workload: 256 independent nucleotide jobs (tips log-uniform in
[50, 4000], sites log-uniform in [1000, 30000], depth 0.02, 1% N / 0.5% gaps,
own seed per job), all resident in HBM, reconstructed in ONE persistent
launch per step (tt_batch_marginal).  A step = one full marginal pass
(postorder, root likelihood, preorder with argmax states) over all jobs.

Metric: pattern x branch updates per second, U = P * (nodes - 1) per job.
Printed beside it: the plain path (one tt_marginal graph per tree, trees in
sequence), the roofline of the batched kernel from the algorithmic bytes per
update, and the FP64 numpy oracle timed on the host as the CPU baseline.

Also measured: pattern compression of a device-resident 5000 x 30000
alignment (the `compression` object; alone with --workload compress), and
marginal / joint passes with a site-specific model (--workload site).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload mix|large-5000|compress|site|...]
"""
import argparse
import json
import multiprocessing as mp
import os
import platform
import sys
import time

# the CPU baseline times one BLAS thread per process: set before numpy loads
for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ.setdefault(_v, "1")

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD datasheet


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mix")
    ap.add_argument("--jobs", type=int, default=256)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true", help="skip the single-tree (drop-in) passes")
    ap.add_argument("--workers", type=int, default=0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="time the batched launches only (no per-tree plain passes; for counter runs)")
    ap.add_argument("--no-compress", action="store_true", help="skip the pattern-compression measurement")
    ap.add_argument("--no-site", action="store_true", help="skip the site-specific model measurement")
    ap.add_argument("--serve-rate", default="250,1000,4000,16000",
                    help="request arrival rates (per second) to sweep, comma-separated")
    ap.add_argument("--serve-requests", type=int, default=600)
    ap.add_argument("--serve-pool", type=int, default=64)
    ap.add_argument("--serve-batch", default="",
                    help="minimum group sizes for device-batched serving replays (comma list; empty: none)")
    ap.add_argument("--branch-opt", type=int, default=0, metavar="K",
                    help="also time the device branch-length sweeps (joint and marginal) on the first K trees "
                         "against the FP64 oracle's per-branch Brent")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES between two device buffers 5 times (HBM counter calibration)")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "batch_marginal_mix256_traffic.md"),
                    help="measured HBM bytes per launch of the batched kernel (from rocprofv3 --pmc)")
    return ap.parse_args()


def n_workers(req):
    if req:
        return req
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def gen_preps(specs, workers):
    from treetime_amd.synth import prepared_job

    if workers <= 1 or len(specs) == 1:
        return [prepared_job(s) for s in specs]
    # big jobs first so the pool's tail is short
    order = sorted(range(len(specs)), key=lambda i: -specs[i]["n_tips"] * specs[i]["L"])
    ctx = mp.get_context("fork")  # no GPU has been touched yet in this process
    with ctx.Pool(workers) as pool:
        res = pool.map(prepared_job, [specs[i] for i in order], chunksize=1)
    out = [None] * len(specs)
    for i, r in zip(order, res):
        out[i] = r
    return out


def tree_arrays(prep):
    parent = prep["parent"]
    is_tip = np.ones(len(parent), dtype=bool)
    is_tip[parent[parent >= 0]] = False
    return parent, is_tip


def updates(prep):
    return prep["tip_codes"].shape[1] * (len(prep["parent"]) - 1)


def algorithmic_bytes(prep):
    """Minimal compulsory HBM bytes of one FP32 marginal pass: per pattern,
    18n+7 per internal-child branch, 2n+3 per tip branch, plus 4 (weights)
    and 4n (root) per pattern."""
    n = prep["n_states"]
    parent, is_tip = tree_arrays(prep)
    nonroot = parent >= 0
    n_tip_br = int((is_tip & nonroot).sum())
    n_int_br = int((~is_tip & nonroot).sum())
    P = prep["tip_codes"].shape[1]
    return P * (n_int_br * (18 * n + 7) + n_tip_br * (2 * n + 3) + 4 + 4 * n)


def joint_algorithmic_bytes(prep):
    """Minimal compulsory HBM bytes of one joint pass with FP32 messages: per
    pattern, 8n+c+3 per internal-child branch (the child's message row written
    and read back, 4n each; its argmax row written, c bytes: 1 for n <= 4 at 2
    bits per parent state, else n; the backtrack reads the parent state and
    one argmax byte and writes the state, 3), 3 per tip branch (the character
    code read by the postorder and again by the log-likelihood score of the
    decoded states, plus the parent state), plus 8 (weights) and 9 (root:
    logL, state)."""
    n = prep["n_states"]
    parent, is_tip = tree_arrays(prep)
    nonroot = parent >= 0
    n_tip_br = int((is_tip & nonroot).sum())
    n_int_br = int((~is_tip & nonroot).sum())
    P = prep["tip_codes"].shape[1]
    cx = 1 if n <= 4 else n
    return P * (n_int_br * (8 * n + cx + 3) + 3 * n_tip_br + 8 + 9)


def loglik_algorithmic_bytes(prep):
    """Minimal compulsory HBM bytes of one log-likelihood pass (postorder +
    root): per pattern, 6n+2 per internal-child branch (the child's row read,
    4n, half of the parent's row write, 2n+2), 2n+3 per tip branch, plus 8
    (weights), 4n (root row) and 8 (per-pattern logL)."""
    n = prep["n_states"]
    parent, is_tip = tree_arrays(prep)
    nonroot = parent >= 0
    n_tip_br = int((is_tip & nonroot).sum())
    n_int_br = int((~is_tip & nonroot).sum())
    P = prep["tip_codes"].shape[1]
    return P * (n_int_br * (6 * n + 2) + n_tip_br * (2 * n + 3) + 16 + 4 * n)


def oracle_marginal(prep):
    """FP64 oracle (CPU baseline / checker) on one prepared job."""
    return oracle_tree(prep).marginal()


def oracle_tree(prep):
    """The FP64 oracle's TreeAnc over one prepared job."""
    from oracle import fp64 as O

    parent, is_tip = tree_arrays(prep)
    names = ["n%d" % i for i in range(len(parent))]
    root, nodes = O.build_tree(parent, prep["t"], names)
    alpha = {4: "nuc_nogap", 20: "aa_nogap"}[prep["n_states"]]
    chars = prep["tip_codes"].view("S1").astype("U1")
    tips = np.where(is_tip)[0]
    comp = {names[v]: chars[k] for k, v in enumerate(tips)}
    g = O.GTRModel(alpha, mu=prep["mu"], pi=prep["pi"], W=prep["W"])
    return O.OracleTreeAnc(root, g, comp, prep["multiplicity"], full_length=1e300)


def cpu_info():
    """The host the CPU baseline ran on."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    env = {v: os.environ.get(v) for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    return {"cpu_model": model, "affinity_cores": cores, "blas_threads_env": env}


def _oracle_job(prep):
    t0 = time.perf_counter()
    oracle_marginal(prep)
    return updates(prep), time.perf_counter() - t0


def cpu_baseline(preps, workers):
    """FP64 numpy oracle marginal passes on the host, jobs across the size
    range: (a) one process, one BLAS thread, the jobs at the centres of the
    four quartiles of the updates per job; (b) `workers` processes, one BLAS
    thread each, over every job below the 90th percentile (largest first),
    run to completion: aggregate updates/s over the wall time."""
    order = sorted(range(len(preps)), key=lambda i: updates(preps[i]))
    n = len(order)
    picks = sorted({order[min(n - 1, int(q * n))] for q in (0.125, 0.375, 0.625, 0.875)})
    t0 = time.perf_counter()
    done_u = sum(_oracle_job(preps[i])[0] for i in picks)
    single = (done_u / (time.perf_counter() - t0), time.perf_counter() - t0, picks)
    par = None
    if workers > 1:
        jobs = [order[j] for j in range(int(0.9 * n) - 1, -1, -1)]
        ctx = mp.get_context("fork")  # no GPU has been touched yet in this process
        t0 = time.perf_counter()
        with ctx.Pool(workers) as pool:
            done_u = sum(u for u, _ in pool.imap_unordered(_oracle_job, [preps[i] for i in jobs], chunksize=1))
        wall = time.perf_counter() - t0
        par = (done_u / wall, wall, len(jobs))
    return single, par


def serve_bench(eng, pool, args):
    """Poisson request stream over a resident pool: 70% branch re-evaluation
    (root path only), 20% marginal, 10% joint; HBM budget 200 GB.  The same
    stream is replayed with every branch request recomputed from scratch,
    and with same-kind pass requests fused into batched launches once at
    least B of them are admitted together (``--serve-batch``)."""
    from treetime_amd.serve import Request, Server, make_stream, summarize

    for t in pool:  # switch to update mode (keeps per-node normalisers)
        node = int(t.nonroot_nodes()[0])
        t.update_branch(node, float(t.t_cur[node]))
    rates = [float(x) for x in str(args.serve_rate).split(",") if x.strip()]
    out = {"pool": len(pool), "mix": "70% branch / 20% marginal / 10% joint", "hbm_budget_GB": 200,
           "requests_per_rate": args.serve_requests, "sweep": []}
    for rate in rates:
        rng = np.random.default_rng(4000)
        reqs = make_stream(pool, args.serve_requests, rate, rng)

        def replay(scratch, device_batch=0):
            rr = [Request(r.kind, r.tree, r.arrival, r.node, r.scale, r.rid) for r in reqs]
            srv = Server(pool, engine=eng, slots=16, hbm_budget=200 * 10**9, scratch=scratch,
                         device_batch=device_batch)
            w0 = time.perf_counter()
            done = srv.run(rr, realtime=True)
            out = summarize(done, time.perf_counter() - w0)
            if device_batch:
                out["batched_launches"] = srv.batches
                out["requests_in_batches"] = sum(1 for r in done if r.meta.get("batched"))
            return out

        rp = replay(False)
        sc = replay(True)
        log(f"[bench] serving @ {rate:.0f} req/s offered: p50 {rp['p50_ms']:.2f} ms p99 {rp['p99_ms']:.2f} ms "
            f"({rp['throughput_req_s']:.0f} req/s served); from scratch p50 {sc['p50_ms']:.2f} p99 {sc['p99_ms']:.2f} "
            f"({sc['throughput_req_s']:.0f} req/s)")
        entry = {"rate_req_s": rate, "root_path": rp, "from_scratch": sc}
        for b in [int(x) for x in str(args.serve_batch).split(",") if x.strip()]:
            db = replay(False, device_batch=b)
            log(f"[bench] serving @ {rate:.0f} req/s, device batches of >= {b}: p50 {db['p50_ms']:.2f} ms "
                f"p99 {db['p99_ms']:.2f} ms ({db['throughput_req_s']:.0f} req/s served, "
                f"{db['requests_in_batches']} requests in {db['batched_launches']} batched launches)")
            entry[f"device_batched_min{b}"] = db
        out["sweep"].append(entry)
    return out


def copy_probe(nbytes):
    """A plain device-to-device copy of known size: what FETCH_SIZE and
    WRITE_SIZE read for it scales the counters of the batched kernel."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copied {nbytes} bytes 5 times")
    return 0


def branch_opt_bench(trees, preps, torch):
    """Device branch-length sweeps (all branches, one launch each) against
    the FP64 oracle's scipy Brent on a sample of branches (host, 1 core)."""
    from oracle import fp64 as O

    n_br = sum(t.n_nodes - 1 for t in trees)
    codes = np.full(256, -1, dtype=np.int16)
    for i, ch in enumerate("ACGT"):
        codes[ord(ch)] = i
    for t in trees:
        t.marginal(want_logL=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in trees:
        t.optimize_branch_lengths_marginal(-1, 4.0, 1e-10)
    dt_m = time.perf_counter() - t0
    dt_j = float("nan")
    if trees[0].n == 4:
        for t in trees:
            t.joint(want_logL=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in trees:
            t.optimize_branch_lengths(codes, -1, 4.0, 1e-10, with_counts=False)
        dt_j = time.perf_counter() - t0
    log(f"[bench] branch sweeps on {len(trees)} trees ({n_br} branches): marginal {dt_m * 1e3:.1f} ms "
        f"({n_br / dt_m:.3e} branches/s), joint {dt_j * 1e3:.1f} ms ({n_br / dt_j:.3e} branches/s)")
    # oracle: marginal profiles of the first tree, then Brent per branch for ~10 s
    prep = preps[0]
    ta = oracle_tree(prep)
    ta.marginal()
    og = ta.gtr
    done, t0 = 0, time.perf_counter()
    for nd in O.preorder(ta.root):
        if nd.up is None:
            continue
        og.optimal_t_compressed((nd.marginal_outgroup_LH, nd.marginal_subtree_LH), prep["multiplicity"],
                                profiles=True, tol=1e-10)
        done += 1
        if time.perf_counter() - t0 > 10:
            break
    dt = time.perf_counter() - t0
    log(f"[bench] oracle marginal Brent: {done} branches of tree 0 (P={prep['tip_codes'].shape[1]}) in {dt:.1f}s "
        f"= {done / dt:.2f} branches/s (1 core)")


def synthetic_alignment(N, L, n_variable, seed):
    """(N, L) uint8 nucleotide alignment of the large single-tree shape: a
    random root sequence, n_variable columns carrying mutations on a random
    subset of rows, 2% N, and leading / trailing gaps on 1% of the rows."""
    rng = np.random.default_rng(seed)
    acgt = np.frombuffer(b"ACGT", dtype=np.uint8)
    A = np.repeat(rng.choice(acgt, size=L)[None, :], N, axis=0)
    cols = rng.choice(L, size=n_variable, replace=False)
    for c in cols:
        rows = rng.random(N) < rng.uniform(0.001, 0.3)
        rows[rng.integers(N)] = True
        A[rows, c] = rng.choice(acgt, size=int(rows.sum()))
    k = N * L // 50
    A[rng.integers(0, N, k), rng.integers(0, L, k)] = ord("N")
    for r in rng.choice(N, size=max(1, N // 100), replace=False):
        A[r, : rng.integers(1, 500)] = ord("-")
        A[r, L - rng.integers(1, 500):] = ord("-")
    return A


def compress_bench(eng, torch, steps, N=5000, L=30000, n_variable=2600, cpu_seconds=5.0):
    """Pattern compression (tt_compress_alignment) of a device-resident
    5000 x 30000 alignment with the overhang fill, timed with HIP events per
    call (each call waits for its pattern count once, mid-way); the FP64
    oracle's column-by-column compression on a sample of columns beside it."""
    from oracle import fp64 as O
    from treetime_amd.prepare import fill_overhangs
    from treetime_amd.sequence_data import compress_columns

    A = synthetic_alignment(N, L, n_variable, 4242)
    pitch = (L + 15) // 16 * 16
    host = np.zeros((N, pitch), dtype=np.uint8)
    host[:, :L] = A
    d = torch.from_numpy(host).cuda()
    pat = torch.empty(L, dtype=torch.int32, device="cuda")
    comp = torch.empty(N * L, dtype=torch.uint8, device="cuda")
    mult = torch.empty(L, dtype=torch.float64, device="cuda")
    P = eng.compress_alignment_device(d, L, "N", pat, comp, mult, fill_overhangs=True)
    # parity of this exact input against the numpy compression (itself pinned to the oracle)
    rp, rc = compress_columns(fill_overhangs(A, "N"), "N")
    ok = (P == rc.shape[1] and np.array_equal(pat.cpu().numpy(), rp)
          and np.array_equal(comp[:N * P].cpu().numpy().reshape(N, P), rc))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        eng.compress_alignment_device(d, L, "N", pat, comp, mult, fill_overhangs=True)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    n_var = int((mult[:P].cpu().numpy() == 1).sum())
    alg = N * L + N * n_var + N * P + 4 * L + 8 * P
    # CPU: the oracle's column loop on the first columns for about cpu_seconds
    names = ["s%d" % i for i in range(N)]
    F = fill_overhangs(A, "N").view("S1").astype("U1")
    cols, t0 = 250, time.perf_counter()
    while True:
        O.compress_alignment(names, {nm: F[i, :cols] for i, nm in enumerate(names)}, "N")
        dt = time.perf_counter() - t0
        if dt > cpu_seconds or cols >= L:
            break
        cols, t0 = min(L, cols * 2), time.perf_counter()
    out = {"workload": f"compress-{N}x{L}", "patterns": P, "variable_patterns": n_var, "ms_per_call": ms,
           "columns_per_s": L / (ms / 1e3), "algorithmic_bytes": alg, "achieved_GBs": alg / (ms / 1e3) / 1e9,
           "frac_of_hbm_peak": alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, "parity_vs_numpy": bool(ok),
           "cpu_oracle_columns_per_s": cols / dt, "cpu_sample": f"oracle column loop over the first {cols} columns"}
    log(f"[bench] compress {N}x{L}: P={P} ({n_var} variable) {ms:.3f} ms/call, {out['achieved_GBs']:.0f} GB/s "
        f"algorithmic, parity {ok}; oracle {cols / dt:.0f} columns/s")
    del d, comp
    torch.cuda.empty_cache()
    return out


def site_bench(eng, torch, steps, n_tips=1000, L=5000, seed=77, cpu=True):
    """Marginal and joint passes with a site-specific GTR model
    (GTR_site_specific.random, approximate=True) on an uncompressed
    1000-tip x 5000-site alignment; the FP64 oracle's SiteGTRModel pass on
    the same input as the parity check and the CPU baseline."""
    from oracle import fp64 as O
    from treetime_amd import GTR_site_specific, synth
    from treetime_amd.engine import DeviceTree, code_lut
    from treetime_amd.prepare import prepare

    job = synth.make_job(n_tips, L, 0.05, 4, seed, 0.01, 0.005)
    prep = prepare(job["parent"], job["branch_length"], job["chars"], job["gtr"], job["ambiguous"], compress=False)
    g = GTR_site_specific.random(L=L, alphabet="nuc_nogap", rng=np.random.default_rng(seed))
    dev = DeviceTree(eng, prep["parent"], prep["t"], prep["tip_codes"], code_lut(g.profile_map, 4),
                     prep["multiplicity"], g.Pi.mean(axis=1), None, None, None, None, site=g.device_model())
    s = torch.cuda.Stream()
    sp = s.cuda_stream
    lh = dev.marginal()
    ljt = dev.joint()
    res = {}
    for name, fn in (("marginal", dev.marginal), ("joint", dev.joint)):
        fn(stream=sp, want_logL=False)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(steps):
            fn(stream=sp, want_logL=False)
        e1.record(s)
        torch.cuda.synchronize()
        res[name] = e0.elapsed_time(e1) / steps
    U = L * (len(prep["parent"]) - 1)
    out = {"workload": f"site-{n_tips}x{L}", "updates_per_pass": U, "marginal_ms": res["marginal"],
           "marginal_updates_per_s": U / (res["marginal"] / 1e3), "joint_ms": res["joint"],
           "joint_updates_per_s": U / (res["joint"] / 1e3)}
    if cpu:
        parent = prep["parent"]
        names = ["n%d" % i for i in range(len(parent))]
        root, nodes = O.build_tree(parent, prep["t"], names)
        is_tip = np.ones(len(parent), dtype=bool)
        is_tip[parent[parent >= 0]] = False
        chars = prep["tip_codes"].view("S1").astype("U1")
        comp = {names[v]: chars[k] for k, v in enumerate(np.where(is_tip)[0])}
        og = O.SiteGTRModel("nuc_nogap", mu=g.mu, pi=g.Pi, W=g.W)
        ta = O.OracleTreeAnc(root, og, comp, prep["multiplicity"], full_length=1e300)
        t0 = time.perf_counter()
        ref = ta.marginal()
        dt = time.perf_counter() - t0
        ref_j = ta.joint()
        out.update({"logL_rel_err_vs_fp64_oracle": abs(lh - ref) / abs(ref),
                    "joint_logL_rel_err_vs_fp64_oracle": abs(ljt - ref_j) / abs(ref_j),
                    "cpu_oracle_marginal_s": dt, "cpu_oracle_updates_per_s": U / dt})
    log(f"[bench] site-specific {n_tips}x{L}: marginal {res['marginal']:.3f} ms "
        f"({out['marginal_updates_per_s']:.3e} U/s), joint {res['joint']:.3f} ms"
        + (f"; logL rel err {out['logL_rel_err_vs_fp64_oracle']:.1e}, oracle {dt:.1f} s" if cpu else ""))
    dev.close()
    return out


def hbm_ceilings(torch, nbytes=4 << 30, reps=5):
    """Streaming rates of the library's own kernels (tt_hbm_probe: 16-byte
    accesses, every CU busy), GB/s of bytes moved: read-only sweep, copy
    (read + write) and write-only sweep."""
    import ctypes as C

    from treetime_amd import _native

    lib = _native.load()
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    ms = C.c_double()
    out = {}
    for kind, name, moved in ((0, "read", nbytes), (1, "copy", 2 * nbytes), (2, "write", nbytes)):
        _native.check(lib.tt_hbm_probe(kind, b.data_ptr(), a.data_ptr(), nbytes, reps, C.byref(ms),
                                       torch.cuda.current_stream().cuda_stream))
        out[name] = moved / (ms.value / 1e3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return out


def copy_ceiling(torch, nbytes=2 << 30, reps=5):
    """Streaming rates of plain device kernels (GB/s, bytes moved per second):
    copy (read + write) and write-only (fill) -- the measured ceilings beside
    the datasheet peak."""
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)

    def rate(fn, moved):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return moved * reps / (e0.elapsed_time(e1) / 1e3) / 1e9

    out = {"copy": rate(lambda: b.copy_(a), 2 * nbytes), "write": rate(lambda: b.fill_(1.0), nbytes)}
    del a, b
    torch.cuda.empty_cache()
    return out


def measured_traffic(path, workload):
    """HBM bytes per launch recorded for ``workload`` in a profile summary
    (line 'traffic_bytes_per_launch <workload> <bytes>'), else None."""
    try:
        with open(path) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == "traffic_bytes_per_launch" and w[1] == workload:
                    return float(w[2])
    except OSError:
        pass
    return None


SINGLE_TREES = ("plain-200", "large-5000", "large-5000-deep", "aa-1000")


def single_spec(name):
    from treetime_amd import synth

    n_tips, L, depth, n_states, seed, pa, pg = synth.WORKLOADS[name]
    return dict(n_tips=n_tips, L=L, depth=depth, n_states=n_states, seed=seed, p_ambig=pa, p_gap=pg, name=name)


def oracle_logl_of(name, prep):
    """The FP64 oracle's marginal and joint logL of a single-tree workload:
    computed here for the 200-tip tree, taken from the committed oracle
    digests (tests/golden/configs_golden.npz, checked against the inputs'
    hash) for the large ones, whose oracle runs take minutes."""
    if name == "plain-200":
        return oracle_marginal(prep), oracle_tree(prep).joint()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import config_digest as CD

    path = os.path.join(ROOT, "tests", "golden", "configs_golden.npz")
    if not os.path.exists(path):
        return None, None
    with np.load(path) as z:
        if f"{name}/hash" not in z.files or str(z[f"{name}/hash"]) != CD.input_hash(prep):
            return None, None
        return float(z[f"{name}/m_logL"]), float(z[f"{name}/j_logL"])


def single_tree_bench(eng, torch, named, reps=50):
    """The plain (drop-in) passes on one tree at a time: ms per marginal,
    log-likelihood and joint pass (HIP events over `reps` passes on one
    stream), updates/s, the algorithmic bytes rate against the HBM peak, and
    logL against the FP64 oracle."""
    from treetime_amd.prepare import device_tree

    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    out = {}
    for name, prep in named.items():
        dt = device_tree(eng, prep)
        U = updates(prep)
        bytes_of = {"marginal": algorithmic_bytes(prep), "loglik": loglik_algorithmic_bytes(prep),
                    "joint": joint_algorithmic_bytes(prep)}
        res = {"tips": int((len(prep["parent"]) + 1) // 2), "patterns": int(prep["tip_codes"].shape[1]),
               "states": int(prep["n_states"]), "updates_per_pass": int(U)}
        got = {}
        for kind in ("marginal", "loglik", "joint"):
            f = getattr(dt, kind)
            got[kind] = f(stream=sp)
            for _ in range(3):
                f(stream=sp, want_logL=False)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                f(stream=sp, want_logL=False)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            res[kind] = {"ms": ms, "updates_per_s": U / (ms / 1e3), "algorithmic_bytes": int(bytes_of[kind]),
                         "achieved_GBs": bytes_of[kind] / (ms / 1e3) / 1e9,
                         "frac_of_hbm_peak": bytes_of[kind] / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
        dt.close()
        ref_m, ref_j = oracle_logl_of(name, prep)
        res["logL_rel_err_vs_fp64_oracle"] = {
            "marginal": abs(got["marginal"] - ref_m) / abs(ref_m) if ref_m is not None else None,
            "joint": abs(got["joint"] - ref_j) / abs(ref_j) if ref_j is not None else None}
        out[name] = res
        log(f"[bench] single tree {name}: marginal {res['marginal']['ms']:.4f} ms, loglik {res['loglik']['ms']:.4f} ms, "
            f"joint {res['joint']['ms']:.4f} ms; logL rel err {res['logL_rel_err_vs_fp64_oracle']}")
    return out


def main():
    args = parse()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described, not built",
                          "n_gpus": args.gpus}))
        log("bench.py: only --gpus 1 is supported (one MI355X per box)")
        return 2
    from treetime_amd import synth

    if args.workload == "site":
        import torch

        from treetime_amd.engine import Engine
        eng = Engine(0, 0)
        print(json.dumps({"site_specific": site_bench(eng, torch, args.steps, cpu=not args.no_cpu)}), flush=True)
        eng.close()
        return 0
    if args.workload == "compress":
        import torch

        from treetime_amd.engine import Engine
        eng = Engine(0, 0)
        print(json.dumps({"compression": compress_bench(eng, torch, args.steps, cpu_seconds=0 if args.no_cpu else 5.0)}),
              flush=True)
        eng.close()
        return 0
    if args.workload == "mix":
        specs = [synth.mix_job_spec(j) for j in range(args.jobs)]
        wname = f"mix{args.jobs}-nuc-marginal-batched"
    else:
        n_tips, L, depth, n_states, seed, pa, pg = synth.WORKLOADS[args.workload]
        specs = [dict(n_tips=n_tips, L=L, depth=depth, n_states=n_states, seed=seed, p_ambig=pa, p_gap=pg)]
        wname = args.workload
    singles = list(SINGLE_TREES) if (args.workload == "mix" and not args.no_single) else []
    t0 = time.perf_counter()
    all_preps = gen_preps(specs + [single_spec(nm) for nm in singles], n_workers(args.workers))
    preps, single_preps = all_preps[:len(specs)], dict(zip(singles, all_preps[len(specs):]))
    log(f"[bench] generated {len(all_preps)} jobs in {time.perf_counter() - t0:.1f}s")

    # CPU baseline first: the worker processes fork before anything touches the GPU
    cpu = None
    if not args.no_cpu:
        info = cpu_info()
        workers = n_workers(args.workers)
        (rate1, secs1, picks), par = cpu_baseline(preps, workers)
        log(f"[bench] cpu baseline: 1 core {rate1:.3e} U/s over {secs1:.1f}s ({len(picks)} jobs)"
            + (f"; {workers} cores {par[0]:.3e} U/s over {par[1]:.1f}s ({par[2]} jobs)" if par else ""))
        cpu = {"value": par[0] if par else rate1, "unit": "pattern*branch updates/s",
               "cores": workers if par else 1, "kind": "port",
               "sample": (f"FP64 numpy oracle marginal passes, one BLAS thread per process: {workers} processes "
                          f"over the {par[2]} jobs of the mix below its 90th percentile of updates per job, "
                          f"{par[1]:.1f}s wall" if par else
                          f"FP64 numpy oracle marginal passes on {len(picks)} jobs"),
               "single_core_value": rate1,
               "single_core_sample": f"jobs {picks} (centres of the quartiles of updates per job), {secs1:.1f}s",
               **info}

    import torch

    from treetime_amd.engine import Engine, batch_info, batch_marginal, batch_run
    from treetime_amd.prepare import device_tree

    eng = Engine(0, 0)
    t0 = time.perf_counter()
    trees = [device_tree(eng, p) for p in preps]
    torch.cuda.synchronize()
    upload_s = time.perf_counter() - t0
    log(f"[bench] {len(trees)} trees resident ({eng.usage()[0] / 2**30:.1f} GiB) in {upload_s:.1f}s")
    U = sum(updates(p) for p in preps)
    B = sum(algorithmic_bytes(p) for p in preps)
    BJ = sum(joint_algorithmic_bytes(p) for p in preps)
    BL = sum(loglik_algorithmic_bytes(p) for p in preps)
    P_total = sum(p["tip_codes"].shape[1] for p in preps)

    ceilings = hbm_ceilings(torch)
    torch_ceilings = copy_ceiling(torch)
    ceiling = ceilings["copy"]
    log(f"[bench] streaming ceilings (GB/s), own kernels: read {ceilings['read']:.0f}, copy {ceilings['copy']:.0f}, "
        f"write {ceilings['write']:.0f}; torch: copy {torch_ceilings['copy']:.0f}, fill {torch_ceilings['write']:.0f}")
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    # correctness: batched logL equals the plain pass per tree
    plain_logl = np.array([t.marginal(stream=sp) for t in trees])
    batch_logl = batch_marginal(eng, trees, stream=sp)
    rel = float(np.max(np.abs(batch_logl - plain_logl) / np.abs(plain_logl)))
    log(f"[bench] batched vs plain logL max rel diff {rel:.2e}")
    shape = batch_info(eng, trees)
    log(f"[bench] autotuned launch shape {shape}")
    small = int(np.argmin([updates(p) for p in preps]))
    ref = oracle_marginal(preps[small])
    err_fp64 = abs(plain_logl[small] - ref) / abs(ref)
    log(f"[bench] job {small}: logL rel err vs FP64 oracle {err_fp64:.2e}")

    # batched: warmup + timed
    for _ in range(args.warmup):
        batch_marginal(eng, trees, stream=sp, want_logL=False)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        batch_marginal(eng, trees, stream=sp, want_logL=False)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = ev0.elapsed_time(ev1) / args.steps
    # per-step distribution: the same steps again, one event pair each
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for a, b in evs:
        a.record(stream)
        batch_marginal(eng, trees, stream=sp, want_logL=False)
        b.record(stream)
    torch.cuda.synchronize()
    step_ms = np.array([a.elapsed_time(b) for a, b in evs])
    log(f"[bench] batched: {ms:.3f} ms/step (events), wall {wall / args.steps * 1e3:.3f} ms/step")

    # batched log-likelihood only (postorder + root, the re-evaluation pass)
    ms_ll = float("nan")
    if args.plain_steps > 0:
        batch_run(eng, trees, "loglik", stream=sp, want_logL=False)
        torch.cuda.synchronize()
        l0, l1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        l0.record(stream)
        for _ in range(args.plain_steps):
            batch_run(eng, trees, "loglik", stream=sp, want_logL=False)
        l1.record(stream)
        torch.cuda.synchronize()
        ms_ll = l0.elapsed_time(l1) / args.plain_steps
        log(f"[bench] batched loglik: {ms_ll:.3f} ms/step")

    # batched joint: every tree's joint pass on one stream
    ms_joint = ms_joint_plain = float("nan")
    if args.plain_steps > 0:
        batch_run(eng, trees, "joint", stream=sp, want_logL=False)
        torch.cuda.synchronize()
        j0, j1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        j0.record(stream)
        for _ in range(args.plain_steps):
            batch_run(eng, trees, "joint", stream=sp, want_logL=False)
        j1.record(stream)
        torch.cuda.synchronize()
        ms_joint = j0.elapsed_time(j1) / args.plain_steps
        log(f"[bench] joint, all trees: {ms_joint:.3f} ms/step")
    if args.plain_steps > 0 and not args.batched_only:
        # the same joint passes one tree at a time (each tree's captured graph)
        for t in trees:
            t.joint(stream=sp, want_logL=False)
        torch.cuda.synchronize()
        j0.record(stream)
        for _ in range(args.plain_steps):
            for t in trees:
                t.joint(stream=sp, want_logL=False)
        j1.record(stream)
        torch.cuda.synchronize()
        ms_joint_plain = j0.elapsed_time(j1) / args.plain_steps
        log(f"[bench] joint, plain: {ms_joint_plain:.3f} ms/step")
    for t in trees:  # return the joint buffers (tens of GB on the mix) before the serving leg
        t.trim()

    # plain: one tree per launch sequence, trees in sequence
    ms_plain = float("nan")
    if args.plain_steps > 0 and not args.batched_only:
        for t in trees:
            t.marginal(stream=sp, want_logL=False)
        torch.cuda.synchronize()
        p0, p1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        p0.record(stream)
        for _ in range(args.plain_steps):
            for t in trees:
                t.marginal(stream=sp, want_logL=False)
        p1.record(stream)
        torch.cuda.synchronize()
        ms_plain = p0.elapsed_time(p1) / args.plain_steps
        log(f"[bench] plain: {ms_plain:.3f} ms/step")

    if args.branch_opt:
        branch_opt_bench(trees[: args.branch_opt], preps[: args.branch_opt], torch)

    compression = None
    if not args.no_compress:
        compression = compress_bench(eng, torch, 10, cpu_seconds=0 if args.no_cpu else 5.0)
    site = None
    if not args.no_site and args.workload == "mix":
        site = site_bench(eng, torch, 10, cpu=not args.no_cpu)

    serving = None
    if not args.no_serve and len(trees) > 1:
        serving = serve_bench(eng, trees[: args.serve_pool], args)

    single = single_tree_bench(eng, torch, single_preps) if single_preps else None

    value = U / (ms / 1e3)
    achieved = B / (ms / 1e3) / 1e9
    out = {
        "metric": "pattern*branch updates/s (marginal reconstruction)",
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": wname,
            "jobs": len(preps),
            "patterns_total": int(P_total),
            "updates_per_step": int(U),
            "precision": "f32 rows, messages and preorder; f64 normaliser products, root and final sum",
            "plain_ms_per_step": ms_plain,
            "batched_loglik_ms_per_step": ms_ll,
            "batched_loglik_value": U / (ms_ll / 1e3),
            "joint_ms_per_step": ms_joint,
            "plain_joint_ms_per_step": ms_joint_plain,
            "joint_value": U / (ms_joint / 1e3),
            "joint_algorithmic_bytes_per_step": int(BJ),
            "joint_achieved_GBs": BJ / (ms_joint / 1e3) / 1e9,
            "loglik_algorithmic_bytes_per_step": int(BL),
            "loglik_achieved_GBs": BL / (ms_ll / 1e3) / 1e9,
            "plain_value": U / (ms_plain / 1e3),
            "batched_vs_plain": ms_plain / ms,
            "batched_vs_plain_logL_max_rel_diff": rel,
            "logL_rel_err_vs_fp64_oracle": err_fp64,
            "wall_ms_per_step": wall / args.steps * 1e3,
            "step_ms_median": float(np.median(step_ms)),
            "step_ms_p99": float(np.percentile(step_ms, 99)),
            "resident_upload_s": upload_s,
            "launch_shape": shape,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.traffic_file, wname),
                     "copy_ceiling": ceiling, "frac_of_copy_ceiling": achieved / ceiling,
                     "streaming_ceilings": {**ceilings, "torch_copy": torch_ceilings["copy"],
                                            "torch_fill": torch_ceilings["write"]},
                     "algorithmic_bytes_per_launch": int(B)},
        "cpu_baseline": cpu,
        "single_tree": single,
        "serving": serving,
        "compression": compression,
        "site_specific": site,
    }
    print(json.dumps(out), flush=True)
    for t in trees:
        t.close()
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

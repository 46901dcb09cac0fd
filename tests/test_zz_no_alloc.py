"""The allocation-free entry point tt_prob_t_profiles_noalloc against the
allocating tt_prob_t_profiles (identical values) and the FP64 oracle, on two
calls in a row, the second with a smaller input; and its size limit."""
import ctypes as C

import numpy as np
import pytest

from oracle import fp64 as O
from treetime_amd import _native


def _case(rng, P, n, K):
    W = rng.lognormal(0, 0.5, (n, n))
    W = 0.5 * (W + W.T)
    np.fill_diagonal(W, 0)
    g = O.GTRModel("nuc" if n == 5 else "nuc_nogap", pi=rng.dirichlet(np.full(n, 5.0)), W=W)
    pp = rng.dirichlet(np.ones(n), P)
    pc = rng.dirichlet(np.ones(n), P)
    mult = rng.integers(1, 5, P).astype(np.float64)
    ts = np.linspace(0.001, 0.4, K)
    Qs = np.ascontiguousarray(np.stack([g.expQt(t) for t in ts]))
    gap = 4 if n == 5 else -1
    ref = np.array([g.prob_t_profiles((pp, pc), mult, t, return_log=True, ignore_gaps=gap >= 0) for t in ts])
    return pp, pc, mult, Qs, gap, ref


def _call(lib, fn, dpp, dpc, dm, P, n, Qs, gap, stream):
    out = np.empty(len(Qs))
    _native.check(getattr(lib, fn)(dpp.data_ptr(), dpc.data_ptr(), dm.data_ptr(), P, n, _native.ptr(Qs, C.c_double),
                                   len(Qs), gap, _native.ptr(out, C.c_double), stream))
    return out


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry_point():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = _native.load()
    rng = np.random.default_rng(2024)
    stream = torch.cuda.current_stream().cuda_stream
    # second call smaller: fewer patterns, fewer matrices, other alphabet
    for P, n, K in [(60000, 4, 125), (3001, 5, 17)]:
        print(f"noalloc case P={P} n={n} K={K}", flush=True)
        pp, pc, mult, Qs, gap, ref = _case(rng, P, n, K)
        dpp, dpc, dm = (torch.from_numpy(a).to("cuda") for a in (pp, pc, mult))
        got = _call(lib, "tt_prob_t_profiles_noalloc", dpp, dpc, dm, P, n, Qs, gap, stream)
        base = _call(lib, "tt_prob_t_profiles", dpp, dpc, dm, P, n, Qs, gap, stream)
        assert np.array_equal(got, base)
        np.testing.assert_allclose(got, ref, rtol=1e-10, atol=0)


@pytest.mark.gpu
def test_noalloc_refuses_sizes_beyond_its_scratch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = _native.load()
    P, n, K = 1 << 20, 4, 200  # 200 * (16 + 4096 + 1) doubles > 524288
    d = torch.zeros(16, dtype=torch.float64, device="cuda")  # never read: the size check comes first
    Qs = np.zeros((K, n, n))
    out = np.empty(K)
    rc = lib.tt_prob_t_profiles_noalloc(d.data_ptr(), d.data_ptr(), d.data_ptr(), P, n, _native.ptr(Qs, C.c_double),
                                        K, -1, _native.ptr(out, C.c_double), None)
    assert rc == _native.TT_EINVAL
    assert b"scratch" in lib.tt_last_error()

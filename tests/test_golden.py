"""Golden vectors (tests/golden/oracle_small.npz, written by
tests/golden/make_oracle_golden.py): FP64 oracle outputs of two seeded small
configurations, a 4-state and a 20-state job.  The CPU test pins the oracle
to its committed outputs; the GPU test holds the HIP engine to them with the
tolerances of test_gpu_parity.py."""
import os

import numpy as np
import pytest

import treetime_amd as tt
from oracle import fp64 as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_small.npz")
CASES = [("nuc", "nuc_nogap"), ("aa", "aa_nogap")]


def load(tag):
    z = np.load(GOLDEN)
    return {k[len(tag) + 1:]: z[k] for k in z.files if k.startswith(tag + "_")}


def oracle_from(g, alphabet):
    parent = g["parent"]
    names = ["n%d" % i for i in range(len(parent))]
    root, nodes = O.build_tree(parent, g["t"], names)
    is_tip = np.ones(len(parent), dtype=bool)
    is_tip[parent[parent >= 0]] = False
    chars = g["tip_codes"].view("S1").astype("U1")
    comp = {names[v]: chars[k] for k, v in enumerate(np.where(is_tip)[0])}
    model = O.GTRModel(alphabet, mu=float(g["mu"]), pi=g["pi"], W=g["W"])
    return O.OracleTreeAnc(root, model, comp, g["multiplicity"], full_length=1e300), nodes


@pytest.mark.parametrize("tag,alphabet", CASES)
def test_oracle_reproduces_golden(tag, alphabet):
    g = load(tag)
    ta, nodes = oracle_from(g, alphabet)
    assert abs(ta.marginal() - float(g["logL"])) <= 1e-12 * abs(float(g["logL"]))
    assert np.allclose(ta.sequence_LH, g["sequence_LH"], rtol=1e-12, atol=1e-12)
    internal = g["internal"]
    assert np.allclose(np.stack([nodes[i].marginal_profile for i in internal]), g["profiles"], atol=1e-12)
    assert np.array_equal(np.stack([nodes[i].seq_idx for i in internal]), g["marginal_states"])
    assert abs(ta.joint() - float(g["joint_logL"])) <= 1e-12 * abs(float(g["joint_logL"]))
    assert np.array_equal(np.stack([nodes[i].seq_idx for i in internal]), g["joint_states"])


@pytest.mark.gpu
@pytest.mark.parametrize("tag,alphabet", CASES)
def test_engine_matches_golden(gpu_engine, tag, alphabet):
    from treetime_amd.engine import DeviceTree, code_lut

    g = load(tag)
    model = tt.GTR(alphabet=alphabet)
    n = len(model.alphabet)
    dev = DeviceTree(gpu_engine, g["parent"], g["t"], g["tip_codes"], code_lut(model.profile_map, n),
                     g["multiplicity"], g["pi"], g["eigvals"], g["v"], g["v_inv"], float(g["mu"]))
    lh = dev.marginal()
    ref = float(g["logL"])
    assert abs(lh - ref) <= 1e-7 * abs(ref)
    seq = dev.sequence_lh()
    assert np.all(np.abs(seq - g["sequence_LH"]) <= 5e-7 * np.maximum(1.0, np.abs(g["sequence_LH"])))
    internal = g["internal"]
    prof = dev.profiles()[internal]
    assert np.max(np.abs(prof - g["profiles"])) <= 2e-6
    srt = np.sort(g["profiles"], axis=2)
    sure = srt[:, :, -1] - srt[:, :, -2] > 1e-5
    assert np.array_equal(dev.states()[internal][sure], g["marginal_states"][sure])
    jl = dev.joint()
    assert abs(jl - float(g["joint_logL"])) <= 1e-9 * abs(float(g["joint_logL"]))
    jsure = g["joint_margin"] > 1e-5
    assert np.array_equal(dev.states()[internal][jsure], g["joint_states"][jsure])
    dev.close()

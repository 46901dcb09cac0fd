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


# ----------------------------------------------------------- profile tables --
PROFILE_MAPS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "profile_maps.json")


def _reference_tables():
    import json

    with open(PROFILE_MAPS) as f:
        return json.load(f)


@pytest.mark.parametrize("alphabet", ["nuc", "nuc_nogap", "aa", "aa_nogap"])
def test_profile_maps_equal_reference_tables(alphabet):
    """The oracle's and the facade's alphabets and character profiles equal
    the literal tables of seq_utils.py:19-122 (tests/golden/profile_maps.json,
    read from the reference source with ast by make_profile_maps.py): the
    same characters, the same multi-hot vectors."""
    from treetime_amd import seq_utils as S

    ref = _reference_tables()
    alpha = ref["alphabets"][alphabet]
    assert list(O.ALPHABETS[alphabet]) == alpha
    assert list(S.alphabets[alphabet]) == alpha
    table = {c: np.array(v, dtype=float) for c, v in ref["profile_maps"][alphabet].items()}
    for name, pm in [("oracle", O.profile_map(alphabet)), ("facade", S.profile_maps[alphabet])]:
        assert sorted(pm) == sorted(table), name
        for c, v in table.items():
            assert np.array_equal(np.asarray(pm[c], dtype=float), v), (name, c)


def test_alphabet_synonyms_equal_reference():
    from treetime_amd import seq_utils as S

    assert S.alphabet_synonyms == _reference_tables()["alphabet_synonyms"]


@pytest.mark.parametrize("alphabet", ["nuc", "nuc_nogap", "aa", "aa_nogap"])
def test_device_code_lut_follows_reference_tables(alphabet):
    """The 256-entry code table the device reads (engine.code_lut) holds the
    reference's profile of every character of the tables, and all ones for
    every other byte (extend_profile, seq_utils.py:126-136)."""
    from treetime_amd.engine import code_lut
    from treetime_amd import seq_utils as S

    table = _reference_tables()["profile_maps"][alphabet]
    n = len(_reference_tables()["alphabets"][alphabet])
    lut = code_lut(S.profile_maps[alphabet], n)
    for b in range(256):
        c = chr(b)
        want = np.array(table[c], dtype=np.float32) if c in table else np.ones(n, dtype=np.float32)
        assert np.array_equal(lut[b], want), (alphabet, c)

"""From a tree + tip alignment to the device inputs of one ``DeviceTree``.

Steps (reference semantics in brackets):
  1. fill overhangs: leading/trailing '-' become the ambiguous character
     (seq_utils.py:196-202, SequenceData default fill_overhangs=True);
  2. pattern compression (sequence_data.py:325-464, see
     sequence_data.compress_columns);
  3. effective branch lengths t = max(1e-3 / L_full, branch_length)
     (treeanc.py:752-760, config.py:7).
"""
import numpy as np

from . import config as ttconf
from .engine import DeviceTree, code_lut
from .sequence_data import compress_columns


def fill_overhangs(codes, ambiguous):
    """Vectorised fill of leading/trailing gaps for a (N, L) uint8 matrix."""
    gap = codes == ord("-")
    ng = ~gap
    L = codes.shape[1]
    any_ng = ng.any(axis=1)
    first = np.where(any_ng, ng.argmax(axis=1), L)
    last = np.where(any_ng, L - 1 - ng[:, ::-1].argmax(axis=1), -1)
    cols = np.arange(L)
    over = (cols[None, :] < first[:, None]) | (cols[None, :] > last[:, None])
    out = codes.copy()
    out[over] = ord(ambiguous)
    return out


def effective_lengths(branch_length, full_length):
    return np.maximum(ttconf.MIN_BRANCH_LENGTH / full_length, np.asarray(branch_length, dtype=float))


def prepare(parent, branch_length, tip_chars, gtr, ambiguous, do_fill_overhangs=True, compress=True, engine=None):
    """tip_chars: (n_tips, L) 'S1' or uint8, rows in ascending tip node index.
    With an engine, the overhang fill and the compression run on the device
    (Engine.compress_alignment), otherwise in numpy; the results are equal.

    Returns a dict of DeviceTree inputs plus the compression maps.
    """
    codes = np.ascontiguousarray(tip_chars).view(np.uint8).reshape(tip_chars.shape)
    L = codes.shape[1]
    fill = do_fill_overhangs and ambiguous is not None
    if compress and engine is not None:
        pat, comp, mult, _ = engine.compress_alignment(codes, ambiguous, fill_overhangs=fill)
    else:
        if fill:
            codes = fill_overhangs(codes, ambiguous)
        if compress:
            pat, comp = compress_columns(codes, ambiguous)
        else:
            pat, comp = np.arange(L), codes
        mult = np.bincount(pat, minlength=comp.shape[1]).astype(np.float64)
    return {
        "parent": np.asarray(parent, dtype=np.int32),
        "t": effective_lengths(branch_length, L),
        "tip_codes": np.ascontiguousarray(comp),
        "lut": code_lut(gtr.profile_map, gtr.n_states),
        "multiplicity": mult,
        "full_to_compressed": pat,
        "pi": gtr.Pi,
        "eigvals": gtr.eigenvals,
        "v": gtr.v,
        "v_inv": gtr.v_inv,
        "mu": gtr.mu,
        "W": gtr.W,
        "n_states": gtr.n_states,
        "full_length": L,
    }


def device_tree(engine, prep):
    return DeviceTree(engine, prep["parent"], prep["t"], prep["tip_codes"], prep["lut"], prep["multiplicity"],
                      prep["pi"], prep["eigvals"], prep["v"], prep["v_inv"], prep["mu"])

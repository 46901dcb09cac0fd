"""Writes tests/golden/profile_maps.json: the literal alphabets and profile
maps of the reference's treetime/seq_utils.py (lines 19-122), read from
the source text with ``ast`` (the reference is never imported).

Every ``np.array([...], dtype=...)`` in the two dict literals is read as its
list argument; nothing else of the file is evaluated.

Run from the repository root:
    python tests/golden/make_profile_maps.py [/root/reference/treetime/seq_utils.py]
"""
import ast
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "profile_maps.json")
DEFAULT_SRC = "/root/reference/treetime/seq_utils.py"


def _value(node):
    """Literal value of a dict entry: np.array(<list>, ...) -> list."""
    if isinstance(node, ast.Call) and getattr(node.func, "attr", None) == "array":
        return ast.literal_eval(node.args[0])
    return ast.literal_eval(node)


def _dict(node):
    if not isinstance(node, ast.Dict):
        raise ValueError("expected a dict literal")
    out = {}
    for k, v in zip(node.keys, node.values):
        key = ast.literal_eval(k)
        out[key] = _dict(v) if isinstance(v, ast.Dict) else _value(v)
    return out


def read_tables(path):
    tree = ast.parse(open(path).read(), filename=path)
    found = {}
    for stmt in tree.body:
        if isinstance(stmt, ast.Assign) and len(stmt.targets) == 1 and isinstance(stmt.targets[0], ast.Name):
            name = stmt.targets[0].id
            if name in ("alphabets", "profile_maps", "alphabet_synonyms"):
                found[name] = _dict(stmt.value)
    missing = {"alphabets", "profile_maps", "alphabet_synonyms"} - set(found)
    if missing:
        raise ValueError("not found in %s: %s" % (path, sorted(missing)))
    return found


def main(src=DEFAULT_SRC):
    tables = read_tables(src)
    tables["source"] = "treetime/seq_utils.py (TreeTime 0.12.1), lines 5-122"
    with open(OUT, "w") as f:
        json.dump(tables, f, indent=1, sort_keys=True)
    print("wrote", OUT, {k: len(v) for k, v in tables["profile_maps"].items()})


if __name__ == "__main__":
    main(*sys.argv[1:])

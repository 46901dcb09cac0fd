"""Minimal rooted-tree objects with the Bio.Phylo API subset TreeAnc uses.

Biopython is not a dependency here.  ``Tree``/``Clade`` offer what the
reference's TreeAnc touches on ``Bio.Phylo`` trees (find_clades with
preorder/postorder/level order, get_terminals, get_nonterminals,
count_terminals, ladderize, total_branch_length) plus a Newick reader and
writer.  Any object with the same duck-typed surface (a real Bio.Phylo tree
included) is accepted by ``TreeAnc``.
"""
import re

import numpy as np


class Clade:
    def __init__(self, name=None, branch_length=None, clades=None, confidence=None):
        self.name = name
        self.branch_length = branch_length
        self.clades = list(clades) if clades else []
        self.confidence = confidence

    def is_terminal(self):
        return not self.clades

    def __iter__(self):
        return iter(self.clades)

    def __len__(self):
        return len(self.clades)

    def __repr__(self):
        return f"Clade(name={self.name!r}, branch_length={self.branch_length!r})"

    def find_clades(self, order="preorder"):
        if order == "preorder":
            stack = [self]
            while stack:
                n = stack.pop()
                yield n
                stack.extend(reversed(n.clades))
        elif order == "postorder":
            stack = [(self, False)]
            while stack:
                n, done = stack.pop()
                if done or not n.clades:
                    yield n
                else:
                    stack.append((n, True))
                    stack.extend((c, False) for c in reversed(n.clades))
        elif order == "level":
            queue = [self]
            i = 0
            while i < len(queue):
                n = queue[i]
                i += 1
                yield n
                queue.extend(n.clades)
        else:
            raise ValueError(f"unknown traversal order {order!r}")

    def get_terminals(self, order="preorder"):
        return [c for c in self.find_clades(order) if not c.clades]

    def get_nonterminals(self, order="preorder"):
        return [c for c in self.find_clades(order) if c.clades]

    def count_terminals(self):
        return sum(1 for c in self.find_clades() if not c.clades)

    def total_branch_length(self):
        return sum(c.branch_length or 0.0 for c in self.find_clades())

    def ladderize(self, reverse=False):
        """Sort children by their number of tips, smallest first (stable),
        recursively -- Bio.Phylo's ladderize, which TreeAnc.prepare_tree
        applies (treeanc.py:456-457)."""
        counts = {}
        for n in self.find_clades("postorder"):
            counts[id(n)] = 1 if not n.clades else sum(counts[id(c)] for c in n.clades)
        for n in self.find_clades("preorder"):
            if n.clades:
                n.clades.sort(key=lambda c: counts[id(c)], reverse=reverse)


class Tree:
    def __init__(self, root=None, rooted=True):
        self.root = root if root is not None else Clade()
        self.rooted = rooted

    # delegate the traversal API to the root clade
    def find_clades(self, order="preorder"):
        return self.root.find_clades(order)

    def get_terminals(self, order="preorder"):
        return self.root.get_terminals(order)

    def get_nonterminals(self, order="preorder"):
        return self.root.get_nonterminals(order)

    def count_terminals(self):
        return self.root.count_terminals()

    def total_branch_length(self):
        return self.root.total_branch_length()

    def ladderize(self, reverse=False):
        self.root.ladderize(reverse)

    def __iter__(self):
        return iter(self.root.clades)

    @classmethod
    def from_newick(cls, text):
        return cls(parse_newick(text))

    def to_newick(self, fmt="%1.8g"):
        return write_newick(self.root, fmt) + ";"


_TOKEN = re.compile(r"\s*(\(|\)|,|;|:|'[^']*'|\[[^\]]*\]|[^()\[\],:;']+)")


def parse_newick(text):
    """Parse one Newick tree.  Numeric labels of internal nodes are read as
    confidence values (as Bio.Phylo does); other labels become names."""
    tokens = [t for t in _TOKEN.findall(text.strip()) if t.strip() and not t.startswith("[")]
    pos = 0

    def label_to(node, lab, internal):
        lab = lab.strip()
        if lab.startswith("'") and lab.endswith("'"):
            node.name = lab[1:-1]
            return
        if internal:
            try:
                node.confidence = float(lab)
                return
            except ValueError:
                pass
        node.name = lab

    def parse_clade():
        nonlocal pos
        node = Clade()
        internal = False
        if tokens[pos] == "(":
            internal = True
            pos += 1
            while True:
                node.clades.append(parse_clade())
                if tokens[pos] == ",":
                    pos += 1
                    continue
                if tokens[pos] == ")":
                    pos += 1
                    break
                raise ValueError(f"unexpected token {tokens[pos]!r} in Newick")
        if pos < len(tokens) and tokens[pos] not in ("(", ")", ",", ";", ":"):
            label_to(node, tokens[pos], internal)
            pos += 1
        if pos < len(tokens) and tokens[pos] == ":":
            pos += 1
            node.branch_length = float(tokens[pos])
            pos += 1
        return node

    root = parse_clade()
    if pos < len(tokens) and tokens[pos] != ";":
        raise ValueError("trailing tokens after Newick tree")
    return root


def write_newick(node, fmt="%1.8g"):
    s = ""
    if node.clades:
        s = "(" + ",".join(write_newick(c, fmt) for c in node.clades) + ")"
    if node.name:
        s += node.name
    if node.branch_length is not None:
        s += ":" + (fmt % node.branch_length)
    return s


def flatten(root):
    """Preorder numbering of a tree: (nodes, parent[int32], branch_length[f64]).

    Root is node 0; the children of every node get ascending indices in
    their list order, which is the child order the engine uses.
    """
    nodes = list(root.find_clades("preorder"))
    idx = {id(n): i for i, n in enumerate(nodes)}
    parent = np.full(len(nodes), -1, dtype=np.int32)
    bl = np.zeros(len(nodes), dtype=np.float64)
    for i, n in enumerate(nodes):
        bl[i] = n.branch_length if n.branch_length else 0.0
        for c in n.clades:
            parent[idx[id(c)]] = i
    return nodes, parent, bl

"""Tree XML load/save in the reference's format (src/padenti/tree_impl.hpp:
103-234, Boost property_tree write_xml): nonexistent nodes omitted, numbers
with 6 significant digits unless ``lossless`` (float max_digits10 = 9)."""
import re

import numpy as np

from . import Tree


def _fmt(v, lossless):
    if isinstance(v, (np.integer, int)):
        return str(int(v))
    return ("%.9g" if lossless else "%g") % float(v)


def save(tree, path, idx=-1, lossless=False):
    tid = tree.id if idx == -1 else idx
    parts = ['<?xml version="1.0" encoding="utf-8"?>\n<Trees><TreeNumber>1</TreeNumber><MaxDepth>%d</MaxDepth>'
             '<Tree%d>' % (tree.depth, tid)]
    short = np.issubdtype(tree.features.dtype, np.integer)
    for n in range(tree.n_nodes):
        lc = int(tree.left_children[n])
        if lc == -2:
            continue
        parts.append("<Node%d><LeftChild>%d</LeftChild>" % (n, lc))
        if lc != -1:
            vals = [int(v) if short else v for v in tree.features[n]] + [
                int(tree.thresholds[n]) if short else tree.thresholds[n]]
            parts.append("<SplitParameters>%s</SplitParameters>" % " ".join(_fmt(v, lossless) for v in vals))
        parts.append("<Histogram>%s</Histogram></Node%d>" % (
            " ".join(_fmt(v, lossless) for v in tree.posteriors[n]), n))
    parts.append("</Tree%d></Trees>" % tid)
    with open(path, "w") as f:
        f.write("".join(parts))


def load(path, feature_kind, n_classes, idx=-1, tree_id=0):
    text = open(path).read()
    depth = int(re.search(r"<MaxDepth>\s*(\d+)\s*</MaxDepth>", text).group(1))
    tid = tree_id if idx == -1 else idx
    m = re.search(r"<Tree%d>(.*)</Tree%d>" % (tid, tid), text, re.S)
    tree = Tree(tid, depth, feature_kind, n_classes)
    if m is None:
        return tree
    body = m.group(1)
    for nm in re.finditer(r"<Node(\d+)>(.*?)</Node\1>", body, re.S):
        n, inner = int(nm.group(1)), nm.group(2)
        lc = int(re.search(r"<LeftChild>\s*(-?\d+)\s*</LeftChild>", inner).group(1))
        tree.left_children[n] = lc
        h = re.search(r"<Histogram>(.*?)</Histogram>", inner, re.S)
        if h:
            tree.posteriors[n] = np.array(h.group(1).split(), dtype=np.float32)[:n_classes]
        if lc >= 0:
            sp = re.search(r"<SplitParameters>(.*?)</SplitParameters>", inner, re.S).group(1).split()
            fd = tree.features.shape[1]
            tree.features[n] = np.array(sp[:fd], dtype=np.float64).astype(tree.features.dtype)
            tree.thresholds[n] = np.array([sp[fd]], dtype=np.float64).astype(tree.thresholds.dtype)[0]
    return tree

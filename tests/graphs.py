"""Small fixed graphs shared by the CPU and GPU tests (arc arrays in SoA
insertion order)."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WIKI_VOTE = os.path.join(HERE, "golden", "wiki_vote_sample.txt")


def arcs(pairs):
    src = np.array([p[0] for p in pairs], dtype=np.uint32)
    dst = np.array([p[1] for p in pairs], dtype=np.uint32)
    return src, dst


def undirected(pairs):
    out = []
    for u, v in pairs:
        out.append((u, v))
        out.append((v, u))
    return arcs(out)


def wiki_vote():
    """The reference's fixture test/fixtures/wiki_vote_sample.txt as a
    directed graph with ids in first-seen order (read_edgelist semantics)."""
    ids = {}
    pairs = []
    with open(WIKI_VOTE) as fh:
        for line in fh:
            t = line.split()
            if len(t) < 2 or t[0].startswith("#"):
                continue
            for lab in t[:2]:
                if lab not in ids:
                    ids[lab] = len(ids)
            pairs.append((ids[t[0]], ids[t[1]]))
    s, d = arcs(pairs)
    return len(ids), s, d


def random_multigraph(rng, n, m, self_loops=True):
    src = rng.integers(0, n, size=m).astype(np.uint32)
    dst = rng.integers(0, n, size=m).astype(np.uint32)
    if not self_loops:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    return n, src, dst

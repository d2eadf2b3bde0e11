"""Writes nx_modularity.json: networkx's modularity of given partitions on small
weighted graphs (self-loops, zero weights, resolution != 1).  networkx's
`modularity` is the textbook Newman-Girvan formula (a library routine, not the
reference); these vectors pin the oracle's FP64 modularity.

Run from the repository root:  python tests/golden/make_nx_fixtures.py
"""
import json
import os

import networkx as nx
import numpy as np
from networkx.algorithms.community import modularity

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nx_modularity.json")


def main():
    rng = np.random.Generator(np.random.PCG64(12345))
    cases = []
    for case in range(40):
        n = int(rng.integers(2, 30))
        e = int(rng.integers(1, 3 * n))
        src = rng.integers(0, n, e).tolist()
        dst = rng.integers(0, n, e).tolist()
        # merge duplicates so the edge list is a simple graph with loops
        seen = {}
        for a, b in zip(src, dst):
            key = (min(a, b), max(a, b))
            seen[key] = seen.get(key, 0) + 1
        edges = sorted(seen)
        if case % 2:
            w = [float(np.float32(rng.uniform(0.5, 4.0))) for _ in edges]
        else:
            w = [float(int(rng.integers(1, 5))) for _ in edges]
        comm = rng.integers(0, max(1, n // 3), n).tolist()
        gamma = [1.0, 0.5, 2.0][case % 3]
        G = nx.Graph()
        G.add_nodes_from(range(n))
        for (a, b), ww in zip(edges, w):
            G.add_edge(a, b, weight=ww)
        groups = {}
        for v, c in enumerate(comm):
            groups.setdefault(c, set()).add(v)
        q = modularity(G, list(groups.values()), weight="weight", resolution=gamma)
        cases.append({"n": n, "src": [a for a, _ in edges], "dst": [b for _, b in edges],
                      "w": w, "float": bool(case % 2), "comm": comm, "gamma": gamma, "q": q})
    with open(OUT, "w") as f:
        json.dump({"generator": "networkx %s modularity" % nx.__version__, "cases": cases}, f)


if __name__ == "__main__":
    main()

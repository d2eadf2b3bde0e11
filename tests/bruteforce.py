"""Pure-Python statement of the reference's Louvain semantics for tiny graphs.

Written independently of oracle/louvain_oracle.c (dicts and lists, no sorting
tricks) so that it can pin the oracle on small inputs.  Weights are exact
Python integers in units; gains use the same FP64 expression the reference
semantics define, W - ((gamma*deg[c])*f) with f = k_i/(2m).

Each step cites the reference lines it states:
  CSR and m        LouvainCUDA/NeighbourCompute.cu:46,66-68,82
  k_i, loop_i      LouvainCUDA/InitKernel.cu:43-56,70-76,119-125
  one pass         LouvainCUDA/OneLevelKernel.cu:43-49,265-316,330,358,407-421
  convergence      LouvainCUDA/OneLevelKernel.cu:528,875-896
  merge-isolated   LouvainCUDA/OneLevelKernel.cu:138-171,921-944
  renumber         LouvainCUDA/RenumberKernel.cu:117-133,191,202-203
  coarsen          LouvainCUDA/InducedGraphKernel.cu:38-45,91-99
  level loop       LouvainCUDA/LouvainCPP.cpp:334-507
"""
from __future__ import annotations

from fractions import Fraction


def unit_scale(weights):
    """Smallest s >= 0 with every w * 2^s an integer (exact rationals)."""
    s = 0
    for w in weights:
        f = Fraction(float(w))
        while (f * (1 << s)).denominator != 1:
            s += 1
    return s


def to_units(weights, s):
    return [int(Fraction(float(w)) * (1 << s)) for w in weights]


def q_value(sum_in, sum_deg2, m, gamma):
    if m <= 0:
        return 0.0
    dm = float(m)
    hi, lo = divmod(sum_deg2, 1 << 64)
    d2 = float(hi) * 18446744073709551616.0 + float(lo)
    return sum_in / dm - gamma * (d2 / ((4.0 * dm) * dm))


def one_level(N, EL, gamma, lower, min_gain, trace):
    m = sum(w for _, _, w in EL)
    # row entries in the reference's sorted order: by target, stable
    rows = [[] for _ in range(N)]
    entries = [(s, t, w) for s, t, w in EL] + [(t, s, w) for s, t, w in EL]
    for s, t, w in entries:
        rows[s].append((t, w))
    for i in range(N):
        rows[i].sort(key=lambda e: e[0])  # Python sort is stable
    k = [0] * N
    loop = [0] * N
    for i in range(N):
        for j, w in rows[i]:
            v = 1 if w == 0 else w
            if j == i:
                v *= 2
                loop[i] = w
            k[i] += v
    C = list(range(N))
    deg = list(k)
    inn = list(loop)
    size = [1] * N

    def Q():
        return q_value(sum(inn), sum(d * d for d in deg), m, gamma)

    q_init = Q()
    curQ = -1.0
    passes = []
    while True:
        best_c = [None] * N
        ngroups = [0] * N
        nxt = list(C)
        ddeg = [0] * N
        din = [0] * N
        dsize = [0] * N
        for i in range(N):
            if not rows[i]:
                continue
            W = {}
            for j, w in rows[i]:
                c = C[j]
                W[c] = W.get(c, 0) + (0 if j == i else w)
            ngroups[i] = len(W)
            f = k[i] / (2.0 * m) if m else float("inf")
            bc, bg, bw = C[i], 0.0, 0
            for c in sorted(W):
                if W[c] == 0:
                    continue
                g = W[c] - ((gamma * deg[c]) * f)
                if g > bg or (g == bg and bg != 0 and c < bc):
                    bc, bg, bw = c, g, W[c]
            best_c[i] = bc
            own = C[i]
            if size[own] == 1 and size[bc] == 1 and bc > own:
                bc = own
            if bc != own:
                nxt[i] = bc
                ddeg[own] -= k[i]
                ddeg[bc] += k[i]
                din[own] -= bw + loop[i]
                din[bc] += bw + loop[i]
                dsize[own] -= 1
                dsize[bc] += 1
        qp = Q()
        if qp - curQ < min_gain:
            # post-loop: discard deltas, keep proposals, merge isolated
            ddeg = [0] * N
            din = [0] * N
            dsize = [0] * N
            for i in range(N):
                if ngroups[i] == 0:
                    continue
                own = C[i]
                if size[own] != 1 or ngroups[i] != 1:
                    continue
                c0 = best_c[i]
                if c0 != own and own < c0:
                    nxt[i] = c0
                    ddeg[own] -= k[i]
                    ddeg[c0] += k[i]
                    din[own] -= loop[i]
                    din[c0] += loop[i]
                    dsize[own] -= 1
                    dsize[c0] += 1
            for c in range(N):
                deg[c] += ddeg[c]
                inn[c] += din[c]
                size[c] += dsize[c]
            passes.append((qp, list(nxt)))
            C = nxt
            break
        passes.append((qp, list(nxt)))
        curQ = max(qp, lower)
        for c in range(N):
            deg[c] += ddeg[c]
            inn[c] += din[c]
            size[c] += dsize[c]
        C = nxt
    return dict(m=m, k=k, loop=loop, q_init=q_init, q=Q(), passes=passes, C=C, rows=rows)


def renumber(N, C, ids):
    order = []
    first = {}
    for i in range(N):
        if C[i] not in first:
            first[C[i]] = len(first)
    newc = [first[C[i]] for i in range(N)]
    perm = sorted(range(N), key=lambda i: (C[i], i))
    node = [ids[p] for p in perm]
    comm = [newc[p] for p in perm]
    return newc, len(first), node, comm


def coarsen(EL, newc):
    acc = {}
    for s, t, w in EL:
        key = (newc[s], newc[t])
        acc[key] = acc.get(key, 0) + (1 if w == 0 else w)
    return [(a, b, acc[(a, b)]) for (a, b) in sorted(acc)]


def louvain(n, edges, ids=None, gamma=1.0, min_gain=1e-7):
    """edges: list of (s, t, w_units) over positions.  Returns per-level dicts
    and the final assignment of every position."""
    if ids is None:
        ids = list(range(n))
    EL = list(edges)
    N = n
    levels = []
    mod = None
    while True:
        lvl = one_level(N, EL, gamma, 0.0 if mod is None else mod, min_gain, True)
        if mod is None:
            mod = lvl["q_init"]
            lvl = one_level(N, EL, gamma, mod, min_gain, True)
        newc, K, node, comm = renumber(N, lvl["C"], ids if not levels else list(range(N)))
        lvl.update(newc=newc, K=K, node=node, comm=comm, EL=EL, N=N, lower=mod)
        levels.append(lvl)
        if len(levels) > 1 and lvl["q"] - mod < min_gain:
            break
        mod = lvl["q"]
        EL = coarsen(EL, newc)
        N = K
    final = []
    for i in range(n):
        f = levels[0]["newc"][i]
        for L in levels[1:]:
            f = L["newc"][f]
        final.append(f)
    return levels, final

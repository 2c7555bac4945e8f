"""Peel sequence for a pedigree (host pre-computation).

Produces the operation table the peel kernels consume, with the reference's
semantics (peel_sequence_generator.cc, peeling.h):
  * the graph: each person is adjacent to its parents, children and mates
    (build_simple_graph, :188-214);
  * eliminating a node joins its current neighbours (eliminate_node, :354-367);
  * an operation's type follows set_type (:131-164): LAST when everyone else
    is peeled, CHILD when neither parent is peeled, PARENT when some child and
    some mate are unpeeled, otherwise PARTNER;
  * consumed tables are the earlier unused tables whose cutset lies inside the
    op's cutset plus its peel node, first match first (find_prev_functions,
    :13-41); PARENT ops list their cutset members that are its children
    (peeling.h:145-163).
The order itself is searched differently: greedy minimum-cutset elimination
with seeded random tie-breaks, keeping the cheapest legitimate order (the
is_legit condition, :396-425) over a number of tries.  Any legitimate order
gives the same likelihood; only the cost (sum of 4^|cutset|) differs.
"""
from __future__ import annotations

from typing import List

import numpy as np

from .case import PeelOps, OP_CHILD, OP_PARENT, OP_PARTNER, OP_LAST


def _graph(N, mother, father):
    nb = [[] for _ in range(N)]
    children = [[] for _ in range(N)]
    mates = [set() for _ in range(N)]
    for i in range(N):
        if mother[i] >= 0:
            children[mother[i]].append(i)
            children[father[i]].append(i)
            mates[mother[i]].add(int(father[i]))
            mates[father[i]].add(int(mother[i]))

    def add(lst, x):
        if x not in lst:
            lst.append(x)

    for i in range(N):
        if mother[i] >= 0:
            add(nb[i], int(mother[i]))
            add(nb[i], int(father[i]))
        for c in children[i]:
            add(nb[i], c)
        for m in sorted(mates[i]):
            add(nb[i], m)
    return nb, children, mates


def _eliminate(cuts, node):
    cs = list(cuts[node])
    for a in cs:
        if node in cuts[a]:
            cuts[a].remove(node)
        for b in cs:
            if a != b and b not in cuts[a]:
                cuts[a].append(b)


def _legit(order, N, mother, father, children, mates):
    peeled = np.zeros(N, bool)
    for n in order:
        if mother[n] >= 0 and not (peeled[mother[n]] or peeled[father[n]]):
            if children[n] and any(not peeled[m] for m in mates[n]) and any(not peeled[c] for c in children[n]):
                return False
        peeled[n] = True
    return True


def _greedy(nb, rng):
    N = len(nb)
    cuts = [list(x) for x in nb]
    alive = set(range(N))
    order = []
    cost = 0
    while alive:
        sizes = {}
        for n in alive:
            sizes.setdefault(len(cuts[n]), []).append(n)
        best = min(sizes)
        cand = sorted(sizes[best])
        n = cand[int(rng.integers(len(cand)))]
        order.append(n)
        cost += 4 ** len(cuts[n])
        _eliminate(cuts, n)
        alive.remove(n)
    return order, cost


def finalise(order: List[int], N, mother, father) -> PeelOps:
    nb, children, mates = _graph(N, mother, father)
    cuts = [list(x) for x in nb]
    peeled = np.zeros(N, bool)
    used = []
    types, peels, cutsets, prevs, kids = [], [], [], [], []
    for n in order:
        cut = list(cuts[n])
        if peeled.sum() == N - 1:
            t = OP_LAST
        elif mother[n] >= 0 and not (peeled[mother[n]] or peeled[father[n]]):
            t = OP_CHILD
        elif children[n] and any(not peeled[m] for m in mates[n]) and any(not peeled[c] for c in children[n]):
            t = OP_PARENT
        else:
            t = OP_PARTNER
        scope = set(cut) | {n}
        pv = []
        while True:
            hit = -1
            for j in range(len(cutsets)):
                if not used[j] and set(cutsets[j]) <= scope:
                    used[j] = True
                    hit = j
                    break
            if hit < 0:
                break
            pv.append(hit)
        ch = [c for c in cut if c in children[n]] if t == OP_PARENT else []
        types.append(t)
        peels.append(n)
        cutsets.append(cut)
        prevs.append(pv)
        kids.append(ch)
        used.append(False)
        _eliminate(cuts, n)
        peeled[n] = True
    return PeelOps.from_lists(types, peels, cutsets, prevs, kids)


def build(N, F, mother, father, seed: int = 1, tries: int = 64) -> PeelOps:
    nb, children, mates = _graph(N, mother, father)
    rng = np.random.default_rng(seed)
    best = None
    for _ in range(tries):
        order, cost = _greedy(nb, rng)
        if not _legit(order, N, mother, father, children, mates):
            continue
        if best is None or cost < best[1]:
            best = (order, cost)
    if best is None:
        raise RuntimeError("no legitimate peel order found")
    return finalise(best[0], N, mother, father)

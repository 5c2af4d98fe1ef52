"""Second, independently written simulator of the recolor rules (pure Python,
tiny graphs only).  It shares no code with oracle/dc_oracle.c: the graph is a
multiset of edge keys, every round is computed as a function
new_state = f(old_state), and sets are Python sets.

Rules (reference paths relative to the reference repo root):
  independent: DynColorHeuristicIndependent/DynColorHeuristicIndependent/
               gpuFunctions_undir.cuh:147-284, Color_main.cu:146-202
  ldf:         DynColorHeuristicBased/DynColorHeuristicBased/
               gpuFunctions_undir.cuh:104-129,222-314, Color_main.cu:160-224
"""
from collections import Counter


def _key(a, b):
    return (a, b) if a < b else (b, a)


class BruteGraph:
    def __init__(self, n, edges, n_owned=None):
        self.n = n
        self.n_owned = n if n_owned is None else n_owned
        self.multi = Counter(_key(a, b) for a, b in edges)
        self.prop = [0] * n
        for a, b in edges:
            self.prop[a] += 1
            self.prop[b] += 1

    def nbrs(self, v):
        out = set()
        for (a, b), c in self.multi.items():
            if c <= 0:
                continue
            if a == v:
                out.add(b)
            elif b == v:
                out.add(a)
        return out

    def frozen(self, v):
        return v >= self.n_owned


def mex(values):
    s = set(values)
    m = 0
    while m in s:
        m += 1
    return m


def run_batch(g, colors, lines, rule):
    """Apply one batch and recolor; returns (colors, trace)."""
    colors = list(colors)
    s0 = list(colors)
    partition = g.n_owned < g.n
    dels, xdels, ins, xins = [], [], [], []
    for a, b, is_ins in lines:
        if g.frozen(a):
            a, b = b, a
        cross = g.frozen(b)
        if is_ins:
            g.multi[_key(a, b)] += 1
            g.prop[a] += 1
            g.prop[b] += 1
            if cross:
                xins.append((a, b))
            elif partition or s0[a] == s0[b]:
                ins.append((a, b))
        else:
            g.prop[a] -= 1
            g.prop[b] -= 1
            (xdels if cross else dels).append((a, b))
    for a, b in dels + xdels:   # deletion wins over any insertion of the pair
        g.multi[_key(a, b)] = 0

    N = {v: g.nbrs(v) for v in range(g.n)}

    def before(u, v):  # u precedes v in the ldf order
        return (g.prop[u], -u) > (g.prop[v], -v)

    trace = []
    prev = [-1] * g.n
    marked = set()
    change = False
    if rule == 0:
        for kind, lst in ((1, dels), (4, xdels)):
            if not lst:
                continue
            T = set()
            for a, b in lst:
                t = a if colors[a] > colors[b] else b
                if not g.frozen(t):
                    T.add(t)
            new = {t: mex(colors[w] for w in N[t]) for t in T}
            ch = [t for t in T if new[t] != colors[t]]
            for t in ch:
                prev[t] = colors[t]
                colors[t] = new[t]
            marked |= set(ch)
            trace.append([kind, len(T), len(ch), 0])
            change = change or bool(ch)
        for kind, lst in ((2, ins), (5, xins)):
            if not lst:
                continue
            T = set()
            for a, b in lst:
                if colors[a] != colors[b]:
                    continue
                ka = float("inf") if g.frozen(a) else len({colors[w] for w in N[a]})
                kb = float("inf") if g.frozen(b) else len({colors[w] for w in N[b]})
                t = a if ka < kb else b
                if not g.frozen(t):
                    T.add(t)
            new = {t: mex(colors[w] for w in N[t]) for t in T}
            nch = sum(1 for t in T if new[t] != colors[t])
            for t in T:
                prev[t] = colors[t]
                colors[t] = new[t]
            marked |= T
            trace.append([kind, len(T), nch, 0])
            change = change or bool(T)
    else:
        A = dels + [e for e in ins]
        if A:
            T = set()
            for a, b in A:
                T.add(a)
                T.add(b)
            new = {t: mex(colors[w] for w in N[t] if before(w, t)) for t in T}
            ch = [t for t in T if new[t] != colors[t]]
            for t in ch:
                colors[t] = new[t]
            marked = set(ch)
            trace.append([3, len(T), len(ch), 0])
            change = bool(ch)
    L = marked
    rounds = 0
    while change:
        rounds += 1
        if rounds > 10000:
            raise RuntimeError("brute: no convergence")
        Cset = set()
        for v in L:
            for w in N[v]:
                if rule == 0:
                    if g.frozen(w):
                        continue
                    if (colors[v] == colors[w] and v < w) or prev[v] < colors[w]:
                        Cset.add(w)
                else:
                    if colors[v] == colors[w]:
                        Cset.add(w if before(v, w) else v)
                    if before(v, w):
                        Cset.add(w)
        if rule == 0:
            new = {w: mex(colors[x] for x in N[w]) for w in Cset}
            nch = sum(1 for w in Cset if new[w] != colors[w])
            for w in Cset:
                prev[w] = colors[w]
                colors[w] = new[w]
            change = bool(Cset)
        else:
            new = {w: mex(colors[x] for x in N[w] if before(x, w)) for w in Cset}
            nch = sum(1 for w in Cset if new[w] != colors[w])
            for w in Cset:
                colors[w] = new[w]
            change = nch > 0
        trace.append([10, len(L), len(Cset), nch])
        L = Cset
    return colors, trace


def greedy(g, rule, colors=None):
    """first-fit greedy in ID order (rule 0) or ldf priority order (rule 1)"""
    colors = list(colors) if colors is not None else [0] * g.n
    order = list(range(g.n_owned))
    if rule == 1:
        order.sort(key=lambda v: (-g.prop[v], v))
    done = set(range(g.n_owned, g.n))
    for v in order:
        colors[v] = mex(colors[w] for w in g.nbrs(v) if w in done)
        done.add(v)
    return colors


def monochromatic(g, colors):
    bad = 0
    for (a, b), c in g.multi.items():
        if c > 0 and colors[a] == colors[b]:
            bad += c
    return bad

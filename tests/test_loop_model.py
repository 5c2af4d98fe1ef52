"""The incremental independent-rule loop, as a CPU model, against the direct
re-derivation of the reference loop (HI/Color_main.cu:185-202 with
findEligibleNeighbors and recolorNeighbor, HI/gpuFunctions_undir.cuh:229-284).

The HIP loop keeps K[w] = #{slots (w, v): v in L, the marking predicate holds
for (v, w)} and moves it by deltas of the vertices whose state changed (D),
with a full recount by pull when D is large.  This model follows the same
steps on small random graphs (multi-edges, frozen ghosts) and checks, after
every round, that K equals a recount and that the trace and colours equal the
direct loop.  A mistake in a delta term, in the D membership rule or in the
candidate set fails here on the CPU."""
import random

import pytest


def marks(v, cv, pv, w, cw):
    return (cv == cw and v < w) or pv < cw


def mex(c, nbrs):
    s = {c[w] for w in nbrs}
    m = 0
    while m in s:
        m += 1
    return m


def direct(n_owned, adj, col, prev, L0):
    c, p, L, tr = col[:], prev[:], sorted(set(L0)), []
    while True:
        C = sorted({w for v in L for w in adj[v]
                    if w < n_owned and marks(v, c[v], p[v], w, c[w])})
        new = {w: mex(c, adj[w]) for w in C}
        tr.append((len(L), len(C), sum(new[w] != c[w] for w in C)))
        for w in C:
            p[w], c[w] = c[w], new[w]
        if not C:
            return c, p, tr
        L = C


def incremental(n, n_owned, adj, col, prev, L0, pull_frac):
    c, p, K = col[:], prev[:], [0] * n
    total = sum(len(a) for a in adj)
    L = set(L0)
    cand = set()
    for v in L:
        for w in adj[v]:
            if w < n_owned and marks(v, c[v], p[v], w, c[w]):
                K[w] += 1
                cand.add(w)
    cand |= L
    tr = []
    while True:
        C = {w for w in cand if K[w] > 0}
        new = {w: mex(c, adj[w]) for w in C}
        old, D = {}, set()
        for w in C:
            so, po, inl = c[w], p[w], w in L
            p[w], c[w] = so, new[w]
            if not inl or po != so or new[w] != so:
                old[w] = (so, po, inl)
                K[w] = 0
                D.add(w)
        for x in L - C:
            old[x] = (c[x], p[x], True)
            D.add(x)
        tr.append((len(L), len(C), sum(old[w][0] != c[w] for w in C if w in old)))
        if not C:
            break
        R = set()
        if sum(len(adj[x]) for x in D) > pull_frac * total:
            # heavy round: every count by pull
            for y in range(n_owned):
                K[y] = sum(1 for x in adj[y] if x in C and marks(x, c[x], p[x], y, c[y]))
                if K[y] and y not in C:
                    R.add(y)
        else:
            for x in D:
                so, po, inl = old[x]
                inc = x in C
                cnt = 0
                for y in adj[x]:
                    if y >= n_owned:
                        continue
                    if y not in D:
                        to = inl and marks(x, so, po, y, c[y])
                        tn = inc and marks(x, c[x], p[x], y, c[y])
                        if to != tn:
                            K[y] += 1 if tn else -1
                            if tn:
                                R.add(y)
                    if y in C and marks(y, c[y], p[y], x, c[x]):
                        cnt += 1
                K[x] = cnt
        for w in range(n_owned):
            assert K[w] == sum(1 for v in adj[w] if v in C and marks(v, c[v], p[v], w, c[w]))
        L, cand = C, C | D | R
    assert all(k == 0 for k in K)
    return c, p, tr


@pytest.mark.parametrize("pull_frac", [2.0, 0.0, 0.05])
def test_incremental_loop_equals_direct(pull_frac):
    rng = random.Random(11)
    for trial in range(400):
        n = rng.randint(2, 24)
        n_owned = n if rng.random() < 0.7 else rng.randint(1, n)
        adj = [[] for _ in range(n)]
        pr = rng.choice([0.1, 0.3, 0.6])
        for a in range(n):
            for b in range(a + 1, n):
                if a >= n_owned and b >= n_owned:
                    continue
                k = (2 if rng.random() < 0.1 else 1) if rng.random() < pr else 0
                for _ in range(k):
                    adj[a].append(b)
                    adj[b].append(a)
        col = [rng.randint(0, 5) for _ in range(n)]
        prev = [rng.choice([-1] + list(range(6))) for _ in range(n)]
        L0 = [v for v in range(n_owned) if rng.random() < 0.3]
        want = direct(n_owned, adj, col, prev, L0)
        got = incremental(n, n_owned, adj, col, prev, L0, pull_frac)
        assert got[2] == want[2], trial
        assert got[0] == want[0] and got[1] == want[1], trial

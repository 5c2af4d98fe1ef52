"""The incremental heuristic-based (ldf) loop, as a CPU model, against the
direct re-derivation of the reference loop (HB/Color_main.cu:203-224 with
findEligibleNeighbors and recolorNeighbor, HB/gpuFunctions_undir.cuh:259-314).

The HIP loop keeps, with the priorities fixed for the batch,
  K[w]  = predecessor slots of w whose neighbour is in L,
  cc[v] = predecessor slots of v whose neighbour has v's colour,
  pm[v] = mex of v's predecessors' colours, cached (-1: stale),
so C = {w : K[w] > 0} u {v in L : cc[v] > 0}.  Between rounds the rows of
D = L xor C add +-1 to their successors' K, and a vertex x that changes
colour a -> b moves its successors' cc and drops a successor's cached mex m
only when the change can move it (a < m or b == m).  This model follows those
steps on small random graphs (multi-edges, ties in the priorities) and checks
after every round that K and cc equal recounts, that every cached mex still
held equals the true mex, and that the trace and colours equal the direct
loop.  Dropping too few mexes, a wrong delta or a wrong membership rule fails
here on the CPU."""
import random

import pytest


def mex(s):
    m = 0
    while m in s:
        m += 1
    return m


def prior(prop, u, v):
    """u precedes v (HB/gpuFunctions_undir.cuh:104-113)."""
    return prop[u] > prop[v] or (prop[u] == prop[v] and u < v)


def pred_mex(adj, prop, c, v):
    return mex({c[x] for x in adj[v] if prior(prop, x, v)})


def direct(adj, prop, col, L0):
    c, L, tr = col[:], sorted(set(L0)), []
    while True:
        C = set()
        for v in L:
            for w in adj[v]:
                pv = prior(prop, v, w)
                if c[v] == c[w]:
                    C.add(w if pv else v)
                if pv:
                    C.add(w)
        new = {w: pred_mex(adj, prop, c, w) for w in C}
        changed = [w for w in C if new[w] != c[w]]
        tr.append((len(L), len(C), len(changed)))
        for w in changed:
            c[w] = new[w]
        if not changed:
            return c, tr
        L = sorted(C)


def keeps_mex(a, b, m):
    """A predecessor's change a -> b leaves a cached mex m in place."""
    return not (a < m or b == m)


def incremental(adj, prop, col, L0, keeps=keeps_mex):
    n = len(adj)
    c = col[:]
    pm = [pred_mex(adj, prop, c, v) for v in range(n)]
    cc = [sum(1 for x in adj[v] if prior(prop, x, v) and c[x] == c[v]) for v in range(n)]
    K = [0] * n
    L = set(L0)
    for x in L:
        for y in adj[x]:
            if prior(prop, x, y):
                K[y] += 1
    tr = []
    while True:
        C = {w for w in range(n) if K[w] > 0 or (w in L and cc[w] > 0)}
        new = {}
        for w in C:
            if pm[w] < 0:
                pm[w] = pred_mex(adj, prop, c, w)
            if pm[w] != c[w]:
                new[w] = pm[w]
        for w in new:
            cc[w] = 0   # its changed predecessors re-count it below
        tr.append((len(L), len(C), len(new)))
        D = L ^ C
        for x in D:
            for y in adj[x]:
                if prior(prop, x, y):
                    K[y] += 1 if x in C else -1
        for x, b in new.items():
            a = c[x]
            for y in adj[x]:
                if not prior(prop, x, y):
                    continue
                if y in new:
                    cc[y] += 1 if b == new[y] else 0
                else:
                    cc[y] += (1 if b == c[y] else 0) - (1 if a == c[y] else 0)
                m = pm[y]
                if m >= 0 and not keeps(a, b, m):
                    pm[y] = -1
        for x, b in new.items():
            c[x] = b
        # every count and every cached mex still held is exact for the next round
        for v in range(n):
            assert K[v] == sum(1 for x in adj[v] if x in C and prior(prop, x, v))
            assert cc[v] == sum(1 for x in adj[v] if prior(prop, x, v) and c[x] == c[v])
            if pm[v] >= 0:
                assert pm[v] == pred_mex(adj, prop, c, v), (v, pm[v])
        if not new:
            return c, tr
        L = C


def random_case(seed):
    rnd = random.Random(seed)
    n = rnd.randint(6, 40)
    m = rnd.randint(n, 4 * n)
    adj = [[] for _ in range(n)]
    for _ in range(m):
        u, v = rnd.randrange(n), rnd.randrange(n)
        if u == v:
            continue
        adj[u].append(v)   # repeats make multi-edges, as appended insertions do
        adj[v].append(u)
    # priorities: the running degree counter, with ties (HB/all_structure_undir.cuh:111-119)
    prop = [len(a) + rnd.choice((0, 0, 1, -1)) for a in adj]
    col = [rnd.randrange(0, 6) for _ in range(n)]
    L0 = rnd.sample(range(n), rnd.randint(1, max(1, n // 3)))
    return adj, prop, col, L0


@pytest.mark.parametrize("seed", range(60))
def test_incremental_ldf_loop_equals_direct(seed):
    adj, prop, col, L0 = random_case(seed)
    c1, t1 = direct(adj, prop, col, L0)
    c2, t2 = incremental(adj, prop, col, L0)
    assert t1 == t2
    assert c1 == c2


def test_drop_condition_covers_every_mex_move():
    # one predecessor's change a -> b moves a successor's mex only when
    # a < m or b == m; and the condition does keep many mexes in place
    moved = kept = 0
    for seed in range(100):
        adj, prop, col, L0 = random_case(seed)
        c = col[:]
        pm = [pred_mex(adj, prop, c, v) for v in range(len(adj))]
        for x in range(len(adj)):
            a = c[x]
            for b in range(8):
                if b == a:
                    continue
                c[x] = b
                for y in adj[x]:
                    if not prior(prop, x, y):
                        continue
                    if pm[y] != pred_mex(adj, prop, c, y):
                        assert not keeps_mex(a, b, pm[y])
                        moved += 1
                    elif keeps_mex(a, b, pm[y]):
                        kept += 1
                c[x] = a
    assert moved > 0 and kept > 0


@pytest.mark.parametrize("wrong", [lambda a, b, m: b != m,          # drops only on b == m
                                   lambda a, b, m: not a < m])     # drops only on a < m
def test_model_fails_with_a_weaker_condition(wrong):
    # a planted mistake in the drop condition leaves a stale mex that the
    # model's per-round check (or the trace) catches for some case
    caught = False
    for seed in range(60):
        adj, prop, col, L0 = random_case(seed)
        try:
            c2, t2 = incremental(adj, prop, col, L0, keeps=wrong)
        except AssertionError:
            caught = True
            break
        c1, t1 = direct(adj, prop, col, L0)
        if (c1, t1) != (c2, t2):
            caught = True
            break
    assert caught

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


def incremental(n, n_owned, adj, col, prev, L0, pull_frac, split=False):
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
        if sum(len(adj[x]) for x in D) > pull_frac * total and split:
            # heavy round, split pull: the slot caches still hold every
            # member's colour from before the round (its prev), and a slot
            # whose cached colour is above y's is skipped; then each member
            # whose colour went down adds its terms on such slots (new colour
            # equal to y's, x < y)
            for y in range(n_owned):
                K[y] = sum(1 for x in adj[y] if x in C and not p[x] > c[y]
                           and marks(x, c[x], p[x], y, c[y]))
            for x in C:
                if c[x] < p[x]:
                    for y in adj[x]:
                        if x < y < n_owned and c[y] == c[x]:
                            K[y] += 1
            R = {y for y in range(n_owned) if K[y] and y not in C}
        elif sum(len(adj[x]) for x in D) > pull_frac * total:
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


@pytest.mark.parametrize("pull_frac,split", [(2.0, False), (0.0, False), (0.05, False),
                                             (0.0, True), (0.05, True)])
def test_incremental_loop_equals_direct(pull_frac, split):
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
        got = incremental(n, n_owned, adj, col, prev, L0, pull_frac, split)
        assert got[2] == want[2], trial
        assert got[0] == want[0] and got[1] == want[1], trial


# ---------------------------------------------------------------- settle --
# The library stops the loop once the colours are settled: a round changed no
# colour and every owned vertex holds the mex of its neighbours' colours.
# Then each later round recolours every member to the colour it has (its mex
# over an unchanged snapshot), so the result is the current colouring.  The
# check below runs whole batches with the reference's rules (deletion round,
# insertion round on edges equal-coloured at batch start, loop;
# HI/Color_main.cu:146-202) on random multigraphs with frozen ghosts, and
# asserts that no round after a settled one changes a colour, and that the
# rule fires on a fair share of batches.

def _settle_batches(rng, n, m, nb, bsz, n_owned):
    edges = []
    for _ in range(m):
        a, b = rng.randrange(n), rng.randrange(n)
        if a != b and not (a >= n_owned and b >= n_owned):
            edges.append((a, b))

    def adjacency():
        adj = [[] for _ in range(n)]
        for a, b in edges:
            adj[a].append(b)
            adj[b].append(a)
        return adj

    adj = adjacency()
    c = [0] * n
    for v in range(n):   # greedy in ID order; ghosts keep fixed colours
        if v >= n_owned:
            c[v] = rng.randrange(4)
            continue
        used = {c[w] for w in adj[v] if w < v or w >= n_owned}
        k = 0
        while k in used:
            k += 1
        c[v] = k
    fired = violations = 0
    for _ in range(nb):
        dels, ins = [], []
        for _ in range(bsz):
            if rng.random() < 0.5 and edges:
                dels.append(edges[rng.randrange(len(edges))])
            else:
                a, b = rng.randrange(n), rng.randrange(n)
                if a != b and not (a >= n_owned and b >= n_owned):
                    ins.append((a, b))
        eq0 = [c[a] == c[b] for a, b in ins]
        gone = {(min(a, b), max(a, b)) for a, b in dels}
        edges = [e for e in edges if (min(e), max(e)) not in gone] + ins
        adj = adjacency()
        p = [-1] * n
        marked, change = [], False
        targets = []
        for a, b in dels:
            t = a if c[a] > c[b] else b
            if t < n_owned and t not in targets:
                targets.append(t)
        new = [mex(c, adj[t]) for t in targets]
        for t, m_ in zip(targets, new):
            if m_ != c[t]:
                p[t], c[t] = c[t], m_
                marked.append(t)
                change = True
        targets = []
        for (a, b), e0 in zip(ins, eq0):
            if not e0 or c[a] != c[b]:
                continue
            ka = len({c[w] for w in adj[a]}) if a < n_owned else 1 << 30
            kb = len({c[w] for w in adj[b]}) if b < n_owned else 1 << 30
            t = a if ka < kb else b
            if t < n_owned and t not in targets:
                targets.append(t)
        new = [mex(c, adj[t]) for t in targets]
        for t, m_ in zip(targets, new):
            p[t], c[t] = c[t], m_
            marked.append(t)
        change = change or bool(targets)
        if not change:
            continue
        L, rounds = sorted(set(marked)), []
        while True:
            C = sorted({w for v in L for w in adj[v]
                        if w < n_owned and marks(v, c[v], p[v], w, c[w])})
            new = {w: mex(c, adj[w]) for w in C}
            changed = sum(new[w] != c[w] for w in C)
            for w in C:
                p[w], c[w] = c[w], new[w]
            settled = changed == 0 and all(mex(c, adj[v]) == c[v] for v in range(n_owned))
            rounds.append((changed, settled))
            if not C:
                break
            L = C
        for r, (_, settled) in enumerate(rounds):
            if settled:
                fired += 1
                violations += sum(ch for ch, _ in rounds[r + 1:])
                break
    return fired, violations


def test_settle_rule():
    rng = random.Random(5)
    fired = 0
    for _ in range(12):
        for n, m, nb, bsz, own in [(60, 300, 5, 30, 60), (200, 1500, 4, 80, 200),
                                   (120, 900, 4, 60, 90)]:
            f, viol = _settle_batches(rng, n, m, nb, bsz, own)
            assert viol == 0
            fired += f
    assert fired >= 20


def test_settle_rule_arbitrary_states():
    # loops started from arbitrary colourings (unstable vertices anywhere):
    # the rule must look at every owned vertex, not only at the members
    rng = random.Random(7)
    fired = 0
    for trial in range(600):
        n = rng.randint(4, 30)
        n_owned = n if rng.random() < 0.7 else rng.randint(2, n)
        adj = [[] for _ in range(n)]
        pr = rng.choice([0.15, 0.3, 0.5])
        for a in range(n):
            for b in range(a + 1, n):
                if (a < n_owned or b < n_owned) and rng.random() < pr:
                    adj[a].append(b)
                    adj[b].append(a)
        c = [rng.randint(0, 6) for _ in range(n)]
        p = [rng.choice([-1] + list(range(7))) for _ in range(n)]
        L = sorted(v for v in range(n_owned) if rng.random() < 0.2)
        rounds = []
        while True:
            C = sorted({w for v in L for w in adj[v]
                        if w < n_owned and marks(v, c[v], p[v], w, c[w])})
            new = {w: mex(c, adj[w]) for w in C}
            changed = sum(new[w] != c[w] for w in C)
            for w in C:
                p[w], c[w] = c[w], new[w]
            settled = changed == 0 and all(mex(c, adj[v]) == c[v] for v in range(n_owned))
            rounds.append((changed, settled))
            if not C or len(rounds) > 200:
                break
            L = C
        for r, (_, settled) in enumerate(rounds):
            if settled:
                fired += 1
                assert sum(ch for ch, _ in rounds[r + 1:]) == 0, trial
                break
    assert fired >= 50

"""Trace comparison between the HIP library and the oracle.

The library's independent-rule loop may stop once the colours are settled
(trace kind 11, include/dyncolor.h): a round changed no colour and every owned
vertex holds the mex of its neighbours' colours, so every round the reference
still runs recolours each member to the colour it has.  Its trace is then the
oracle's trace up to that round, followed by the kind-11 entry; the oracle's
remaining rounds must all be loop rounds that change nothing.  That last
condition is what this module checks: a settle taken too early shows up as a
skipped oracle round with a colour change."""
LOOP, SETTLED = 10, 11


def same_trace(gpu, ref):
    gpu = [[int(x) for x in e] for e in gpu]
    ref = [[int(x) for x in e] for e in ref]
    if gpu and gpu[-1][0] == SETTLED:
        k = len(gpu) - 1
        if gpu[:k] != ref[:k]:
            return False
        runs = sum(1 for e in gpu[:k] if e[0] == LOOP)
        rest = ref[k:]
        return (gpu[-1][1] == runs and len(rest) > 0
                and all(e[0] == LOOP and e[3] == 0 for e in rest))
    return gpu == ref

"""Write tests/golden/sample.json from the reference's sample input files.

Inputs are data files shipped with the reference (graph, initial coloring and
change list of the 10-vertex sample).  The reference records no expected
outputs, so the expected colors/traces below were derived BY HAND from the
reference rules (round by round, on paper) and are independent of the oracle:

independent rule, coloring prevColor.txt of the heuristic-based directory:
  deletion round: targets {4, 1} (4: 2 -> 0, 1: 1 -> 0)
  insertion round: edge (0,5) equal colors, SC sizes 2 vs 5 -> target 0: 0 -> 2
  loop: L={0,1,4} C={2,3,4,7,9} (3:2->1, 4:0->4)
        L=5 C={0,3,4,6,7,8,9} (6:2->1, 7:3->2)
        L=7 C={0,4,7,8} (4:4->3, 8:4->3)
        L=4 C={4,8} no change; L=2 C={} stop
ldf rule (priority 5>0>7>1>4>8>9>6>3>2 from the line counters):
  change round: endpoints {0,1,2,3,4,5,6}, 5 change
  loop: L=5 C={2,3,4,6,9} 4 change; L=5 C={3,6,9} 2 change; L=3 C={} stop

Run once in a container that has the reference mounted; the JSON is committed.
"""
import json
import os

REF = "/root/reference"
HB = os.path.join(REF, "DynColorHeuristicBased", "DynColorHeuristicBased")
HI = os.path.join(REF, "DynColorHeuristicIndependent", "DynColorHeuristicIndependent")


def ints(path):
    rows = []
    with open(path) as f:
        for line in f:
            p = line.split()
            if p:
                rows.append([int(x) for x in p])
    return rows


def main():
    graph = ints(os.path.join(HB, "graph.txt"))
    changes = ints(os.path.join(HB, "cE.txt"))
    col_hb = ints(os.path.join(HB, "prevColor.txt"))
    col_hi = ints(os.path.join(HI, "prevColor.txt"))
    n = 10
    c_hb = [0] * n
    for v, c in col_hb:
        c_hb[v] = c
    c_hi = [0] * n
    for v, c in col_hi:
        c_hi[v] = c
    out = {
        "n": n,
        "edges": [[a, b] for a, b, _w in graph],
        "changes": [[a, b, ins] for a, b, _w, ins in changes],
        "colors_hb_dir": c_hb,
        "colors_hi_dir": c_hi,
        "expected": {
            "independent_hb_colors": {
                "colors": [2, 0, 1, 1, 3, 0, 1, 2, 3, 1],
                "trace": [[1, 2, 2, 0], [2, 1, 1, 0], [10, 3, 5, 2], [10, 5, 7, 2],
                          [10, 7, 4, 2], [10, 4, 2, 0], [10, 2, 0, 0]],
            },
            "ldf_hb_colors": {
                "colors": [1, 0, 2, 2, 1, 0, 2, 3, 4, 2],
                "trace": [[3, 7, 5, 0], [10, 5, 5, 4], [10, 5, 3, 2], [10, 3, 0, 0]],
                "prop": [6, 5, 2, 3, 5, 7, 4, 6, 5, 5],
            },
        },
    }
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "sample.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

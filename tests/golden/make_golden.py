"""Writes tests/golden/dp_golden.json: graphs with their exact treedepth and
canonical decomposition (parent array).

Expected values come from the CPU restatement in oracle/ (faithful
DynamicCPU for n <= 16, cross-checked against DynamicCPUImprov and the
flat-table DP; the flat-table DP alone above that) and, where one exists,
from the closed form, which the script asserts.  The reference has no Python
and no numeric golden files, so these vectors are the oracle's, pinned by the
closed forms and by the three-way agreement.

Run from the repo root: python tests/golden/make_golden.py
"""
import json
import math
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle_lib  # noqa: E402
from gisbdw_amd import graphs as G  # noqa: E402


def closed_form(name, g):
    if name.startswith("P"):
        return math.ceil(math.log2(g.n + 1))
    if name.startswith("C") and not name.startswith("Chordal"):
        return 1 + math.ceil(math.log2(g.n))
    if name.startswith("K") and "," not in name:
        return g.n
    if name.startswith("S"):
        return 2
    if name.startswith("K") and "," in name:
        a, b = map(int, name[1:].split(","))
        return min(a, b) + 1
    if name.startswith("E"):
        return 1
    return None


def main():
    o = oracle_lib.Oracle()
    cases = []
    fams = [G.path(n) for n in (1, 2, 3, 4, 7, 8, 10, 12, 14, 16, 21)]
    fams += [G.cycle(n) for n in range(3, 17)]
    fams += [G.complete(n) for n in (1, 2, 5, 9)]
    fams += [G.star(n) for n in (3, 6, 11)]
    fams += [G.complete_bipartite(a, b) for a, b in ((2, 3), (3, 5), (4, 4))]
    fams += [G.edgeless(5)]
    fams += [G.chordal_cycle(n) for n in range(10, 16)]
    fams += [G.sparse_connected(n, 7) for n in (8, 9, 10, 14, 15, 16, 17, 18)]
    fams += [G.halin(n, 3) for n in (10, 14)]
    fams += [G.workload_dp_single(0)]
    for g in fams:
        td, par = o.dp_fast(g)
        if g.n <= 16 and g.is_connected():
            td2, par2, _ = o.dyncpu(g)
            td3, par3, _ = o.dyncpu_improv(g)
            assert td == td2 == td3 and par == par2 == par3, g.name
        assert o.certify(g, par, td), g.name
        cf = closed_form(g.name, g)
        if cf is not None:
            assert cf == td, (g.name, cf, td)
        cases.append({"name": g.name, "n": g.n, "edges": g.edges(), "treedepth": td,
                      "parent": par, "closed_form": cf})
    out = Path(__file__).resolve().parent / "dp_golden.json"
    out.write_text(json.dumps({"generator": "tests/golden/make_golden.py", "cases": cases},
                              indent=0))
    print(f"wrote {len(cases)} cases to {out}")


if __name__ == "__main__":
    main()

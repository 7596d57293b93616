"""The C++ adapters (include/tdhip.hpp) through the example CLI, on the
GPU: the reference's P3 known answer (td 2, root 1) and a cycle."""
import subprocess
from pathlib import Path

import pytest

from gisbdw_amd import graphs as G

pytestmark = pytest.mark.gpu
CLI = Path(__file__).resolve().parent.parent / "gisbdw_amd" / "examples" / "solve_cli"


def run(alg, g, tmp_path):
    f = tmp_path / "g.gviz"
    f.write_text(G.format_graph_text(g))
    r = subprocess.run([str(CLI), "-a", alg, str(f)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    out = dict(line.split(": ") for line in r.stdout.splitlines() if ": " in line)
    edges = [tuple(map(int, line.rstrip(";").split(" -- "))) for line in r.stdout.splitlines()
             if " -- " in line]
    return out, edges


@pytest.mark.parametrize("alg", ["dynGPU", "bnbGPU"])
def test_cli_p3_and_cycle(alg, tmp_path, oracle):
    out, edges = run(alg, G.path(3), tmp_path)
    assert out["Treedepth"] == "2" and out["Root"] == "1" and out["Certified"] == "1"
    assert sorted(tuple(sorted(e)) for e in edges) == [(0, 1), (1, 2)]
    g = G.cycle(11)
    out, edges = run(alg, g, tmp_path)
    assert int(out["Treedepth"]) == oracle.dp_fast(g)[0]


@pytest.mark.parametrize("alg,kind", [("highestDegreeHeur", 2), ("spanningTreeHeur", 3),
                                      ("varianceHeur", 4)])
def test_cli_heuristics(alg, kind, tmp_path, oracle):
    """The reference CLI's heuristic options: the same depth and tree as the
    oracle's restatement of that heuristic, reported as uncertified."""
    for g in (G.cycle(11), G.grid(3, 4), G.gnp_connected(14, 0.3, 2)):
        out, edges = run(alg, g, tmp_path)
        td, parent, _ = oracle.heuristic(g, kind)
        assert int(out["Treedepth"]) == td and out["Certified"] == "0", g.name
        assert sorted(edges) == sorted((v, p) for v, p in enumerate(parent) if p >= 0), g.name


def test_cli_bottom_up(tmp_path, oracle):
    """bottomUpHeur (Treedepth/main.cpp:203-204): the oracle's bottom-up depth
    and forest, reported as uncertified."""
    for g in (G.cycle(11), G.grid(3, 4), G.gnp_connected(14, 0.3, 2)):
        out, edges = run("bottomUpHeur", g, tmp_path)
        td, order, parent, root, _, _ = oracle.bottom_up(g, history=False)
        assert int(out["Treedepth"]) == td and out["Certified"] == "0", g.name
        assert sorted(edges) == sorted((v, p) for v, p in enumerate(parent) if p >= 0), g.name

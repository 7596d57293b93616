"""The C++ adapters (include/tdhip.hpp) through the example CLI, on the
GPU, in the reference CLI's form (-a ALGORITHM -o OUTPUT_DIR files...): the
reference's P3 known answer (td 2, root 1), a cycle, the heuristics, and the
.out statistics / .gviz decomposition files it writes."""
import json
import subprocess
from pathlib import Path

import pytest

from gisbdw_amd import graphs as G

pytestmark = pytest.mark.gpu
CLI = Path(__file__).resolve().parent.parent / "gisbdw_amd" / "examples" / "solve_cli"


def run(alg, g, tmp_path):
    f = tmp_path / "g.txt"
    f.write_text(G.format_graph_text(g))
    out_dir = tmp_path / "out"
    out_dir.mkdir(exist_ok=True)
    for old in out_dir.iterdir():
        old.unlink()
    r = subprocess.run([str(CLI), "-a", alg, "-o", str(out_dir), str(f)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = dict(line.split(": ", 1) for line in r.stdout.splitlines() if ": " in line)
    assert r.stdout.startswith("Processing graph g\n")
    stats_files = sorted(out_dir.glob("g_*.out"))
    assert len(stats_files) == 1
    stats = json.loads(stats_files[0].read_text())
    assert stats["algorithmType"] == alg and stats["graphName"] == "g"
    assert stats["treedepth"] == out["Treedepth"] and stats["root"] == out["Root"]
    assert stats["vertices"] == str(g.n) and stats["edges"] == out["Edges"]
    gv = (out_dir / (stats_files[0].name + ".gviz")).read_text().splitlines()
    assert gv[0] == "graph G {" and gv[-1] == "}"
    edges = [tuple(map(int, line.rstrip(" ;").split("--"))) for line in gv if "--" in line]
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


def test_cli_several_files_and_cpu_names(tmp_path, oracle):
    """Several graph files in one call, one .out/.gviz pair each; the
    reference's CPU algorithm names give the same exact answer (computed on
    the GPU)."""
    gs = [G.cycle(9), G.grid(3, 4), G.gnp_connected(12, 0.3, 4)]
    files = []
    for i, g in enumerate(gs):
        f = tmp_path / f"graph{i}.txt"
        f.write_text(G.format_graph_text(g))
        files.append(str(f))
    out_dir = tmp_path / "many"
    out_dir.mkdir()
    r = subprocess.run([str(CLI), "-a", "dynCPU", "-o", str(out_dir), *files],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for i, g in enumerate(gs):
        stats = [json.loads(p.read_text()) for p in out_dir.glob(f"graph{i}_*.out")]
        assert len(stats) == 1
        assert int(stats[0]["treedepth"]) == oracle.dp_fast(g)[0]
    for alg in ("dynCPUImprov", "bnbCPU"):
        out, _ = run(alg, G.cycle(11), tmp_path)
        assert int(out["Treedepth"]) == oracle.dp_fast(G.cycle(11))[0]


def test_cli_usage_errors(tmp_path):
    r = subprocess.run([str(CLI), "-a", "dynGPU"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stdout
    f = tmp_path / "g.txt"
    f.write_text(G.format_graph_text(G.path(4)))
    r = subprocess.run([str(CLI), "-a", "noSuchAlg", "-o", str(tmp_path), str(f)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Wrong algorithm" in r.stderr

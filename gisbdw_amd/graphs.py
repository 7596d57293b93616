"""Graphs, seeded generators and the reference's graph text format.

A graph is ``n`` vertices ``0..n-1`` with adjacency stored as ``n`` uint64
bitmask rows (symmetric, no self loops), the layout every entry point of the
C-ABI takes.

Generators use SplitMix64 (Steele, Lea, Flood 2014), so every workload is
reproducible from its seed.  The reference's families follow
Treedepth/common/utils/graph_gen.hpp:16-83 (Path, Cycle, Complete,
ChordalCycle, SpanningTree, Halin, the reseed-until-connected G(n,p)
convention); the reference draws its random graphs with Boost's
sorted_erdos_renyi_iterator and std::minstd_rand, which are absent here, so the
random draws themselves are this module's own.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from typing import Iterable, List, Sequence, Tuple

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 generator (public-domain algorithm by S. Vigna)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def random(self) -> float:
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def randbelow(self, m: int) -> int:
        """Uniform integer in [0, m) by rejection (no modulo bias)."""
        if m <= 0:
            raise ValueError("randbelow needs m > 0")
        lim = (1 << 64) - ((1 << 64) % m)
        while True:
            x = self.next_u64()
            if x < lim:
                return x % m

    def shuffle(self, a: list) -> None:
        for i in range(len(a) - 1, 0, -1):
            j = self.randbelow(i + 1)
            a[i], a[j] = a[j], a[i]


@dataclass
class Graph:
    n: int
    adj: List[int] = field(default_factory=list)
    name: str = ""

    def __post_init__(self):
        if not self.adj:
            self.adj = [0] * self.n
        if len(self.adj) != self.n:
            raise ValueError("adjacency must have n rows")

    # ---- construction -------------------------------------------------
    @classmethod
    def from_edges(cls, n: int, edges: Iterable[Tuple[int, int]], name: str = "") -> "Graph":
        g = cls(n, [0] * n, name)
        for a, b in edges:
            g.add_edge(a, b)
        return g

    def add_edge(self, a: int, b: int) -> None:
        if a == b:
            raise ValueError("self loops are not allowed")
        if not (0 <= a < self.n and 0 <= b < self.n):
            raise ValueError("edge endpoint out of range")
        self.adj[a] |= 1 << b
        self.adj[b] |= 1 << a

    def has_edge(self, a: int, b: int) -> bool:
        return bool(self.adj[a] >> b & 1)

    def remove_edge(self, a: int, b: int) -> None:
        self.adj[a] &= ~(1 << b)
        self.adj[b] &= ~(1 << a)

    # ---- views -----------------------------------------------------------
    def edges(self) -> List[Tuple[int, int]]:
        out = []
        for a in range(self.n):
            m = self.adj[a] >> (a + 1)
            b = a + 1
            while m:
                if m & 1:
                    out.append((a, b))
                m >>= 1
                b += 1
        return out

    @property
    def m(self) -> int:
        return sum(bin(r).count("1") for r in self.adj) // 2

    def degree(self, v: int) -> int:
        return bin(self.adj[v]).count("1")

    def adj_array(self) -> np.ndarray:
        return np.array(self.adj, dtype=np.uint64)

    def components(self) -> List[int]:
        rest = (1 << self.n) - 1
        out = []
        while rest:
            low = rest & -rest
            comp = low
            frontier = low
            while frontier:
                v = (frontier & -frontier).bit_length() - 1
                frontier &= frontier - 1
                nb = self.adj[v] & rest & ~comp
                comp |= nb
                frontier |= nb
            out.append(comp)
            rest &= ~comp
        return out

    def is_connected(self) -> bool:
        return self.n > 0 and len(self.components()) == 1

    def relabel(self, perm: Sequence[int]) -> "Graph":
        """Vertex v becomes perm[v]."""
        g = Graph(self.n, [0] * self.n, self.name)
        for a, b in self.edges():
            g.add_edge(perm[a], perm[b])
        return g

    def validate(self) -> None:
        if not (1 <= self.n <= 64):
            raise ValueError("n must be in [1, 64]")
        full = (1 << self.n) - 1
        for v, row in enumerate(self.adj):
            if row & ~full:
                raise ValueError(f"row {v} names a vertex >= n")
            if row >> v & 1:
                raise ValueError(f"self loop at {v}")
            m = row
            while m:
                u = (m & -m).bit_length() - 1
                m &= m - 1
                if not self.adj[u] >> v & 1:
                    raise ValueError(f"asymmetric edge {v}-{u}")


# ---------------------------------------------------------------------------
# deterministic families (graph_gen.hpp:28-83)
# ---------------------------------------------------------------------------
def path(n: int) -> Graph:
    return Graph.from_edges(n, [(i, i + 1) for i in range(n - 1)], f"P{n}")


def cycle(n: int) -> Graph:
    g = path(n)
    g.add_edge(0, n - 1)
    g.name = f"C{n}"
    return g


def complete(n: int) -> Graph:
    return Graph.from_edges(n, [(i, j) for i in range(n) for j in range(i)], f"K{n}")


def chordal_cycle(n: int) -> Graph:
    g = cycle(n)
    g.add_edge(0, n // 3)
    g.add_edge(0, 2 * n // 3)
    g.name = f"ChordalC{n}"
    return g


def star(n: int) -> Graph:
    return Graph.from_edges(n, [(0, i) for i in range(1, n)], f"S{n}")


def complete_bipartite(a: int, b: int) -> Graph:
    return Graph.from_edges(a + b, [(i, a + j) for i in range(a) for j in range(b)], f"K{a},{b}")


def edgeless(n: int) -> Graph:
    return Graph(n, [0] * n, f"E{n}")


def grid(rows: int, cols: int) -> Graph:
    e = []
    for r in range(rows):
        for c in range(cols):
            v = r * cols + c
            if c + 1 < cols:
                e.append((v, v + 1))
            if r + 1 < rows:
                e.append((v, v + cols))
    return Graph.from_edges(rows * cols, e, f"grid{rows}x{cols}")


def spanning_tree(g: Graph) -> Graph:
    """SpanningTree (graph_gen.hpp:47-62): keep the edges, scanned in
    (i ascending, neighbour ascending) order, that join two sets."""
    uf = list(range(g.n))

    def find(x):
        while uf[x] != x:
            uf[x] = uf[uf[x]]
            x = uf[x]
        return x

    t = Graph(g.n, [0] * g.n, g.name + "_st")
    for i in range(g.n):
        m = g.adj[i]
        while m:
            j = (m & -m).bit_length() - 1
            m &= m - 1
            a, b = find(i), find(j)
            if a != b:
                uf[b] = a
                t.add_edge(i, j)
    return t


# ---------------------------------------------------------------------------
# random families
# ---------------------------------------------------------------------------
def gnp(n: int, p: float, seed: int) -> Graph:
    """G(n, p): every pair (i < j, i-major order) is an edge with prob. p."""
    rng = SplitMix64(seed)
    g = Graph(n, [0] * n, f"G({n},{p})#{seed}")
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < p:
                g.add_edge(i, j)
    return g


def gnp_connected(n: int, p: float, seed: int, max_tries: int = 100000) -> Graph:
    """RandomConnectedGraph convention (graph_gen.hpp:16-24): draw with
    seed+1, seed+2, ... until the graph is connected."""
    s = seed
    for _ in range(max_tries):
        s += 1
        g = gnp(n, p, s)
        if g.is_connected():
            g.name = f"Gc({n},{p})#{seed}"
            return g
    raise RuntimeError("no connected draw found")


def sparse_connected(n: int, seed: int) -> Graph:
    """RandomSparseConnectedGraph (graph_gen.hpp:25-27): p = ln n / n."""
    return gnp_connected(n, math.log(n) / n, seed)


def halin(n: int, seed: int = 0) -> Graph:
    """Halin (graph_gen.hpp:64-77): spanning tree of a sparse connected
    graph, leaves joined into a cycle in ascending order."""
    g = spanning_tree(sparse_connected(n, seed))
    leaves = [v for v in range(n) if g.degree(v) == 1]
    for a, b in zip(leaves, leaves[1:]):
        if not g.has_edge(a, b):
            g.add_edge(a, b)
    if len(leaves) > 2 and not g.has_edge(leaves[0], leaves[-1]):
        g.add_edge(leaves[0], leaves[-1])
    g.name = f"Halin{n}#{seed}"
    return g


def random_permutation(n: int, rng: SplitMix64) -> List[int]:
    p = list(range(n))
    rng.shuffle(p)
    return p


def grid_dropped(rows: int, cols: int, frac: float, seed: int) -> Graph:
    """rows x cols grid minus ceil(frac * |E|) edges drawn uniformly without
    replacement.  May be disconnected."""
    g = grid(rows, cols)
    rng = SplitMix64(seed)
    e = g.edges()
    drop = math.ceil(frac * len(e))
    rng.shuffle(e)
    for a, b in e[:drop]:
        g.remove_edge(a, b)
    g.name = f"grid{rows}x{cols}-{drop}#{seed}"
    return g


def random_cubic(n: int, seed: int, max_tries: int = 100000) -> Graph:
    """Random 3-regular simple graph: configuration model with rejection."""
    if n % 2 or n < 4:
        raise ValueError("3-regular graphs need even n >= 4")
    rng = SplitMix64(seed)
    for _ in range(max_tries):
        pts = [v for v in range(n) for _ in range(3)]
        rng.shuffle(pts)
        g = Graph(n, [0] * n, f"cubic{n}#{seed}")
        ok = True
        for i in range(0, len(pts), 2):
            a, b = pts[i], pts[i + 1]
            if a == b or g.has_edge(a, b):
                ok = False
                break
            g.add_edge(a, b)
        if ok and g.is_connected():
            return g
    raise RuntimeError("no simple connected cubic draw found")


def random_tree(n: int, rng: SplitMix64) -> Graph:
    """Uniform random labelled tree from a random Pruefer sequence."""
    if n == 1:
        return Graph(1, [0])
    if n == 2:
        return Graph.from_edges(2, [(0, 1)])
    seq = [rng.randbelow(n) for _ in range(n - 2)]
    deg = [1] * n
    for x in seq:
        deg[x] += 1
    g = Graph(n, [0] * n)
    import heapq

    leaves = [v for v in range(n) if deg[v] == 1]
    heapq.heapify(leaves)
    for x in seq:
        leaf = heapq.heappop(leaves)
        g.add_edge(leaf, x)
        deg[x] -= 1
        if deg[x] == 1:
            heapq.heappush(leaves, x)
    a = heapq.heappop(leaves)
    b = heapq.heappop(leaves)
    g.add_edge(a, b)
    return g


def tree_plus_edges(n: int, extra_frac: float, seed: int) -> Graph:
    """Random spanning tree plus ceil(extra_frac*(n-1)) extra random edges."""
    rng = SplitMix64(seed)
    g = random_tree(n, rng)
    extra = math.ceil(extra_frac * (n - 1))
    added = 0
    while added < extra:
        a, b = rng.randbelow(n), rng.randbelow(n)
        if a != b and not g.has_edge(a, b):
            g.add_edge(a, b)
            added += 1
    g.name = f"tree{n}+{extra}#{seed}"
    return g


def relabeled(g: Graph, seed: int) -> Graph:
    rng = SplitMix64(seed)
    h = g.relabel(random_permutation(g.n, rng))
    h.name = g.name + f"~{seed}"
    return h


# ---------------------------------------------------------------------------
# the benchmark workloads
# ---------------------------------------------------------------------------
def workload_dp_single(seed: int = 0) -> Graph:
    """One G(20, 0.3) graph, reseeded until connected."""
    return gnp_connected(20, 0.3, seed)


def workload_dp_n30() -> List[Graph]:
    """8 graphs at n = 30: 4 x G(30, 0.2) seeds 0-3, 4 x 5x6 grid with 10% of
    its 49 edges (5) dropped, seeds 0-3 (possibly disconnected)."""
    gs = [gnp(30, 0.2, s) for s in range(4)]
    gs += [grid_dropped(5, 6, 0.10, s) for s in range(4)]
    return gs


def workload_dp_batch(count: int = 4096, seed: int = 0, nmin: int = 12, nmax: int = 24,
                      pmin: float = 0.15, pmax: float = 0.5) -> List[Graph]:
    """`count` G(n,p) graphs, n ~ U{nmin..nmax}, p ~ U[pmin, pmax]; graph i
    drawn from a seed derived from (seed, i)."""
    rng = SplitMix64(seed)
    out = []
    for i in range(count):
        n = nmin + rng.randbelow(nmax - nmin + 1)
        p = pmin + (pmax - pmin) * rng.random()
        out.append(gnp(n, p, rng.next_u64()))
    return out


def workload_bnb(count: int = 256, seed: int = 0) -> List[Graph]:
    """Sparse graphs with n in [40, 64]: a quarter random 3-regular (even
    n), a quarter 7x7 / 8x8 grids with a random relabel, half random trees
    plus ceil(0.1 (n-1)) extra edges."""
    rng = SplitMix64(seed)
    out = []
    q = count // 4
    for i in range(q):
        n = 40 + 2 * rng.randbelow(13)
        out.append(random_cubic(n, rng.next_u64()))
    for i in range(q):
        base = grid(7, 7) if i % 2 == 0 else grid(8, 8)
        out.append(relabeled(base, rng.next_u64()))
    for i in range(count - 2 * q):
        n = 40 + rng.randbelow(25)
        out.append(tree_plus_edges(n, 0.10, rng.next_u64()))
    return out


# ---------------------------------------------------------------------------
# the reference's graph text format (Treedepth/main.cpp:131-143)
# ---------------------------------------------------------------------------
_VERTEX_RE = re.compile(r"^\s*(\d+)\s*;\s*$", re.M)
_EDGE_RE = re.compile(r"(\d+)\s*--\s*(\d+)")


def parse_graph_text(text: str, name: str = "") -> Graph:
    """Vertex lines ``N;`` and edge lines ``a -- b``.  The reference counts
    vertex lines with an anchored regex and lets add_edge grow the vertex set
    to max id + 1; here n = max(vertex-line count, declared ids + 1,
    edge endpoints + 1), so trailing isolated vertices are kept."""
    ids = [int(m.group(1)) for m in _VERTEX_RE.finditer(text)]
    edges = [(int(a), int(b)) for a, b in _EDGE_RE.findall(text)]
    n = max([len(ids)] + [i + 1 for i in ids] + [max(a, b) + 1 for a, b in edges] + [0])
    g = Graph(n, [0] * n, name)
    for a, b in edges:
        if a != b and not g.has_edge(a, b):
            g.add_edge(a, b)
    return g


def format_graph_text(g: Graph) -> str:
    lines = ["graph G {"]
    lines += [f"{v};" for v in range(g.n)]
    lines += [f"{a} -- {b};" for a, b in g.edges()]
    lines.append("}")
    return "\n".join(lines) + "\n"

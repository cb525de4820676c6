"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each check compares the oracle with something it was not derived from:
the reference's own plan classes and binary (golden fixtures, and the live
binary when oracle/_ref is built), closed forms on cliques, and a brute-force
monomorphism count on tiny random graphs.
"""
import itertools
import math
import os
import random

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, clique

# Appendix of closed forms: counts of each pattern in K_n with symmetry
# breaking (each unordered instance once).
def kn_closed_form(q, n):
    C = math.comb
    return {0: C(n, 3), 1: 3 * C(n, 4), 2: 6 * C(n, 4), 3: C(n, 4), 4: 60 * C(n, 5), 7: C(n, 5),
            11: C(n, 6)}[q]


def test_plans_match_reference_plan_classes():
    with open(os.path.join(GOLDEN, "light_plans.txt")) as f:
        golden = f.read()
    mine = "".join(oracle.plan_text(q, cl, o) for o in (True, False) for q in range(14) for cl in range(4))
    assert mine == golden


def test_appendix_plans():
    # LIGHT execute sequences of the six benchmark queries (default compress level)
    ops = {q: oracle.plan_text(q, 3).split("\n")[1] for q in (0, 1, 2, 3, 4, 7)}
    assert ops[0] == "ops: M0 C1 M1 CC2"
    assert ops[1] == "ops: M0 C1 M1 C2 M2 CC3"
    assert ops[2] == "ops: M0 C2 M2 CPC1 CPC3 COUNT4"
    assert ops[3] == "ops: M0 C1 M1 C2 M2 CC3"
    assert ops[4] == "ops: M1 C2 C3 M2 C0 M3 FC0 CC4"
    assert ops[7] == "ops: M0 C1 M1 C2 M2 C3 M3 CC4"


@pytest.mark.parametrize("n", [5, 6, 8])
@pytest.mark.parametrize("q", [0, 1, 2, 3, 4, 7, 11])
def test_cliques_closed_form(q, n):
    if q == 11 and n < 6:
        pytest.skip("needs 6 vertices")
    g = clique(n)
    want = kn_closed_form(q, n)
    assert oracle.wcoj_count(g.row_ptrs, g.cols, q) == want
    for cl in range(4):
        assert oracle.light_count(g.row_ptrs, g.cols, q, cl)[0] == want
    assert oracle.gpsm_count(g.row_ptrs, g.cols, q)[0] == want


# |Aut(P)| of each pattern (number of automorphisms)
AUT = {0: 6, 1: 8, 2: 4, 3: 24, 4: 2, 7: 120}
PATTERN_EDGES = {
    0: [(0, 1), (0, 2), (1, 2)],
    1: [(0, 1), (1, 2), (2, 3), (0, 3)],
    2: [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3)],
    3: [(a, b) for a, b in itertools.combinations(range(4), 2)],
    4: [(0, 1), (0, 2), (1, 2), (1, 3), (2, 4), (3, 4)],
    7: [(a, b) for a, b in itertools.combinations(range(5), 2)],
}


def brute_monomorphisms(adj, q):
    """Injective maps of the pattern into the graph preserving pattern edges."""
    k = 1 + max(max(e) for e in PATTERN_EDGES[q])
    n = len(adj)
    sets = [set(a) for a in adj]
    cnt = 0
    for perm in itertools.permutations(range(n), k):
        if all(perm[b] in sets[perm[a]] for a, b in PATTERN_EDGES[q]):
            cnt += 1
    return cnt


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_brute_force_small_random(seed):
    import subgraphmatchgpu_amd as sme

    rng = random.Random(seed)
    n = 9
    adj = [set() for _ in range(n)]
    for a, b in itertools.combinations(range(n), 2):
        if rng.random() < 0.55:
            adj[a].add(b)
            adj[b].add(a)
    g = sme.from_adjacency([sorted(s) for s in adj])
    for q, aut in AUT.items():
        want = brute_monomorphisms([sorted(s) for s in adj], q)
        assert want % aut == 0
        want //= aut
        assert oracle.wcoj_count(g.row_ptrs, g.cols, q) == want, q
        assert oracle.light_count(g.row_ptrs, g.cols, q, 3)[0] == want, q


def _fixture_cases(ref_counts, max_name=None):
    cases = []
    for name, res in ref_counts["counts"].items():
        for q, c in res["wcoj"].items():
            cases.append((name, int(q), c))
    return cases


SLOW = {("rmat10", 5), ("rmat10", 13), ("rmat10", 4), ("rmat14", 7), ("rmat14", 3), ("rmat14", 2),
        ("rmat12", 1), ("rmat12", 6), ("rmat12", 11), ("rmat12", 7)}


def test_oracle_vs_reference_binary_fixture(ref_counts, fixture_graphs):
    """WCOJ and LIGHT restatements against the reference binary's counts."""
    for name, q, want in _fixture_cases(ref_counts):
        if (name, q) in SLOW:
            continue
        g = fixture_graphs[name]
        assert oracle.wcoj_count(g.row_ptrs, g.cols, q) == want, (name, q)
        assert oracle.light_count(g.row_ptrs, g.cols, q, 3)[0] == want, (name, q)


def test_reference_algorithms_agree(ref_counts):
    for name, res in ref_counts["counts"].items():
        for algo in ("bfs", "rbfs", "vf2"):
            for q, c in res.get(algo, {}).items():
                assert c == res["wcoj"][q], (name, algo, q)


def test_triangle_levels_closed_form(fixture_graphs):
    """Q0 LIGHT levels: M0 = n, C1 = M1 = nnz/2 (each undirected edge once,
    v1 > v0), CC2 = triangles."""
    for name in ("rmat8", "rmat12", "rmat14", "er300"):
        g = fixture_graphs[name]
        cnt, lv = oracle.light_count(g.row_ptrs, g.cols, 0, 3)
        assert lv == [g.n, g.nnz // 2, g.nnz // 2, cnt]


def kn_level_forms(q, n):
    """Per-exec-level outputs of the default (SPECIAL) LIGHT plan on K_n,
    derived by hand from each op's semantics:
      Q1 C2 = sum over v0 < v1 of |{v2 > v0, v2 != v1}| = 2 C(n,3);
      Q2 CPC1 = CPC3 = sum over v0 < v2 of |N(v0) & N(v2)| = C(n,2)(n-2);
      Q4 C3 = all of N(v1) (n(n-1)), C0 = M3 = C(n,2)(n-2) = 3 C(n,3),
         FC0 drops v3 from cand0: 3 C(n,3)(n-3) = 12 C(n,4)."""
    C = math.comb
    c2, c3, c4, c5 = C(n, 2), C(n, 3), C(n, 4), C(n, 5)
    return {0: [n, c2, c2, c3],
            1: [n, c2, c2, 2 * c3, 2 * c3, 3 * c4],
            2: [n, c2, c2, 3 * c3, 3 * c3, 6 * c4],
            3: [n, c2, c2, c3, c3, c4],
            4: [n, c2, n * (n - 1), c2, 3 * c3, 3 * c3, 12 * c4, 60 * c5],
            7: [n, c2, c2, c3, c3, c4, c4, c5]}[q]


@pytest.mark.parametrize("n", [5, 6, 8, 11])
@pytest.mark.parametrize("q", [0, 1, 2, 3, 4, 7])
def test_light_levels_closed_form_cliques(q, n):
    g = clique(n)
    count, levels = oracle.light_count(g.row_ptrs, g.cols, q, 3)
    assert levels == kn_level_forms(q, n)
    assert count == kn_closed_form(q, n)


def test_census_matches_reference_binary(ref_counts, fixture_graphs):
    """The independent census (degree-ordered clique listing, per-edge
    triangle support) reproduces the reference binary's triangle, diamond,
    4-clique and 5-clique counts on every fixture graph."""
    for name, res in ref_counts["counts"].items():
        g = fixture_graphs[name]
        c = oracle.census(g.row_ptrs, g.cols, kmax=5)
        assert c["edges"] == g.nnz // 2
        for q, key in (("0", "triangles"), ("2", "diamonds"), ("3", "k4"), ("7", "k5")):
            if q in res["wcoj"]:
                assert c[key] == res["wcoj"][q], (name, q)


def test_census_cliques_and_brute_force():
    import subgraphmatchgpu_amd as sme

    for n in (4, 7, 10):
        c = oracle.census(clique(n).row_ptrs, clique(n).cols, kmax=5)
        assert (c["triangles"], c["diamonds"], c["k4"], c["k5"]) == (
            kn_closed_form(0, n), kn_closed_form(2, n), kn_closed_form(3, n), kn_closed_form(7, n))
    rng = random.Random(7)
    n = 9
    adj = [set() for _ in range(n)]
    for a, b in itertools.combinations(range(n), 2):
        if rng.random() < 0.6:
            adj[a].add(b)
            adj[b].add(a)
    lists = [sorted(s) for s in adj]
    g = sme.from_adjacency(lists)
    c = oracle.census(g.row_ptrs, g.cols, kmax=5)
    for q, key in ((0, "triangles"), (2, "diamonds"), (3, "k4"), (7, "k5")):
        assert c[key] == brute_monomorphisms(lists, q) // AUT[q], q


def test_light_levels_from_census(fixture_graphs):
    """On R-MAT graphs the clique and diamond plans' level outputs are whole
    graph counts: Q3 [n, m, m, T, T, K4], Q7 [.., K4, K4, K5], Q2 [n, m, m,
    3T, 3T, D] (each triangle has three chord candidates)."""
    for name in ("rmat8", "rmat12"):
        g = fixture_graphs[name]
        c = oracle.census(g.row_ptrs, g.cols, kmax=5)
        m, T = c["edges"], c["triangles"]
        assert oracle.light_count(g.row_ptrs, g.cols, 3, 3) == (c["k4"], [g.n, m, m, T, T, c["k4"]])
        assert oracle.light_count(g.row_ptrs, g.cols, 7, 3) == (
            c["k5"], [g.n, m, m, T, T, c["k4"], c["k4"], c["k5"]])
        assert oracle.light_count(g.row_ptrs, g.cols, 2, 3) == (c["diamonds"], [g.n, m, m, 3 * T, 3 * T,
                                                                               c["diamonds"]])


def test_level_outputs_consistent(fixture_graphs):
    """Every compress level gives the same total; GPSM's last level is the
    count and its first level is the start range."""
    g = fixture_graphs["rmat8"]
    for q in range(14):
        totals = {oracle.light_count(g.row_ptrs, g.cols, q, cl)[0] for cl in range(4)}
        assert len(totals) == 1, q
        c, lv = oracle.gpsm_count(g.row_ptrs, g.cols, q)
        assert c in totals and lv[0] == g.n


def test_start_range_additivity(fixture_graphs):
    g = fixture_graphs["rmat10"]
    for q in (0, 2, 3, 4, 7):
        whole = oracle.light_count(g.row_ptrs, g.cols, q, 3)[0]
        parts = sum(oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=a, v_end=b)[0]
                    for a, b in ((0, 100), (100, 517), (517, g.n)))
        assert parts == whole


def test_per_vertex_counts_add_up(fixture_graphs):
    """Per-start-vertex counts sum to the range counts and equal one-vertex
    ranges."""
    g = fixture_graphs["rmat10"]
    for q in (0, 1, 2, 3, 4, 7):
        pv = oracle.light_count_per_vertex(g.row_ptrs, g.cols, q, 100, 900)
        assert int(pv.sum()) == oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=100, v_end=900)[0]
        for v in (100, 101, 350, 899):
            assert int(pv[v - 100]) == oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=v, v_end=v + 1)[0]


@pytest.mark.skipif(not oracle.ref_available(), reason="reference CPU binary not built")
def test_oracle_vs_live_reference(tmp_path):
    import subgraphmatchgpu_amd as sme

    g = sme.gen_rmat(9, 6, seed=99)
    p = str(tmp_path / "g.bin")
    sme.write_bin(p, g)
    for q in (0, 1, 2, 3, 4, 7, 8, 9):
        want, _ = oracle.ref_count(p, q, algo=0, threads=4)
        assert oracle.wcoj_count(g.row_ptrs, g.cols, q) == want, q
        assert oracle.light_count(g.row_ptrs, g.cols, q, 3)[0] == want, q


def test_planted_mistake_is_caught():
    """A plausible oracle mistake (dropping the symmetry-breaking orders)
    must change the count the pins compare against."""
    g = clique(6)
    assert oracle.wcoj_count(g.row_ptrs, g.cols, 0, enable_ordering=False) == 6 * kn_closed_form(0, 6)
    assert oracle.wcoj_count(g.row_ptrs, g.cols, 0, enable_ordering=False) != kn_closed_form(0, 6)


# Pattern edges of the listed queries (src/query/common/Query.h:84-282).
PATTERN_EDGES = {
    0: [(0, 1), (0, 2), (1, 2)],
    1: [(0, 1), (1, 2), (2, 3), (0, 3)],
    2: [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3)],
    3: [(a, b) for a in range(4) for b in range(a + 1, 4)],
    4: [(0, 1), (0, 2), (1, 2), (1, 3), (2, 4), (3, 4)],
    7: [(a, b) for a in range(5) for b in range(a + 1, 5)],
}


def check_match_list(g, q, rows, total):
    """rows are valid embeddings, one per instance (distinct edge images),
    lexicographically sorted, and as many as the pinned total."""
    rp, cols = g.row_ptrs, g.cols
    adj = [set(cols[rp[u]:rp[u + 1]].tolist()) for u in range(g.n)]
    images = set()
    for r in rows.tolist():
        assert len(set(r)) == len(r)
        for a, b in PATTERN_EDGES[q]:
            assert r[b] in adj[r[a]], (q, r)
        images.add(frozenset(frozenset((r[a], r[b])) for a, b in PATTERN_EDGES[q]))
    assert len(images) == len(rows) == total
    assert rows.tolist() == sorted(rows.tolist())


@pytest.mark.parametrize("q", sorted(PATTERN_EDGES))
def test_gpsm_list_pinned(fixture_graphs, q):
    g = fixture_graphs["er300" if q == 4 else "rmat8"]
    vb, ve = 0, g.n
    rows, total = oracle.gpsm_list(g.row_ptrs, g.cols, q, vb, ve)
    assert total == oracle.gpsm_count(g.row_ptrs, g.cols, q, v_begin=vb, v_end=ve)[0]
    assert len(rows) == total > 0
    check_match_list(g, q, rows, total)
    part, t2 = oracle.gpsm_list(g.row_ptrs, g.cols, q, vb, ve, cap=5)
    assert t2 == total and part.tolist() == rows[:5].tolist()

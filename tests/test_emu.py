"""The engine's own sources (kernels + host orchestration) executed on the CPU
by the fiber-based stand-in runtime in tests/emu, under AddressSanitizer and
UBSan, compared with the oracle.  This catches index arithmetic errors before
any kernel reaches the GPU."""
import os
import subprocess

import pytest

import oracle
import subgraphmatchgpu_amd as sme
from conftest import ROOT, clique

EMU_DIR = os.path.join(ROOT, "tests", "emu")
EMU = os.path.join(EMU_DIR, "emu_driver")


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-C", EMU_DIR, "-s", "all"], check=True)
    return EMU


def run_emu(emu, path, algo, q, cl, vb, ve, budget, ordering=1, hub_min_degree=None, env=None):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", **(env or {}))
    cmd = [emu, path, str(algo), str(q), str(cl), str(ordering), str(vb), str(ve), str(budget)]
    if hub_min_degree is not None:
        cmd.append(str(hub_min_degree))
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    parts = r.stdout.split()
    cnt = int(parts[0].split("=")[1])
    lv = [int(x) for x in parts[1].split("=")[1].split(",")]
    return cnt, lv, int(parts[2].split("=")[1])


def hubs_graph(n=400, hubs=4, spokes=150, extra=1500, seed=5):
    """A sparse G(n, m) plus a few vertices of low id joined to 150 others
    each: upper neighbourhoods above 64 (bit rows in device memory) and
    heavy 4-cycle apexes, with few cliques."""
    import numpy as np

    rng = np.random.default_rng(seed)
    adj = [set() for _ in range(n)]
    for h in range(hubs):
        for v in rng.choice(np.arange(hubs, n), size=spokes, replace=False):
            adj[h].add(int(v))
            adj[int(v)].add(h)
    for a, b in rng.integers(0, n, size=(extra, 2)):
        if a != b:
            adj[a].add(int(b))
            adj[b].add(int(a))
    return sme.from_adjacency([sorted(x) for x in adj])


@pytest.fixture(scope="module")
def graphs(tmp_path_factory):
    d = tmp_path_factory.mktemp("emu")
    out = {}
    for name, g in {"rmat6": sme.gen_rmat(6, 8, seed=3), "rmat7": sme.gen_rmat(7, 6, seed=3),
                    "er60": sme.gen_er(60, 240, seed=2), "k7": clique(7),
                    "hubs": hubs_graph()}.items():
        p = str(d / (name + ".bin"))
        sme.write_bin(p, g)
        out[name] = (g, p)
    return out


@pytest.mark.parametrize("q", list(range(14)))
def test_emu_light_all_compress_levels(emu, graphs, q):
    for name in ("rmat6", "er60"):
        g, p = graphs[name]
        for cl in range(4):
            want = oracle.light_count(g.row_ptrs, g.cols, q, cl)
            got = run_emu(emu, p, 6, q, cl, 0, g.n, 1 << 30)
            assert got[:2] == (want[0], want[1]), (name, cl)


@pytest.mark.parametrize("q", [0, 1, 2, 3, 4, 7, 9, 13])
def test_emu_gpsm(emu, graphs, q):
    g, p = graphs["rmat7"]
    want, wl = oracle.gpsm_count(g.row_ptrs, g.cols, q)
    for algo in (4, 5):  # GPU_GPSM, GPU_GPSM_REUSE
        cnt, lv, _ = run_emu(emu, p, algo, q, 0, 0, g.n, 1 << 30)
        assert cnt == want
        # the engine's levels alternate C/M; the M levels are GPSM's level sizes
        assert [lv[0]] + lv[2::2] == wl[:len(lv[2::2]) + 1]


def test_emu_tiny_budget_invariance(emu, graphs):
    g, p = graphs["rmat7"]
    for q in (1, 3, 4, 7):
        want = oracle.light_count(g.row_ptrs, g.cols, q, 3)
        seen = set()
        for budget in (40_000, 1 << 30):
            # batching is the level engine's (cliques too, through the
            # level engine here)
            cnt, lv, batches = run_emu(emu, p, 6, q, 3, 0, g.n, budget, env={"TUNE": "0,0,0,4,0,0,0,1"})
            assert (cnt, lv) == (want[0], want[1]), (q, budget)
            seen.add(batches)
            # and the clique bitmap path under the same budget
            assert run_emu(emu, p, 6, q, 3, 0, g.n, budget)[:2] == (want[0], want[1]), (q, budget)
        assert len(seen) > 1, "budgets did not change the batching"


def test_emu_ranges_and_empty(emu, graphs):
    g, p = graphs["rmat7"]
    for q in (0, 4):
        for a, b in ((0, 0), (5, 5), (0, 17), (17, 90), (90, g.n)):
            want = oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=a, v_end=b)
            assert run_emu(emu, p, 6, q, 3, a, b, 1 << 30)[:2] == (want[0], want[1])


def test_emu_hub_bitmaps(emu, graphs):
    """Membership through hub bitmaps (low threshold: most lists are hubs)."""
    g, p = graphs["rmat7"]
    for q in (0, 1, 2, 3, 4, 7, 13):
        want = oracle.light_count(g.row_ptrs, g.cols, q, 3)
        for thr in (1, 8):
            assert run_emu(emu, p, 6, q, 3, 0, g.n, 1 << 30, hub_min_degree=thr)[:2] == \
                (want[0], want[1]), (q, thr)


def test_emu_buckets(emu, graphs):
    """Per-start-vertex-bucket counts of one pass equal separate range counts."""
    g, p = graphs["rmat7"]
    bounds = [0, 3, 3, 17, 40, 41, 90, g.n]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", BUCKETS=",".join(map(str, bounds)))
    for q in (0, 1, 2, 4, 7):
        for algo in (6, 4):
            r = subprocess.run([emu, p, str(algo), str(q), "3", "1", "0", str(g.n), str(1 << 30)],
                               capture_output=True, text=True, env=env, timeout=600)
            assert r.returncode == 0, r.stderr[-2000:]
            got = [int(x) for x in r.stdout.split("buckets=")[1].split(",")]
            count = oracle.light_count if algo == 6 else oracle.gpsm_count
            want = [count(g.row_ptrs, g.cols, q, v_begin=a, v_end=b)[0]
                    for a, b in zip(bounds[:-1], bounds[1:])]
            assert got == want, (q, algo)


def test_emu_buckets_without_table(emu, graphs, tmp_path):
    """Buckets over a range too large for the vertex->bucket table under a
    40 KB budget (the rmat7 edges plus isolated vertices up to 6,000): the
    binary search over the bounds gives the same counts."""
    import numpy as np

    g0, _ = graphs["rmat7"]
    rp = np.concatenate([g0.row_ptrs, np.full(6000 - g0.n, g0.row_ptrs[-1], dtype=np.uint64)])
    g = sme.Graph(rp, g0.cols.copy())
    p = str(tmp_path / "padded.bin")
    sme.write_bin(p, g)
    bounds = [0, 3, 17, 40, 41, 90, 128, 5000, 6000]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", BUCKETS=",".join(map(str, bounds)))
    for q in (0, 1):
        for algo in (6, 4):
            r = subprocess.run([emu, p, str(algo), str(q), "3", "1", "0", str(g.n), "40000"],
                               capture_output=True, text=True, env=env, timeout=600)
            assert r.returncode == 0, r.stderr[-2000:]
            got = [int(x) for x in r.stdout.split("buckets=")[1].split(",")]
            count = oracle.light_count if algo == 6 else oracle.gpsm_count
            assert got == [count(g.row_ptrs, g.cols, q, v_begin=a, v_end=b)[0]
                           for a, b in zip(bounds[:-1], bounds[1:])], (q, algo)


def test_emu_clique_without_ordering(emu, graphs):
    g, p = graphs["k7"]
    for q in (0, 3, 7):
        want = oracle.light_count(g.row_ptrs, g.cols, q, 3, enable_ordering=False)
        assert run_emu(emu, p, 6, q, 3, 0, g.n, 1 << 30, ordering=0)[:2] == (want[0], want[1])


def test_emu_er_batch_members(emu, tmp_path):
    """Per-member counts of a packed batch of ER graphs in one pass equal the
    WCOJ oracle on each member alone."""
    from conftest import member

    G, offs = sme.gen_er_batch(12, (8, 40), (3.0, 10.0), seed=3)
    p = str(tmp_path / "batch.bin")
    sme.write_bin(p, G)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", BUCKETS=",".join(map(str, offs)))
    for q in (0, 1, 2, 3):
        want = [oracle.wcoj_count(m.row_ptrs, m.cols, q)
                for m in (member(G, offs, k) for k in range(12))]
        r = subprocess.run([emu, p, "6", str(q), "3", "1", "0", str(G.n), str(1 << 30)],
                           capture_output=True, text=True, env=env, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        got = [int(x) for x in r.stdout.split("buckets=")[1].split(",")]
        assert got == want, q
        assert any(want)


@pytest.mark.parametrize("tune", ["4,0,0", "64,3,0", "1000000,1,20000", "0,8,0,2", "16,16,0,4",
                                  "0,0,0,4,8", "8,0,0,2,1000", "0,0,0,4,0,0,4", "8,2,0,4,0,8,16"])
def test_emu_tunings_do_not_change_results(emu, graphs, tune):
    """Lane-per-path counting, tiny grids and tiny batch caps: same counts,
    levels and per-bucket counts."""
    g, p = graphs["rmat7"]
    bounds = [0, 3, 17, 40, 90, g.n]
    for q in (0, 1, 2, 3, 4, 7, 9, 13):
        want, wl = oracle.light_count(g.row_ptrs, g.cols, q, 3)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", TUNE=tune)
        r = subprocess.run([emu, p, "6", str(q), "3", "1", "0", str(g.n), str(1 << 30)],
                           capture_output=True, text=True, env=env, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        parts = r.stdout.split()
        assert int(parts[0].split("=")[1]) == want
        assert [int(x) for x in parts[1].split("=")[1].split(",")] == wl
        env["BUCKETS"] = ",".join(map(str, bounds))
        r = subprocess.run([emu, p, "6", str(q), "3", "1", "0", str(g.n), str(1 << 30)],
                           capture_output=True, text=True, env=env, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        got = [int(x) for x in r.stdout.split("buckets=")[1].split(",")]
        assert got == [oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=a, v_end=b)[0]
                       for a, b in zip(bounds[:-1], bounds[1:])], q


def test_emu_super_chunk_fast_path(emu, tmp_path):
    """Runs of chunks inside one path (long candidate lists) take the
    staged fast path of the verify kernels; counts and levels unchanged,
    with and without hub bitmaps."""
    from conftest import clique

    cases = [(sme.gen_rmat(9, 32, seed=5), ((0, "0,0,0,2", None), (0, "0,0,0,4", "16"),
                                            (2, "0,0,0,4", "16"), (3, "0,8,0,2", None),
                                            (3, "0,0,0,4,16", "16"), (3, "32,0,0,4,8", None))),
             (clique(200), ((0, "0,0,0,4", "16"),))]
    for i, (g, runs) in enumerate(cases):
        p = str(tmp_path / f"g{i}.bin")
        sme.write_bin(p, g)
        for q, tune, hub in runs:
            want = oracle.light_count(g.row_ptrs, g.cols, q, 3)
            env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", TUNE=tune)
            cmd = [emu, p, "6", str(q), "3", "1", "0", str(g.n), str(1 << 30)] + ([hub] if hub else [])
            r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
            assert r.returncode == 0, r.stderr[-2000:]
            parts = r.stdout.split()
            assert int(parts[0].split("=")[1]) == want[0], (i, q, tune)
            assert [int(x) for x in parts[1].split("=")[1].split(",")] == want[1], (i, q, tune)


@pytest.mark.parametrize("merge", ["1", "0"])
def test_emu_server_requests(emu, graphs, merge):
    """Requests through the server (merged bucketed passes, or the plain
    one-at-a-time path) get the oracle's count for their range."""
    import numpy as np

    g, p = graphs["rmat7"]
    rng = np.random.default_rng(7)
    reqs = []
    for _ in range(24):
        q = int(rng.choice([0, 1, 2, 3, 4, 7]))
        a = int(rng.integers(0, g.n))
        reqs.append((q, a, int(rng.integers(a, g.n + 1))))
    reqs += [(0, 5, 5), (3, 0, g.n), (3, 0, g.n)]
    want = [oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=a, v_end=b)[0] for q, a, b in reqs]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", SERVE_MERGE=merge,
               SERVE=",".join(f"{q}:{a}:{b}" for q, a, b in reqs))
    r = subprocess.run([emu, p, "6", "0", "3", "1", "0", str(g.n), str(1 << 30)],
                       capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    got = [int(x) for x in r.stdout.split()[0].split("=")[1].split(",")]
    assert got == want
    passes = int(r.stdout.split()[1].split("=")[1])
    nonempty = sum(b > a for _, a, b in reqs)
    assert passes < nonempty if merge == "1" else passes == nonempty


def test_emu_match_lists(emu, graphs):
    """GPSM match lists (capped and whole) equal the oracle's rows and
    order; a LIGHT plan is rejected as unsupported."""
    import numpy as np

    g, p = graphs["rmat7"]
    for q in (0, 1, 2, 3, 4, 7):
        for algo in (4, 5):
            for cap in ((1 << 20, 17) if q in (0, 3) else (1 << 20,)):
                env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", LIST=str(cap))
                r = subprocess.run([emu, p, str(algo), str(q), "3", "1", "5", "100", str(1 << 30)],
                                   capture_output=True, text=True, env=env, timeout=600)
                assert r.returncode == 0, r.stderr[-2000:]
                parts = r.stdout.split()
                total, rows = int(parts[0].split("=")[1]), int(parts[1].split("=")[1])
                vals = parts[2].split("=")[1]
                got = [int(x) for x in vals.split(",")] if vals else []
                want, wt = oracle.gpsm_list(g.row_ptrs, g.cols, q, 5, 100, cap)
                assert total == wt and rows == len(want)
                assert got == want.reshape(-1).tolist(), (q, algo, cap)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", LIST="5")
    r = subprocess.run([emu, p, "6", "0", "3", "1", "0", "10", str(1 << 30)], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 3


def test_emu_group_count_and_rank_order(emu, graphs):
    """Group counting (segment staged in LDS as a bucket hash, lists
    streamed in tiles of 64) alone and beside the item schedule (cap 8), on
    the original ids and on the degree-ranked copy: counts and levels equal
    the oracle; full-range clique runs (and count-only runs of any plan)
    take the ranked copy, ranges do not; ranked tasks of a partition sum to
    the full count."""
    g, p = graphs["rmat7"]
    base = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")

    def go(q, env, vb=0, ve=None):
        r = subprocess.run([emu, p, "6", str(q), "3", "1", str(vb), str(g.n if ve is None else ve),
                            str(1 << 30)], capture_output=True, text=True, env=dict(base, **env), timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        parts = r.stdout.split()
        lv = parts[1].split("=")[1]
        return int(parts[0].split("=")[1]), [int(x) for x in lv.split(",")], "ranked=1" in r.stdout

    for q in (0, 3, 4, 7):
        want = oracle.light_count(g.row_ptrs, g.cols, q, 3)
        for tune, rank in (("0,0,0,4,0,512", "0"), ("0,0,0,4,0,8", "1"), ("0,0,0,1,0,512", "1")):
            cnt, lv, ranked = go(q, {"TUNE": tune, "RANK": rank})
            assert (cnt, lv) == (want[0], want[1]), (q, tune, rank)
            assert ranked == (rank == "1" and q in (0, 3, 7))
        a, b = 17, 90
        cnt, lv, ranked = go(q, {"TUNE": "0,0,0,4,0,512", "RANK": "1"}, a, b)
        assert not ranked
        assert (cnt, lv) == oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=a, v_end=b)
        cnt, _, ranked = go(q, {"TUNE": "0,0,0,4,0,512", "RANK": "1", "NOLEVELS": "1"})
        assert ranked and cnt == want[0]
        bounds = [0, 40, 41, 100, g.n]
        parts = [go(q, {"TUNE": "0,0,0,4,0,512", "RANK": "1", "RANKED": "1"}, a, b)
                 for a, b in zip(bounds[:-1], bounds[1:])]
        assert sum(x[0] for x in parts) == want[0]
        if q in (0, 3, 7):
            assert [sum(x) for x in zip(*[x[1] for x in parts])] == want[1]


def test_pool_stream_ordered_reuse(emu):
    """Shared pool: a block freed by one stream stays out of the free list
    until its free event completes, even after the freeing stream allocates
    again (otherwise another stream could overwrite data still being read)."""
    r = subprocess.run([os.path.join(EMU_DIR, "pool_test")], capture_output=True, text=True,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"), timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr[-2000:]


@pytest.mark.parametrize("q", [0, 3, 7])
def test_emu_clique_bitmaps(emu, graphs, q):
    """Clique plans over bitmaps of upper neighbourhoods: counts and levels
    against the oracle at every compress level and on a range, on a graph
    with neighbourhoods above 64 (bit rows in device memory, several batches
    under a small budget) and the same as the level engine."""
    g, p = graphs["hubs"]
    for cl in (0, 3):
        want = oracle.light_count(g.row_ptrs, g.cols, q, cl)
        assert run_emu(emu, p, 6, q, cl, 0, g.n, 1 << 30)[:2] == (want[0], want[1])
    want = oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=5, v_end=200)
    assert run_emu(emu, p, 6, q, 3, 5, 200, 1 << 30)[:2] == (want[0], want[1])
    cnt, lv, batches = run_emu(emu, p, 6, q, 3, 0, g.n, 1 << 30, env={"TUNE": "0,0,4096"})
    assert batches > 1
    assert (cnt, lv) == run_emu(emu, p, 6, q, 3, 0, g.n, 1 << 30, env={"TUNE": "0,0,0,4,0,0,0,1"})[:2]


@pytest.mark.parametrize("name", ["rmat7", "hubs", "k7"])
def test_emu_house_census(emu, graphs, name):
    """The house count from per-edge triangle counts and 4-cycle wedges
    (no level outputs, all start vertices) equals the oracle's, on the
    original ids and the ranked copy, with several wedge groups under a
    small budget."""
    g, p = graphs[name]
    want = oracle.light_count(g.row_ptrs, g.cols, 4, 3)[0]
    for env in ({"NOLEVELS": "1"}, {"NOLEVELS": "1", "RANK": "1"}):
        for budget in (1 << 30, 120_000):
            assert run_emu(emu, p, 6, 4, 3, 0, g.n, budget, env=env)[0] == want, (env, budget)
    # the diamond from the same per-edge triangle counts: count and every
    # level output (whole-graph counts) at every compress level; a budget
    # too small for the counts leaves it to the level engine
    for cl in range(4):
        want = oracle.light_count(g.row_ptrs, g.cols, 2, cl)
        for env in ({}, {"RANK": "1"}):
            for budget in (1 << 30, 30_000):
                assert run_emu(emu, p, 6, 2, cl, 0, g.n, budget, env=env)[:2] == (want[0], want[1]), (cl, env)

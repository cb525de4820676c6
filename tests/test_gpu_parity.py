"""HIP engine (libsme.so on the MI355X, through the C ABI) against the oracle
and the reference's golden counts.  Bit-exact: counts and per-level outputs
are integers and must be identical."""
import sys

import numpy as np
import pytest

import oracle
import subgraphmatchgpu_amd as sme
from conftest import clique

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    d = sme.Device(0, budget_bytes=16 << 30)
    yield d
    d.close()


@pytest.fixture(scope="module")
def small(dev, fixture_graphs):
    return {name: (g, dev.upload(g)) for name, g in fixture_graphs.items()
            if name in ("rmat8", "rmat10", "er300")}


@pytest.mark.parametrize("q", list(range(14)))
def test_light_all_compress_levels(dev, small, q):
    for name in ("rmat8", "er300"):
        g, dg = small[name]
        for cl in range(4):
            say(f"case light q={q} cl={cl} graph={name}")
            want = oracle.light_count(g.row_ptrs, g.cols, q, cl)
            res = dev.count(dg, sme.Plan(q, sme.GPU_LIGHT, cl))
            assert (res.count, res.levels) == want, (name, cl)


@pytest.mark.parametrize("q", list(range(14)))
def test_gpsm(dev, small, q):
    g, dg = small["rmat8"]
    want, wl = oracle.gpsm_count(g.row_ptrs, g.cols, q)
    for algo in (sme.GPU_GPSM, sme.GPU_GPSM_REUSE):
        say(f"case gpsm q={q} algo={algo}")
        res = dev.count(dg, sme.Plan(q, algo))
        assert res.count == want
        assert [res.levels[0]] + res.levels[2::2] == wl[:len(res.levels[2::2]) + 1]


@pytest.mark.parametrize("q", [0, 1, 2, 3, 4, 6, 7, 11, 12])
def test_light_reuse_on_off(dev, small, q):
    """Intersection reuse changes the work, never the result."""
    g, dg = small["rmat10"]
    say(f"case reuse q={q}")
    a = dev.count(dg, sme.Plan(q, reuse=True))
    b = dev.count(dg, sme.Plan(q, reuse=False))
    assert (a.count, a.levels) == (b.count, b.levels)
    assert (a.count, a.levels) == oracle.light_count(g.row_ptrs, g.cols, q, 3)


def test_golden_reference_counts(dev, ref_counts, fixture_graphs):
    """Counts printed by the reference CPU binary (tests/golden/ref_counts.json)."""
    for name, res in ref_counts["counts"].items():
        g = fixture_graphs[name]
        dg = dev.upload(g)
        for q, want in res["wcoj"].items():
            say(f"case golden graph={name} q={q}")
            got = dev.count(dg, sme.Plan(int(q), sme.GPU_LIGHT, sme.CL_SPECIAL))
            assert got.count == want, (name, q)
            if int(q) in (0, 3, 7):
                assert dev.count(dg, sme.Plan(int(q), sme.GPU_GPSM)).count == want, (name, q)
        dg.release()


def test_config_small_triangle(dev, ref_counts, fixture_graphs):
    """RMAT scale 14, edge factor 16, triangle: exact count and per-level
    partial instances {n, nnz/2, nnz/2, T}."""
    g = fixture_graphs["rmat14"]
    dg = dev.upload(g)
    res = dev.count(dg, sme.Plan(sme.Q0))
    T = ref_counts["counts"]["rmat14"]["wcoj"]["0"]
    assert res.count == T
    assert res.levels == [g.n, g.nnz // 2, g.nnz // 2, T]
    dg.release()


def test_cliques_closed_form(dev):
    import math

    g = clique(12)
    dg = dev.upload(g)
    C = math.comb
    forms = {0: C(12, 3), 1: 3 * C(12, 4), 2: 6 * C(12, 4), 3: C(12, 4), 4: 60 * C(12, 5),
             7: C(12, 5), 11: C(12, 6)}
    for q, want in forms.items():
        say(f"case clique q={q}")
        for cl in range(4):
            assert dev.count(dg, sme.Plan(q, sme.GPU_LIGHT, cl)).count == want
    dg.release()


def test_batching_invariance(fixture_graphs):
    """A tight device budget forces the level engine into many batches;
    counts and per-level outputs must not change (nor with the clique
    bitmap path under the same budget)."""
    g = fixture_graphs["rmat12"]
    big = sme.Device(0, budget_bytes=8 << 30)
    tiny = sme.Device(0, budget_bytes=3 << 20)
    gb, gt = big.upload(g), tiny.upload(g)
    for engine in (1, 0):
        for d in (big, tiny):
            d.set_tuning(super_chunks=4, clique_level_engine=engine, census_level_engine=engine)
        for q in (0, 2, 3, 4, 7):
            say(f"case batching q={q} level_engine={engine}")
            a = big.count(gb, sme.Plan(q))
            b = tiny.count(gt, sme.Plan(q))
            assert (a.count, a.levels) == (b.count, b.levels)
            if engine or q == 4:
                assert b.batches > a.batches
    big.close()
    tiny.close()


def test_start_ranges(dev, small):
    g, dg = small["rmat10"]
    cuts = [0, 1, 7, 300, 301, 700, g.n]
    for q in (0, 1, 2, 4, 7):
        say(f"case ranges q={q}")
        whole = dev.count(dg, sme.Plan(q)).count
        parts = 0
        for a, b in zip(cuts[:-1], cuts[1:]):
            r = dev.count(dg, sme.Plan(q), a, b)
            want = oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=a, v_end=b)
            assert (r.count, r.levels) == want
            parts += r.count
        assert parts == whole
    assert dev.count(dg, sme.Plan(0), 5, 5).count == 0
    assert dev.count(dg, sme.Plan(0), 10, 3).count == 0


def test_hub_bitmaps(fixture_graphs):
    """Membership through adjacency bitmaps gives the same counts and levels."""
    g = fixture_graphs["rmat10"]
    d = sme.Device(0, budget_bytes=4 << 30)
    d.set_hub_bitmaps(16, 1 << 30)
    dg = d.upload(g)
    assert dg.hubs > 0
    for q in (0, 1, 2, 3, 4, 7):
        say(f"case hubs q={q}")
        want = oracle.light_count(g.row_ptrs, g.cols, q, 3)
        r = d.count(dg, sme.Plan(q))
        assert (r.count, r.levels) == want
    d.close()


def test_buckets_match_ranges(dev, small):
    """One pass with per-start-vertex buckets equals one call per range."""
    g, dg = small["rmat10"]
    bounds = [0, 1, 7, 7, 300, 301, 700, g.n]
    for q in (0, 1, 2, 3, 4, 7):
        for algo in (sme.GPU_LIGHT, sme.GPU_GPSM):
            say(f"case buckets q={q} algo={algo}")
            plan = sme.Plan(q, algo)
            got, whole = dev.count_buckets(dg, plan, bounds)
            want = [dev.count(dg, plan, a, b).count for a, b in zip(bounds[:-1], bounds[1:])]
            assert got.tolist() == want
            assert whole.count == sum(want)


def test_edge_graphs(dev):
    empty = sme.from_adjacency([[] for _ in range(10)])
    dg = dev.upload(empty)
    for q in range(14):
        assert dev.count(dg, sme.Plan(q)).count == 0
    star = sme.from_adjacency([list(range(1, 50))] + [[0] for _ in range(49)])
    ds = dev.upload(star)
    assert dev.count(ds, sme.Plan(0)).count == 0
    assert dev.count(ds, sme.Plan(1)).count == 0


def test_rejects_bad_csr(dev):
    bad = sme.Graph(np.array([0, 2, 3, 4], dtype=np.uint64), np.array([2, 1, 0, 0], dtype=np.uint32))
    with pytest.raises(sme.SmeError) as e:
        dev.upload(bad)
    assert e.value.status == 1


def test_self_loops_handled(dev):
    # a self loop must not satisfy non-equality conditions
    adj = [[0, 1, 2], [0, 1, 2], [0, 1, 2]]
    g = sme.from_adjacency(adj)
    dg = dev.upload(g)
    for q in (0, 1, 2, 3, 4):
        want = oracle.wcoj_count(g.row_ptrs, g.cols, q)
        lw = oracle.light_count(g.row_ptrs, g.cols, q, 3)
        got = dev.count(dg, sme.Plan(q))
        assert (got.count, got.levels) == lw
        assert got.count == want


@pytest.mark.parametrize("tune", [dict(lane_path_max=8), dict(lane_path_max=1 << 20, item_grid=64),
                                  dict(item_grid=1, batch_bytes=1 << 16), dict(super_chunks=2),
                                  dict(super_chunks=16, lane_path_max=8, item_grid=512),
                                  dict(stream_ratio=16), dict(stream_ratio=1000, lane_path_max=32),
                                  dict(wg_path_min=8), dict(wg_path_min=64, lane_path_max=16, group_max=8),
                                  dict(wg_path_min=2, item_grid=3)])
def test_tunings_do_not_change_results(fixture_graphs, tune):
    # clique plans too run through the level engine these tunables shape
    tune = dict(tune, clique_level_engine=1, census_level_engine=1)
    g = fixture_graphs["rmat10"]
    d = sme.Device(0, budget_bytes=2 << 30)
    d.set_tuning(**tune)
    assert {k: d.get_tuning()[k] for k in tune} == tune
    dg = d.upload(g)
    bounds = [0, 5, 100, 400, g.n]
    for q in range(14):
        say(f"case tuning {tune} q={q}")
        want = oracle.light_count(g.row_ptrs, g.cols, q, 3)
        r = d.count(dg, sme.Plan(q))
        assert (r.count, r.levels) == want
        if q in (0, 1, 3, 4, 7):
            got, _ = d.count_buckets(dg, sme.Plan(q), bounds)
            assert got.tolist() == [oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=a, v_end=b)[0]
                                    for a, b in zip(bounds[:-1], bounds[1:])]
    d.close()


_ORACLE_CACHE = {}


def light_oracle(name, g, q, cl=3):
    key = (name, q, cl)
    if key not in _ORACLE_CACHE:
        _ORACLE_CACHE[key] = oracle.light_count(g.row_ptrs, g.cols, q, cl)
    return _ORACLE_CACHE[key]


@pytest.mark.parametrize("super_chunks,stream_ratio", [(2, 0), (4, 0), (16, 0), (4, 8), (2, 1000)])
def test_super_chunks_with_hubs(fixture_graphs, super_chunks, stream_ratio):
    """Long single-path runs (staged fast path) with and without hub rows."""
    g = fixture_graphs["rmat12"]
    for hub in (0, 32):
        d = sme.Device(0, budget_bytes=2 << 30)
        d.set_hub_bitmaps(hub, 1 << 30)
        d.set_tuning(super_chunks=super_chunks, stream_ratio=stream_ratio, clique_level_engine=1,
                     census_level_engine=1)
        dg = d.upload(g)
        for q in (0, 1, 2, 3, 4, 7):
            say(f"case super_chunks={super_chunks} stream_ratio={stream_ratio} hub={hub} q={q}")
            want = light_oracle("rmat12", g, q)
            r = d.count(dg, sme.Plan(q))
            assert (r.count, r.levels) == want
        d.close()


def test_er_batch_members_one_pass(dev):
    """Packed ER graphs: one bucketed pass gives each member's triangle and
    4-cycle count, equal to the WCOJ oracle on the member alone and to one
    pass per member."""
    from conftest import member

    G, offs = sme.gen_er_batch(300, (64, 512), (4.0, 48.0), seed=11)
    dg = dev.upload(G)
    for q in (sme.Q0, sme.Q1):
        say(f"case er-batch q={q}")
        got, whole = dev.count_buckets(dg, sme.Plan(q), offs)
        want = [oracle.wcoj_count(m.row_ptrs, m.cols, q) for m in (member(G, offs, k) for k in range(300))]
        assert got.tolist() == want
        assert whole.count == sum(want)
        for k in range(0, 300, 37):
            assert dev.count(dg, sme.Plan(q), int(offs[k]), int(offs[k + 1])).count == want[k]
    dg.release()


def test_autotune_installs_a_measured_setting(fixture_graphs):
    g = fixture_graphs["rmat12"]
    d = sme.Device(0, budget_bytes=2 << 30)
    dg = d.upload(g)
    plans = [sme.Plan(sme.Q3), sme.Plan(sme.Q2)]
    want = [d.count(dg, p).count for p in plans]
    say("case autotune rmat12")
    best, ms, trials = d.autotune(dg, plans, sample_fraction=0.05)
    assert d.get_tuning() == best
    # the installed setting is a measured one; a setting replaces the
    # current one only when at least 2 % faster, once per dimension
    assert len(trials) >= 8 and dict(best, ms=ms) in trials
    assert ms <= min(t["ms"] for t in trials) / 0.98 ** 6
    assert [d.count(dg, p).count for p in plans] == want
    G, offs = sme.gen_er_batch(64, (64, 256), (4.0, 16.0), seed=2)
    dgb = d.upload(G)
    say("case autotune er-batch")
    best, ms, trials = d.autotune(dgb, [sme.Plan(sme.Q0), sme.Plan(sme.Q1)], bounds=offs, sample_fraction=0.25)
    assert d.get_tuning() == best
    d.close()


_PER_VERTEX = {}


def oracle_prefix(name, g, q, lo=0, hi=None):
    """Prefix sums of the oracle's per-start-vertex LIGHT counts over
    [lo, hi): the count of a range [a, b) inside it is P[b - lo] - P[a - lo]."""
    hi = g.n if hi is None else hi
    key = (name, q, lo, hi)
    if key not in _PER_VERTEX:
        pv = oracle.light_count_per_vertex(g.row_ptrs, g.cols, q, lo, hi)
        _PER_VERTEX[key] = np.concatenate([[0], np.cumsum(pv, dtype=np.uint64)]).astype(np.uint64)
    return _PER_VERTEX[key]


@pytest.mark.parametrize("slots,merge,hub", [(4, True, 0), (4, True, 64), (1, False, 0), (2, True, 0)])
def test_server_results_alone_and_in_the_mix(fixture_graphs, slots, merge, hub):
    """Each request's count from the server (merged bucketed passes running
    concurrently on several streams of one shared pool) equals its count
    alone through sme_count and the oracle's count of its range."""
    g = fixture_graphs["rmat12"]
    rng = np.random.default_rng(100 + slots)
    reqs = []
    for _ in range(120):
        q = int(rng.choice([0, 1, 2, 3, 4, 7], p=[0.3, 0.2, 0.2, 0.15, 0.1, 0.05]))
        ln = max(1, int(rng.uniform(0.01, 1.0) * g.n))
        a = int(rng.integers(0, g.n - ln + 1))
        reqs.append((q, a, a + ln))
    d = sme.Device(0, budget_bytes=2 << 30)
    dg = d.upload(g)
    alone = [d.count(dg, sme.Plan(q), a, b).count for q, a, b in reqs]
    d.close()
    say(f"case server slots={slots} merge={merge} hub={hub}")
    srv = sme.Server(g, mem_cap_bytes=2 << 30, slots=slots, merge=merge, min_reserve_bytes=64 << 20,
                     hub_min_degree=hub, hub_max_bytes=1 << 30 if hub else 0)
    ids = [srv.submit(q, a, b) for q, a, b in reqs]
    got = [srv.wait(i)[0] for i in ids]
    st = srv.stats()
    srv.close()
    assert got == alone
    want = [int(oracle_prefix("rmat12", g, q)[b] - oracle_prefix("rmat12", g, q)[a]) for q, a, b in reqs]
    assert got == want
    assert st["requests_done"] == len(reqs) and st["failed"] == 0
    if merge:
        assert st["passes"] < len(reqs)
    assert st["peak_reserved"] <= 2 << 30


@pytest.fixture(scope="module")
def rmat22():
    """R-MAT scale 22, edge factor 32 (the memory-capped and the served
    configurations' graph), generated once per module."""
    say("generating rmat22 ef32")
    g = sme.gen_rmat(22, 32, seed=1)
    assert g.n == 1 << 22
    return g


# Start-vertex windows of R-MAT 22 ef32 where every pattern has instances and
# the oracle counts a window in about a second (R-MAT degrees fall with the
# number of set bits of the id, not with the id).
R22_WINDOWS = {q: [(2500000, 2501024), (2900000, 2900512)] for q in (0, 1, 2, 3, 7)}
R22_WINDOWS[4] = [(2500000, 2500512)]


def test_server_rmat22_requests_vs_oracle(rmat22):
    """The served configuration's graph stays resident in the server; a
    seeded request mix with the served pattern weights, each request a random
    start range inside a window the oracle finishes in seconds, runs through
    4 concurrent slots with merging.  Every count equals the oracle's count
    of its range and the count of the same request alone through sme_count."""
    g = rmat22
    rng = np.random.default_rng(22)
    reqs = []
    for _ in range(60):
        q = int(rng.choice([0, 1, 2, 3, 4, 7], p=[0.3, 0.2, 0.2, 0.15, 0.1, 0.05]))
        wins = R22_WINDOWS[q]
        lo, hi = wins[int(rng.integers(0, len(wins)))]
        ln = max(1, int(np.ceil(rng.uniform(0.01, 1.0) * (hi - lo))))
        a = int(rng.integers(lo, hi - ln + 1))
        reqs.append((q, a, a + ln))
    say("case server rmat22: oracle counts of the windows")
    want = []
    for q, a, b in reqs:
        lo, hi = next(w for w in R22_WINDOWS[q] if w[0] <= a and b <= w[1])
        P = oracle_prefix("rmat22", g, q, lo, hi)
        want.append(int(P[b - lo] - P[a - lo]))
    assert sum(1 for w in want if w) >= 30
    say("case server rmat22: served")
    srv = sme.Server(g, mem_cap_bytes=64 << 30, slots=4, merge=True, min_reserve_bytes=1 << 30,
                     hub_min_degree=256, hub_max_bytes=4 << 30)
    ids = [srv.submit(q, a, b) for q, a, b in reqs]
    got = [srv.wait(i)[0] for i in ids]
    st = srv.stats()
    srv.close()
    assert got == want
    assert st["requests_done"] == len(reqs) and st["failed"] == 0 and st["passes"] < len(reqs)
    say("case server rmat22: alone")
    d = sme.Device(0, budget_bytes=16 << 30)
    dg = d.upload(g)
    alone = [d.count(dg, sme.Plan(q), a, b).count for q, a, b in reqs]
    d.close()
    assert alone == want


def test_config2_rmat20_full(config_counts):
    """Configuration 2 at full size: R-MAT scale 20, edge factor 16, every
    start vertex, 4-clique and diamond.  Counts and per-level outputs equal
    the committed fixture (independent census, and the reference binary's
    count where it finished), with and without hub bitmaps; GPSM_REUSE agrees
    with LIGHT on a start range."""
    ent = config_counts["rmat20"]
    g = sme.gen_rmat(20, 16, seed=1)
    assert g.checksum() == ent["sha256"]
    for q, ref in ent.get("reference", {}).items():
        assert ref["count"] == ent["light"][q]["count"]
    for hub in (256, 0):
        d = sme.Device(0, budget_bytes=32 << 30)
        if hub:
            d.set_hub_bitmaps(hub, 8 << 30)
        dg = d.upload(g)
        for q in (sme.Q3, sme.Q2):
            say(f"case config2 q={q} hub={hub}")
            r = d.count(dg, sme.Plan(q))
            want = ent["light"][str(q)]
            assert (r.count, r.levels) == (want["count"], want["levels"])
        if hub:
            a, b = 1 << 16, 1 << 20
            say(f"case config2 gpsm_reuse q=3 range=[{a},{b})")
            light = d.count(dg, sme.Plan(sme.Q3), a, b).count
            assert d.count(dg, sme.Plan(sme.Q3, sme.GPU_GPSM_REUSE), a, b).count == light
        d.close()
    # the benchmark's configuration: degree-ranked copy and group counting
    d = sme.Device(0, budget_bytes=32 << 30)
    d.set_hub_bitmaps(256, 8 << 30)
    d.set_rank_order(True)
    dg = d.upload(g)
    for engine in (0, 1):
        d.set_tuning(super_chunks=4, group_max=512, clique_level_engine=engine, census_level_engine=engine)
        for q in (sme.Q3, sme.Q2):
            say(f"case config2 ranked q={q} clique_level_engine={engine}")
            r = d.count(dg, sme.Plan(q))
            want = ent["light"][str(q)]
            assert (r.count, r.levels) == (want["count"], want["levels"])
            # the diamond's census runs on the ranked copy too
            assert r.ranked == (q == sme.Q3 or not engine)
    d.close()


def test_config3_rmat22_slices(rmat22):
    """RMAT scale 22, edge factor 32 (the memory-capped configuration), full
    CSR resident, 5-clique and house on start-vertex slices: counts and
    per-level outputs equal the LIGHT oracle where it finishes in seconds;
    on a heavier slice a 32 GB and a 64 MB budget give identical counts and
    levels (the smaller in more batches) and GPSM_REUSE agrees on the
    count."""
    g = rmat22
    big = sme.Device(0, budget_bytes=32 << 30)
    small = sme.Device(0, budget_bytes=64 << 20)
    gb, gs = big.upload(g), small.upload(g)
    for q, (a, b) in ((sme.Q7, (3 << 20, (3 << 20) + 64)), (sme.Q4, (2500000, 2500256))):
        say(f"case config3 oracle q={q} range=[{a},{b})")
        r = big.count(gb, sme.Plan(q), a, b)
        assert (r.count, r.levels) == oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=a, v_end=b)
    # batching is the level engine's; the clique bitmap path must agree with it
    for d in (big, small):
        d.set_tuning(super_chunks=4, clique_level_engine=1)
    for q, (a, b) in ((sme.Q7, (3 << 20, (3 << 20) + 256)), (sme.Q4, (2500000, 2501024))):
        say(f"case config3 budgets q={q} range=[{a},{b})")
        r1 = big.count(gb, sme.Plan(q), a, b)
        r2 = small.count(gs, sme.Plan(q), a, b)
        assert (r1.count, r1.levels) == (r2.count, r2.levels)
        assert r2.batches > r1.batches
        if q == sme.Q7:
            assert big.count(gb, sme.Plan(q, sme.GPU_GPSM_REUSE), a, b).count == r1.count
            small.set_tuning(super_chunks=4)
            r3 = small.count(gs, sme.Plan(q), a, b)
            assert (r3.count, r3.levels) == (r1.count, r1.levels)
            small.set_tuning(super_chunks=4, clique_level_engine=1)
    big.close()
    small.close()


@pytest.fixture(scope="module")
def ranked_dev(fixture_graphs):
    """A device whose uploads keep the degree-ranked copy."""
    d = sme.Device(0, budget_bytes=8 << 30)
    d.set_rank_order(True)
    graphs = {name: (g, d.upload(g)) for name, g in fixture_graphs.items()
              if name in ("rmat10", "rmat12", "er300")}
    yield d, graphs
    d.close()


@pytest.mark.parametrize("q", list(range(14)))
def test_group_count_all_queries(ranked_dev, q):
    """Counting levels that extend a reused segment with one list, done by
    the group schedule (segment staged in LDS, lists streamed), give the
    oracle's count and levels, alone and mixed with the item schedule (a
    small cap leaves the long segments to it), on the original ids and on
    the ranked copy."""
    d, graphs = ranked_dev
    for name in ("rmat10", "er300"):
        g, dg = graphs[name]
        want = oracle.light_count(g.row_ptrs, g.cols, q, 3)
        for gm in (8, 512):
            d.set_tuning(super_chunks=4, group_max=gm, clique_level_engine=1, census_level_engine=1)
            say(f"case group q={q} graph={name} group_max={gm}")
            r = d.count(dg, sme.Plan(q))
            assert (r.count, r.levels) == want, (name, gm)
            assert r.ranked == (q in (0, 3, 7, 11))
            a, b = g.n // 5, g.n // 2
            r = d.count(dg, sme.Plan(q), a, b)
            assert not r.ranked
            assert (r.count, r.levels) == oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=a, v_end=b)
    d.set_tuning(super_chunks=4)


@pytest.mark.parametrize("q", [0, 1, 2, 3, 4, 7, 11])
def test_rank_order_counts(ranked_dev, q):
    """Over all start vertices a clique's count and levels, and any
    pattern's count alone, are the same on the degree-ranked copy; the
    ranked tasks of a partition of [0, n) sum to the full count (and, for
    a clique, to its levels)."""
    d, graphs = ranked_dev
    g, dg = graphs["rmat12"]
    want, wl = oracle.light_count(g.row_ptrs, g.cols, q, 3)
    for engine in (0, 1):
        d.set_tuning(super_chunks=4, group_max=512, clique_level_engine=engine, census_level_engine=engine)
        say(f"case rank q={q} clique_level_engine={engine}")
        r = d.count_only(dg, sme.Plan(q))
        assert r.ranked and r.count == want
        bounds = [0, 7, g.n // 3, g.n // 3 + 1, g.n - 50, g.n]
        parts = [d.count_ranked(dg, sme.Plan(q), a, b) for a, b in zip(bounds[:-1], bounds[1:])]
        assert sum(p.count for p in parts) == want
        if q in (0, 3, 7, 11):
            assert [sum(x) for x in zip(*[p.levels for p in parts])] == wl
    d.set_tuning(super_chunks=4)


@pytest.mark.parametrize("q", [0, 1, 2, 3, 4, 7])
def test_match_lists(dev, small, q):
    """GPSM match lists equal the oracle's rows, in the same (lexicographic)
    order, whole and capped; LIGHT plans are rejected."""
    name = "er300" if q == 4 else "rmat8"
    g, dg = small[name]
    want, wt = oracle.gpsm_list(g.row_ptrs, g.cols, q)
    for algo in (sme.GPU_GPSM, sme.GPU_GPSM_REUSE):
        say(f"case list q={q} algo={algo} graph={name}")
        rows, total = dev.list_matches(dg, sme.Plan(q, algo), max_matches=1 << 20)
        assert total == wt and np.array_equal(rows, want)
        rows, total = dev.list_matches(dg, sme.Plan(q, algo), 7, g.n // 2, max_matches=11)
        w2, t2 = oracle.gpsm_list(g.row_ptrs, g.cols, q, 7, g.n // 2, 11)
        assert total == t2 and np.array_equal(rows, w2)
    with pytest.raises(sme.SmeError) as e:
        dev.list_matches(dg, sme.Plan(q), max_matches=4)
    assert e.value.status == 5


def test_cli_counts_and_match_list(tmp_path, fixture_graphs):
    """The CLI's result line carries the reference's count, and -l 1 prints
    the GPSM match list in the oracle's order."""
    import subprocess

    g = fixture_graphs["rmat8"]
    p = str(tmp_path / "g.bin")
    sme.write_bin(p, g)
    say("case cli -a 4 -q 3 -l 1")
    r = subprocess.run([sme.CLI_PATH, "-f", p, "-a", "4", "-q", "3", "-m", "1", "-l", "1", "-L", "100000"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    i = next(k for k, line in enumerate(lines) if line.startswith("total_match_count="))
    want, total = oracle.gpsm_list(g.row_ptrs, g.cols, 3)
    assert int(lines[i].split(",")[0].split("=")[1]) == total
    got = [[int(x) for x in line.split()] for line in lines[i + 1:]]
    assert got == want.tolist()


def hub_graph(n, extra_m, seed):
    """Vertex 0 adjacent to every other vertex, plus extra_m random edges
    among the others: vertex 0's upper neighbourhood exceeds what the clique
    bitmap path takes (16,384), so the level engine counts it."""
    rng = np.random.default_rng(seed)
    adj = [set() for _ in range(n)]
    for v in range(1, n):
        adj[0].add(v)
        adj[v].add(0)
    for a, b in rng.integers(1, n, size=(extra_m, 2)):
        if a != b:
            adj[a].add(int(b))
            adj[b].add(int(a))
    return sme.from_adjacency([sorted(s) for s in adj])


@pytest.fixture(scope="module")
def clique_graphs(fixture_graphs):
    # upper neighbourhoods up to 355 (rmat10, low ids are hubs), 153 (dense
    # G(n, m)) and 16,999 (hub17k)
    return {"rmat10": fixture_graphs["rmat10"], "er300d": sme.gen_er(300, 20000, seed=4),
            "hub17k": hub_graph(17000, 60000, 1)}


@pytest.mark.parametrize("q", [0, 3, 7, 11])
def test_clique_bitmaps(clique_graphs, q):
    """Clique plans counted over bitmaps of each start vertex's upper
    neighbourhood (one wave per vertex up to 64 upper neighbours, bit rows in
    device memory beyond that, the level engine past 16,384) give the
    oracle's count and per-level outputs at every compress level, on start
    ranges and per bucket, in several row batches under a small batch cap,
    and the same as the level engine."""
    d = sme.Device(0, budget_bytes=2 << 30)
    for name, g in clique_graphs.items():
        dg = d.upload(g)
        d.set_tuning(super_chunks=4)
        for cl in range(4):
            say(f"case clique bitmaps q={q} graph={name} cl={cl}")
            r = d.count(dg, sme.Plan(q, sme.GPU_LIGHT, cl))
            assert (r.count, r.levels) == oracle.light_count(g.row_ptrs, g.cols, q, cl), cl
        a, b = 1, g.n // 2
        say(f"case clique bitmaps q={q} graph={name} range=[{a},{b})")
        r = d.count(dg, sme.Plan(q), a, b)
        assert (r.count, r.levels) == oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=a, v_end=b)
        bounds = [0, 1, 5, g.n // 3, g.n // 3 + 1, g.n - 7, g.n]
        got, tot = d.count_buckets(dg, sme.Plan(q), bounds)
        assert got.tolist() == [oracle.light_count(g.row_ptrs, g.cols, q, 3, v_begin=x, v_end=y)[0]
                                for x, y in zip(bounds[:-1], bounds[1:])]
        full = d.count(dg, sme.Plan(q))
        assert tot.count == full.count
        # 8 KB: less than the rows of one vertex with 355 upper neighbours
        d.set_tuning(super_chunks=4, batch_bytes=1 << 13)
        r = d.count(dg, sme.Plan(q))
        assert (r.count, r.levels) == (full.count, full.levels)
        assert r.batches > 1
        d.set_tuning(super_chunks=4, clique_level_engine=1)
        r = d.count(dg, sme.Plan(q))
        assert (r.count, r.levels) == (full.count, full.levels)
        d.set_tuning(super_chunks=4)
    d.close()


@pytest.mark.parametrize("ranked", [False, True])
def test_house_census(fixture_graphs, ranked):
    """A house count over all start vertices with no level outputs comes from
    per-edge triangle counts and a degree-ordered 4-cycle wedge walk (light
    apexes hashed in LDS, heavy ones in dense per-apex arrays, in several
    groups under an 8 MB budget).  It equals the oracle's LIGHT count, the
    closed form 60 C(n, 5) on K12 and the level engine's count, on the
    original ids and on the degree-ranked copy; start ranges still go
    through the level engine."""
    graphs = {"rmat10": fixture_graphs["rmat10"], "rmat12": fixture_graphs["rmat12"],
              "er300d": sme.gen_er(300, 20000, seed=4), "k12": clique(12)}
    for budget in (2 << 30, 8 << 20):
        d = sme.Device(0, budget_bytes=budget)
        d.set_rank_order(ranked)
        for name, g in graphs.items():
            dg = d.upload(g)
            want = 60 * 792 if name == "k12" else light_oracle(name, g, sme.Q4)[0]
            say(f"case house census graph={name} ranked={ranked} budget={budget}")
            r = d.count_only(dg, sme.Plan(sme.Q4))
            assert r.count == want
            assert r.ranked == ranked
            d.set_tuning(super_chunks=4, census_level_engine=1)
            assert d.count_only(dg, sme.Plan(sme.Q4)).count == want
            d.set_tuning(super_chunks=4)
            a, b = 3, g.n // 2
            assert d.count_only(dg, sme.Plan(sme.Q4), a, b).count == \
                oracle.light_count(g.row_ptrs, g.cols, sme.Q4, 3, v_begin=a, v_end=b)[0]
            # the diamond from the same per-edge triangle counts, with its
            # level outputs (whole-graph counts) at every compress level
            for cl in range(4):
                say(f"case diamond census graph={name} ranked={ranked} budget={budget} cl={cl}")
                r = d.count(dg, sme.Plan(sme.Q2, sme.GPU_LIGHT, cl))
                assert (r.count, r.levels) == light_oracle(name, g, sme.Q2, cl)
            dg.release()
        d.close()


def test_house_census_rmat16():
    """R-MAT scale 16, edge factor 16 (the serving graph): the house count
    from edge statistics equals the level engine's LIGHT count."""
    g = sme.gen_rmat(16, 16, seed=1)
    d = sme.Device(0, budget_bytes=16 << 30)
    d.set_rank_order(True)
    dg = d.upload(g)
    say("case house census rmat16")
    r = d.count_only(dg, sme.Plan(sme.Q4))
    d.set_tuning(super_chunks=4, census_level_engine=1)
    assert d.count_only(dg, sme.Plan(sme.Q4)).count == r.count
    d.close()

"""The no-allocation entry point (sme_count_static) against sme_count, both
through the C ABI: counts and per-level outputs identical on two calls in a
row (the second on a smaller input, so the shared scratch holds the first
call's contents when it starts), and a call whose one unsplittable batch
exceeds the scratch is rejected with SME_ERR_OOM naming the limit."""
import sys

import numpy as np
import pytest

import subgraphmatchgpu_amd as sme

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_static_scratch_matches_count(fixture_graphs):
    d = sme.Device(0, budget_bytes=4 << 30)
    big, small = fixture_graphs["rmat12"], fixture_graphs["rmat10"]
    dgb, dgs = d.upload(big), d.upload(small)
    for g, dg, q, a, b in ((big, dgb, sme.Q4, 0, big.n), (small, dgs, sme.Q3, 100, 700)):
        say(f"case no-alloc q={q} n={g.n} range=[{a},{b})")
        p = sme.Plan(q)
        want = d.count(dg, p, a, b)
        got = d.count_static(dg, p, a, b)
        assert (got.count, got.levels) == (want.count, want.levels)
        assert got.peak_pool_bytes <= sme.library().sme_static_scratch_bytes()
    d.close()


def test_static_scratch_rejects_an_oversized_batch():
    """A star with 2^24 leaves: the triangle's MATERIALIZE of the centre's
    candidate window is one path whose children cannot be split across
    batches, and they need far more than the 64 MiB scratch."""
    n = (1 << 24) + 1
    leaves = np.arange(1, n, dtype=np.uint32)
    rp = np.concatenate([[0, n - 1], n - 1 + np.arange(1, n, dtype=np.uint64)]).astype(np.uint64)
    cols = np.concatenate([leaves, np.zeros(n - 1, dtype=np.uint32)])
    g = sme.Graph(rp, cols)
    d = sme.Device(0, budget_bytes=4 << 30)
    dg = d.upload(g)
    say("case no-alloc star: expect rejection")
    assert d.count(dg, sme.Plan(sme.Q0)).count == 0
    with pytest.raises(sme.SmeError) as e:
        d.count_static(dg, sme.Plan(sme.Q0))
    assert e.value.status == 2 and "no-allocation scratch" in str(e.value)
    # the entry point still works after the rejection
    assert d.count_static(dg, sme.Plan(sme.Q0), 1, 1000).count == 0
    d.close()

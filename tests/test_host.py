"""Host-side product logic: plans, BIN IO, generators, the C ABI surface."""
import ctypes
import os
import re

import numpy as np
import pytest

import subgraphmatchgpu_amd as sme
from conftest import GOLDEN, ROOT


def test_library_exports_every_declared_symbol():
    lib = sme.library()
    with open(os.path.join(ROOT, "include", "sme.h")) as f:
        text = f.read()
    names = set(re.findall(r"\b(sme_[a-z_0-9]+)\s*\(", text))
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_product_plans_match_reference_plans():
    with open(os.path.join(GOLDEN, "light_plans.txt")) as f:
        golden = f.read()
    mine = "".join(sme.Plan(q, sme.GPU_LIGHT, cl, o).describe()
                   for o in (True, False) for q in range(14) for cl in range(4))
    assert mine == golden


def test_plan_levels_and_errors():
    assert sme.Plan(sme.Q0).levels == 4
    assert sme.Plan(sme.Q4).levels == 8
    assert sme.Plan(sme.Q3, sme.GPU_GPSM).levels == 7  # M0 C1 M1 C2 M2 C3 M3
    with pytest.raises(sme.SmeError) as e:
        sme.Plan(14)
    assert e.value.status == 1
    with pytest.raises(sme.SmeError) as e:
        sme.Plan(0, algo=0)
    assert e.value.status == 5
    with pytest.raises(sme.SmeError):
        sme.Plan(0, compress_level=7)


def test_bin_roundtrip_and_layout(tmp_path):
    g = sme.gen_rmat(8, 8, seed=1)
    p = str(tmp_path / "g.bin")
    sme.write_bin(p, g)
    raw = open(p, "rb").read()
    hdr = np.frombuffer(raw[:32], dtype=np.uint64)
    assert list(hdr) == [4, 8, g.n, g.nnz]  # GraphIO.h:104-127
    assert len(raw) == 32 + 8 * (g.n + 1) + 4 * g.nnz
    h = sme.read_bin(p)
    assert np.array_equal(h.row_ptrs, g.row_ptrs) and np.array_equal(h.cols, g.cols)


def test_bin_errors(tmp_path):
    p = tmp_path / "bad.bin"
    p.write_bytes(np.array([4, 4, 3, 2], dtype=np.uint64).tobytes())
    with pytest.raises(sme.SmeError):
        sme.read_bin(str(p))
    p.write_bytes(np.array([4, 8, 3, 2], dtype=np.uint64).tobytes())  # truncated body
    with pytest.raises(sme.SmeError):
        sme.read_bin(str(p))
    with pytest.raises(sme.SmeError):
        sme.read_bin(str(tmp_path / "missing.bin"))
    # header sizes that the file cannot hold are rejected before any
    # allocation (n = 2^64 - 1 would wrap n + 1 to 0)
    for n, nnz in ((2**64 - 1, 0), (2**40, 2**40), (3, 2**62), (1, 3)):
        p.write_bytes(np.array([4, 8, n, nnz], dtype=np.uint64).tobytes() + bytes(8 * 2))
        with pytest.raises(sme.SmeError) as e:
            sme.read_bin(str(p))
        assert e.value.status == 1


def test_snap_reader(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("# comment\n10 11\n11 12\n10 12\n12 13\n")
    g = sme.read_snap(str(p), directed=False)
    # ids shifted by the minimum id, both directions, lists sorted (GraphIO.h:232-311)
    assert g.n == 4 and g.nnz == 8
    assert list(g.row_ptrs) == [0, 2, 4, 7, 8]
    assert list(g.cols) == [1, 2, 0, 2, 0, 1, 3, 2]


def test_generator_determinism_and_shape(ref_counts):
    for name, spec in ref_counts["graphs"].items():
        if spec["kind"] == "rmat" and spec["scale"] > 12:
            continue
        from conftest import make_fixture_graph

        g = make_fixture_graph(spec)  # asserts the committed checksum
        assert g.n == spec["n"] and g.nnz == spec["nnz"]
        rp, cols = g.row_ptrs, g.cols
        assert rp[0] == 0 and rp[-1] == g.nnz and np.all(np.diff(rp.astype(np.int64)) >= 0)
        for u in range(0, g.n, max(1, g.n // 64)):
            lst = cols[rp[u]:rp[u + 1]]
            assert np.all(np.diff(lst.astype(np.int64)) > 0)  # sorted, no duplicates
            assert u not in set(lst.tolist())                 # no self loops
    # symmetric
    g = sme.gen_rmat(9, 8, seed=4)
    src = np.repeat(np.arange(g.n, dtype=np.uint64), np.diff(g.row_ptrs.astype(np.int64)))
    fwd = set(zip(src.tolist(), g.cols.tolist()))
    assert all((b, a) in fwd for a, b in fwd)
    # different seeds differ; same seed identical
    assert sme.gen_rmat(9, 8, seed=4).checksum() == g.checksum()
    assert sme.gen_rmat(9, 8, seed=5).checksum() != g.checksum()


def test_er_generator_edge_count():
    g = sme.gen_er(1000, 5000, seed=9)
    assert g.n == 1000 and g.nnz == 2 * 5000


def test_from_adjacency_and_clique():
    from conftest import clique

    g = clique(5)
    assert g.n == 5 and g.nnz == 20


def test_cli_usage_runs():
    import subprocess

    r = subprocess.run([sme.CLI_PATH], capture_output=True, text=True)
    assert r.returncode == 1 and "patternmatchgpu" in r.stdout


def test_er_batch_generator_layout():
    from conftest import member

    G, offs = sme.gen_er_batch(64, (16, 200), (4.0, 20.0), seed=7)
    assert offs[0] == 0 and offs[-1] == G.n and len(offs) == 65
    sizes = np.diff(offs.astype(np.int64))
    assert sizes.min() >= 16 and sizes.max() <= 200
    assert len(set(sizes.tolist())) > 10
    rp, cols = G.row_ptrs, G.cols
    for k in range(64):
        a, b = int(offs[k]), int(offs[k + 1])
        seg = cols[int(rp[a]):int(rp[b])]
        assert seg.size == 0 or (seg.min() >= a and seg.max() < b)  # disjoint union
        m = member(G, offs, k)
        deg = 2 * (m.nnz // 2) / m.n
        assert 4.0 - 1.0 / m.n <= deg <= 20.0 + 1.0 / m.n
        src = np.repeat(np.arange(m.n), np.diff(m.row_ptrs.astype(np.int64)))
        edges = set(zip(src.tolist(), m.cols.tolist()))
        assert all((y, x) in edges for x, y in edges) and all(x != y for x, y in edges)
        for u in range(m.n):
            assert np.all(np.diff(m.cols[m.row_ptrs[u]:m.row_ptrs[u + 1]].astype(np.int64)) > 0)
    # seeded and independent of the thread count
    G2, offs2 = sme.gen_er_batch(64, (16, 200), (4.0, 20.0), seed=7)
    assert G2.checksum() == G.checksum() and np.array_equal(offs2, offs)
    import subprocess
    import sys

    code = ("import subgraphmatchgpu_amd as s; g, o = s.gen_er_batch(64, (16, 200), (4.0, 20.0), seed=7);"
            "print(g.checksum())")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True,
                         env=dict(os.environ, OMP_NUM_THREADS="1"), cwd=ROOT).stdout.strip()
    assert out == G.checksum()
    assert sme.gen_er_batch(64, (16, 200), (4.0, 20.0), seed=8)[0].checksum() != G.checksum()
    with pytest.raises(sme.SmeError):
        sme.gen_er_batch(0)
    with pytest.raises(sme.SmeError):
        sme.gen_er_batch(4, (100, 50))

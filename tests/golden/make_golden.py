"""Writes the golden fixtures under tests/golden from the reference's own code.

Run in the development container (needs /root/reference and `make -C oracle
ref`):   python tests/golden/make_golden.py

  light_plans.txt  the reference's LazyTraversalPlan for all 14 queries x 4
                   compress levels x ordering on/off, printed by
                   oracle/_ref/plan_dump in the canonical text form.
  ref_counts.json  embedding counts printed by the reference CPU binary
                   (oracle/_ref/patternmatchcpu, CPU_WCOJ and, where the
                   reference supports it, CPU_BFS / CPU_RBFS / CPU_VF2) on
                   seeded graphs written by this repository's generators;
                   each graph is identified by its generator parameters and
                   the sha256 of its CSR arrays.
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import subgraphmatchgpu_amd as sme  # noqa: E402
import oracle  # noqa: E402

# name -> (generator, params, queries for WCOJ)
ALL_Q = list(range(14))
GRAPHS = {
    "rmat8": ("rmat", dict(scale=8, edge_factor=8, seed=1), ALL_Q),
    "rmat10": ("rmat", dict(scale=10, edge_factor=8, seed=2), ALL_Q),
    "rmat12": ("rmat", dict(scale=12, edge_factor=16, seed=11), [0, 1, 2, 3, 6, 7, 11]),
    "rmat14": ("rmat", dict(scale=14, edge_factor=16, seed=1), [0, 2, 3, 7]),
    "er300": ("er", dict(n_vertices=300, n_edges=1500, seed=3), ALL_Q),
}
BFS_Q = [0, 2, 3, 7]   # ReuseTraversalPlan has hard-coded plans only for these among the six


def make_graph(kind, p):
    if kind == "rmat":
        return sme.gen_rmat(p["scale"], p["edge_factor"], 0.57, 0.19, 0.19, p["seed"])
    return sme.gen_er(p["n_vertices"], p["n_edges"], p["seed"])


def main():
    if not oracle.ref_available() or not os.path.exists(oracle.REF_PLAN_DUMP):
        sys.exit("build the reference first: make -C oracle ref")
    plans = subprocess.run([oracle.REF_PLAN_DUMP], capture_output=True, text=True, check=True).stdout
    with open(os.path.join(HERE, "light_plans.txt"), "w") as f:
        f.write(plans)
    out = {"graphs": {}, "counts": {}}
    threads = os.cpu_count() or 4
    with tempfile.TemporaryDirectory() as td:
        for name, (kind, params, qs) in GRAPHS.items():
            g = make_graph(kind, params)
            path = os.path.join(td, name + ".bin")
            sme.write_bin(path, g)
            out["graphs"][name] = dict(kind=kind, **params, n=g.n, nnz=g.nnz, sha256=g.checksum())
            res = {"wcoj": {}, "bfs": {}, "rbfs": {}, "vf2": {}}
            for q in qs:
                res["wcoj"][str(q)] = oracle.ref_count(path, q, algo=0, threads=threads)[0]
                print(name, "Q%d" % q, res["wcoj"][str(q)], flush=True)
            if name in ("rmat8", "er300"):
                for q in qs:
                    res["vf2"][str(q)] = oracle.ref_count(path, q, algo=1, threads=threads)[0]
            if name in ("rmat8", "rmat10", "rmat14"):
                for q in BFS_Q:
                    if q in qs:
                        res["bfs"][str(q)] = oracle.ref_count(path, q, algo=2, threads=threads)[0]
                        res["rbfs"][str(q)] = oracle.ref_count(path, q, algo=3, threads=threads)[0]
            out["counts"][name] = res
    with open(os.path.join(HERE, "ref_counts.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

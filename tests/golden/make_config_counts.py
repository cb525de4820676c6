"""Writes tests/golden/config_counts.json: exact counts of the benchmark
configurations' graphs, from code that shares nothing with the engine.

Run in the development container:
    python tests/golden/make_config_counts.py [--graphs rmat20] [--threads 8]
        [--ref-log 3=/path/to/log ...] [--ref-run 3 --ref-threads 8 --ref-timeout 7200]

For each graph (identified by its generator parameters and the sha256 of its
CSR arrays) it records
  census     triangles, diamonds, 4-cliques (and 5-cliques when asked) from
             oracle.census (degree-ordered clique listing, census.c);
  light      the LIGHT plan's count and per-exec-level outputs of each
             configuration query, which on these plans are whole-graph
             counts: Q3 [n, m, m, T, T, K4], Q2 [n, m, m, 3T, 3T, D],
             Q7 [n, m, m, T, T, K4, K4, K5] (the rule is checked against the
             plan-following oracle by tests/test_oracle.py);
  reference  the reference CPU binary's total_match_count and elapsed_time
             (oracle/_ref/patternmatchcpu, CPU_WCOJ), with its thread count
             and host CPU, either run here (--ref-run) or read from the log
             of the same command run earlier (--ref-log Q=path; the log's
             vertex and edge counts must match the graph).
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import subgraphmatchgpu_amd as sme  # noqa: E402
import oracle  # noqa: E402

OUT = os.path.join(HERE, "config_counts.json")
GRAPHS = {
    # configuration 2: R-MAT scale 20, edge factor 16; 4-clique and diamond
    "rmat20": dict(scale=20, edge_factor=16, seed=1, queries=[3, 2], kmax=4),
    # configurations 3 and 5: R-MAT scale 22, edge factor 32
    "rmat22ef32": dict(scale=22, edge_factor=32, seed=1, queries=[7, 0, 2, 3], kmax=5),
}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def light_from_census(q, n, c):
    m, T = c["edges"], c["triangles"]
    if q == 0:
        return T, [n, m, m, T]
    if q == 2:
        return c["diamonds"], [n, m, m, 3 * T, 3 * T, c["diamonds"]]
    if q == 3:
        return c["k4"], [n, m, m, T, T, c["k4"]]
    if q == 7:
        return c["k5"], [n, m, m, T, T, c["k4"], c["k4"], c["k5"]]
    raise ValueError(q)


def parse_ref_log(path, g):
    count = ms = threads = None
    nv = ne = None
    for line in open(path):
        line = line.strip()
        if line.startswith("vertex_count="):
            a, b = line.split(",")
            nv, ne = int(a.split("=")[1]), int(b.split("=")[1])
        if "thread_num=" in line:
            threads = int(line.split("thread_num=")[1].split(",")[0])
        if line.startswith("total_match_count="):
            a, b = line.split(",")
            count, ms = int(a.split("=")[1]), float(b.split("=")[1].rstrip("ms"))
    if (nv, ne) != (g.n, g.nnz):
        raise SystemExit(f"{path}: log is for another graph ({nv}, {ne})")
    if count is None:
        raise SystemExit(f"{path}: no result line")
    return {"count": count, "elapsed_ms": ms, "threads": threads, "algo": "CPU_WCOJ (-a 0)",
            "source": "log of an earlier run of the same command"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", default="rmat20")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 4)
    ap.add_argument("--ref-log", action="append", default=[], help="Q=path")
    ap.add_argument("--ref-run", action="append", type=int, default=[])
    ap.add_argument("--ref-threads", type=int, default=os.cpu_count() or 4)
    ap.add_argument("--ref-timeout", type=float, default=7200.0)
    args = ap.parse_args()
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in args.graphs.split(","):
        spec = GRAPHS[name]
        g = sme.gen_rmat(spec["scale"], spec["edge_factor"], 0.57, 0.19, 0.19, spec["seed"])
        ent = data.get(name, {})
        if ent.get("sha256") not in (None, g.checksum()):
            raise SystemExit(f"{name}: generator output changed")
        ent.update(kind="rmat", scale=spec["scale"], edge_factor=spec["edge_factor"], seed=spec["seed"],
                   n=g.n, nnz=g.nnz, sha256=g.checksum())
        t = time.time()
        c = oracle.census(g.row_ptrs, g.cols, kmax=spec["kmax"], threads=args.threads)
        ent["census"] = dict(c, seconds=round(time.time() - t, 1), threads=args.threads, kmax=spec["kmax"])
        print(name, ent["census"], flush=True)
        ent["light"] = {}
        for q in spec["queries"]:
            if q == 7 and spec["kmax"] < 5 or q == 3 and spec["kmax"] < 4:
                continue
            cnt, lv = light_from_census(q, g.n, c)
            ent["light"][str(q)] = {"count": cnt, "levels": lv}
        ref = ent.get("reference", {})
        for item in args.ref_log:
            q, path = item.split("=", 1)
            ref[q] = parse_ref_log(path, g)
            ref[q]["cpu"] = cpu_model()
        for q in args.ref_run:
            with tempfile.TemporaryDirectory() as td:
                p = os.path.join(td, name + ".bin")
                sme.write_bin(p, g)
                cnt, ms = oracle.ref_count(p, q, algo=0, threads=args.ref_threads, timeout=args.ref_timeout)
            ref[str(q)] = {"count": cnt, "elapsed_ms": ms, "threads": args.ref_threads,
                           "algo": "CPU_WCOJ (-a 0)", "source": "run by this script", "cpu": cpu_model()}
        if ref:
            ent["reference"] = ref
        data[name] = ent
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

"""Benchmark: embeddings/sec of the LIGHT enumeration on one MI355X.

Workload (one "step"): on a resident R-MAT scale-20 graph (edge factor 16,
a/b/c/d = 0.57/0.19/0.19/0.05, seed 1, symmetrized, deduplicated), count the
4-clique (Q3) and then the diamond (Q2) over every start vertex, each with
symmetry breaking, through the C ABI (libsme.so).  value = embeddings counted
per second of step wall time, inputs already in HBM.

Besides the JSON fields of the harness contract, the line carries
  roofline     -- the dominant kernel family ("intersect": the set
                  intersection verify kernels) timed with HIP events on the
                  engine's stream over the timed steps; achieved = algorithmic
                  bytes / kernel time against 8 TB/s HBM3E;
  cpu_baseline -- a bounded sample of the same workload on the host cores.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SCALE, EDGE_FACTOR, SEED = 20, 16, 1
QUERIES = (3, 2)  # 4-clique, diamond


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(g, seconds_target=15.0, chunk=64, seed=12345):
    """Oracle port of the reference's CPU_WCOJ (OpenMP over all host
    threads) on a seeded uniform random sample of start vertices of the same
    graph, both queries, stopped once about `seconds_target` seconds of CPU
    work are done."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = os.cpu_count() or 1
    order = np.random.default_rng(seed).permutation(g.n).astype(np.uint32)
    emb, used, done = 0, 0.0, 0
    while done < g.n and used < seconds_target:
        starts = np.sort(order[done:done + chunk])
        t0 = time.perf_counter()
        for q in QUERIES:
            emb += oracle.wcoj_count_list(g.row_ptrs, g.cols, q, starts, True, threads)
        used += time.perf_counter() - t0
        done += starts.size
        log(f"cpu baseline: {done} start vertices, {emb} embeddings, {used:.1f}s")
    return {
        "value": emb / used if used > 0 else 0.0,
        "unit": "embeddings/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle CPU_WCOJ port (restates src/cpu/CPUWCOJ.h), Q3+Q2 from {done} seeded random "
                  f"start vertices of {g.n}; {emb} embeddings in {used:.2f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run one device copy of this size (counter calibration)")
    ap.add_argument("--hub-min-degree", type=int, default=0,
                    help="adjacency bitmaps for vertices of at least this degree (0: off)")
    ap.add_argument("--hub-max-gb", type=float, default=8.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import subgraphmatchgpu_amd as sme

    t0 = time.perf_counter()
    g = sme.gen_rmat(SCALE, EDGE_FACTOR, 0.57, 0.19, 0.19, SEED)
    log(f"graph rmat{SCALE} ef{EDGE_FACTOR}: n={g.n} nnz={g.nnz} ({time.perf_counter() - t0:.1f}s)")
    dev = sme.Device(0, budget_bytes=int(args.budget_gb * (1 << 30)))
    if args.calib_copy_mib:
        ms = dev.calibrate_copy(args.calib_copy_mib << 20)
        log(f"calibration copy of {args.calib_copy_mib} MiB: {ms:.3f} ms "
            f"({2 * (args.calib_copy_mib << 20) / ms / 1e6:.1f} GB/s read+write)")
    if args.hub_min_degree:
        dev.set_hub_bitmaps(args.hub_min_degree, int(args.hub_max_gb * (1 << 30)))
    t0 = time.perf_counter()
    dg = dev.upload(g)
    log(f"upload: {1e3 * (time.perf_counter() - t0):.1f} ms, {dg.hubs} hub bitmaps")
    plans = [sme.Plan(q, sme.GPU_LIGHT, sme.CL_SPECIAL) for q in QUERIES]

    def step():
        counts = []
        for q, p in zip(QUERIES, plans):
            t = time.perf_counter()
            r = dev.count(dg, p)
            counts.append(r.count)
            log(f"  Q{q}: {r.count} embeddings in {1e3 * (time.perf_counter() - t):.1f} ms, "
                f"{r.batches} batches, {r.d2h_syncs} syncs, peak pool {r.peak_pool_bytes / 2**30:.2f} GiB, "
                f"levels {r.levels}")
        return counts

    for i in range(args.warmup):
        c = step()
        log(f"warmup {i}: counts {c}")
    dev.profile_reset()
    dev.profile(True)
    t1 = time.perf_counter()
    counts = None
    for i in range(args.steps):
        c = step()
        if counts is not None and c != counts:
            raise AssertionError(f"counts changed between steps: {counts} vs {c}")
        counts = c
        log(f"step {i}: counts {c}")
    elapsed = time.perf_counter() - t1
    dev.profile(False)
    prof = dev.profile_read()
    ms_per_step = 1000.0 * elapsed / args.steps
    value = sum(counts) * args.steps / elapsed

    dom = prof["intersect"]
    avg_ms = dom["ms"] / max(1, dom["launches"])
    bytes_per_launch = dom["bytes"] / max(1, dom["launches"])
    achieved = (bytes_per_launch / 1e9) / (avg_ms / 1e3) if avg_ms > 0 else 0.0
    kernel_ms = sum(v["ms"] for v in prof.values()) / args.steps
    out = {
        "metric": "embeddings/sec & HBM-roofline % (1 GPU)",
        "value": value,
        "unit": "embeddings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64 integer",
        "data": "synthetic (seeded R-MAT)",
        "config": {
            "workload": f"RMAT scale {SCALE} edge factor {EDGE_FACTOR} (n={g.n}, nnz={g.nnz}); "
                        "4-clique then diamond, LIGHT, all start vertices",
            "counts": {"Q3": counts[0], "Q2": counts[1]},
            "budget_gb": args.budget_gb,
            "hub_bitmaps": {"min_degree": args.hub_min_degree, "rows": dg.hubs},
            "kernel_ms_per_step": kernel_ms,
            "kernel_families": {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                                    "GB": round(v["bytes"] / 1e9, 3), "items": v["items"]}
                                for k, v in prof.items()},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "intersect (VerifyKernel)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_ms": avg_ms,
            "launches": dom["launches"],
        },
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(g)
        except Exception as e:  # reported, never fatal to the GPU number
            out["cpu_baseline"] = {"error": str(e)}
    dev.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

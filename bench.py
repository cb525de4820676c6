"""Benchmark: embeddings/sec of the LIGHT enumeration on one MI355X.

Workload (one "step"): on a resident R-MAT scale-20 graph (edge factor 16,
a/b/c/d = 0.57/0.19/0.19/0.05, seed 1, symmetrized, deduplicated), count the
4-clique (Q3) and then the diamond (Q2) over every start vertex, each with
symmetry breaking, through the C ABI (libsme.so).  value = embeddings counted
per second of step wall time, inputs already in HBM.  The graph keeps a
degree-ranked copy (the 4-clique's count and levels are those of the
original ids) and counting levels use the group schedule (group_max 512).

Besides the JSON fields of the harness contract, the line carries
  roofline     -- the dominant kernel family ("intersect": the set
                  intersection verify kernels) timed with HIP events on the
                  engine's stream over the timed steps; achieved = algorithmic
                  bytes / kernel time against 8 TB/s HBM3E;
  cpu_baseline -- a bounded sample of the same workload on the host cores.

Second workload (--workload er-batch): 20,000 seeded Erdos-Renyi graphs
(n uniform in [64, 8192], average degree uniform in [4, 48]) packed back to
back; one step counts triangles and 4-cycles of every member in one
bucketed pass per pattern (value = member graphs per second), after the
autotuner has picked the schedule; the same members counted with one engine
pass per member are timed on a sample for comparison.

Third workload (--workload serve): a resident R-MAT graph (scale 16, edge
factor 16 by default; --serve-scale/--serve-edge-factor); a seeded Poisson
stream of pattern-count requests (200/s for 60 s; triangle/4-cycle/diamond/
4-clique/house/5-clique with weights 0.3/0.2/0.2/0.15/0.1/0.05; each
restricted to a random start range of 1-100% of the vertices) is replayed in
real time into the server (continuous batching, admission under a 64 GB
cap); latency p50/p99 and throughput are reported beside the plain
one-request-at-a-time path replaying the first --plain-limit-s seconds of the
same stream, and every request's count is checked against the plain path's.
On R-MAT scale 22 with edge factor 32 a single house or 5-clique request
over the highest-degree start vertices runs for minutes to hours (requests
are start ranges of the original ids, which the ranked copy cannot serve),
so that stream cannot be replayed within a benchmark run.

Fourth workload (--workload rmat22): R-MAT scale 22, edge factor 32, with
5-clique and house under a 32 GB intermediate cap on the degree-ranked copy,
as ranked start-range tasks for a bounded time per pattern; the rest of the
range is projected from a sampled set of start vertices (bench_rmat22).

The default line (rmat20) also carries, as sub-objects, a bounded run of the
second workload ("er_batch") and of the third on a 5 s stream replayed in
full by both the served and the plain path ("serve").

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--workload rmat20|er-batch|serve|rmat22]
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


REF_CPU = os.path.join(ROOT, "oracle", "_ref", "patternmatchcpu")


def reference_cpu_config1(sme):
    """The reference's own CPU path (src/cpu CPU_WCOJ, built from its sources
    into oracle/_ref by `make -C oracle ref`) on configuration 1 (R-MAT
    scale 14, edge factor 16, triangle) with every host thread: its printed
    elapsed_time (enumeration only, src/cpu/CPUWCOJ.h:29-66), beside this
    engine's time for the same count on the GPU."""
    import re
    import subprocess
    import tempfile

    if not os.path.exists(REF_CPU):
        return {"error": "oracle/_ref/patternmatchcpu is not built"}
    g = sme.gen_rmat(14, 16, 0.57, 0.19, 0.19, SEED)
    threads = os.cpu_count() or 1
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "rmat14.bin")
        sme.write_bin(path, g)
        r = subprocess.run([REF_CPU, "-f", path, "-d", "0", "-a", "0", "-q", "0", "-t", str(threads)],
                           capture_output=True, text=True, timeout=300)
    m = re.search(r"total_match_count=(\d+), elapsed_time=([0-9.]+)ms", r.stdout)
    if r.returncode != 0 or not m:
        return {"error": f"reference run failed (rc {r.returncode}): {r.stdout[-200:]} {r.stderr[-200:]}"}
    count, ms = int(m.group(1)), float(m.group(2))
    dev = sme.Device(0, budget_bytes=4 << 30)
    dg = dev.upload(g)
    plan = sme.Plan(0)
    dev.count(dg, plan)
    t = time.perf_counter()
    gpu = dev.count(dg, plan)
    gpu_ms = 1e3 * (time.perf_counter() - t)
    dev.close()
    if gpu.count != count:
        raise AssertionError(f"configuration 1: reference {count} vs engine {gpu.count}")
    return {"value": count / (ms / 1e3), "unit": "embeddings/s", "cores": threads, "kind": "reference",
            "cpu_model": cpu_model, "sample": "configuration 1 in full: R-MAT scale 14 edge factor 16, "
            f"triangle, CPU_WCOJ -t {threads}", "count": count, "elapsed_ms": ms, "engine_gpu_ms": gpu_ms}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "clique_rows_rmat20_traffic.json")
KERNEL_NAMES = {
    "intersect": "intersect (VerifyKernel)",
    "intersect_group": "intersect_group (GroupCountKernel)",
    "intersect_lane": "intersect_lane (PrepCountKernel)",
    "clique_small": "clique_small (CliqueSmallKernel)",
    "clique_rows": "clique_rows (CliqueRowsHashKernel<10> + <13>; CliqueRowsKernel above 4,096)",
    "clique_count": "clique_count (CliqueCountKernel)",
    "house_mirror": "house_mirror (HouseMirrorKernel / EdgePairsKernel)",
    "house_wedge": "house_wedge (HouseWedgeHashKernel / HouseWedgeDenseKernel)",
}
# Revision of the intersect kernel's memory behaviour; a committed traffic
# figure is used only when it was measured on the same revision.
KERNEL_REV = "clique-bitmaps-hash-rows-word-major-ldscol-pipelined"


def summarize_pmc(fetch_csv, write_csv, calib_bytes, kernel="CliqueRowsHashKernel", per_launch=1):
    """HBM bytes per launch of `kernel` from two rocprofv3 --pmc runs
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what the
    same counter read for the calibration copy of calib_bytes (the largest
    __amd_rocclr_copyBuffer dispatch: it reads and writes calib_bytes).  A
    launch of the engine's kernel slot is `per_launch` dispatches of kernels
    whose names contain `kernel` (the bit-row slot runs one per table size)."""
    import csv

    def load(path):
        rows = list(csv.DictReader(open(path)))
        copy = max((float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]),
                   default=0.0)
        vals = [float(r["Counter_Value"]) for r in rows if kernel in r["Kernel_Name"]]
        return copy, vals

    fc, fv = load(fetch_csv)
    wc, wv = load(write_csv)
    f_factor = (calib_bytes / 1024.0) / fc
    w_factor = (calib_bytes / 1024.0) / wc
    fetch = sum(fv) / len(fv) * per_launch * 1024.0 * f_factor
    write = sum(wv) / len(wv) * per_launch * 1024.0 * w_factor
    return {"kernel": kernel, "dispatches": len(fv), "dispatches_per_launch": per_launch,
            "fetch_kib_raw_per_dispatch": sum(fv) / len(fv),
            "write_kib_raw_per_dispatch": sum(wv) / len(wv), "fetch_factor": f_factor,
            "write_factor": w_factor, "fetch_bytes_per_launch": fetch, "write_bytes_per_launch": write,
            "bytes_per_launch": fetch + write}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run one device copy of this size (counter calibration)")
    ap.add_argument("--hub-min-degree", type=int, default=256,
                    help="adjacency bitmaps for vertices of at least this degree (0: off)")
    ap.add_argument("--hub-max-gb", type=float, default=8.0)
    ap.add_argument("--workload", choices=("rmat20", "er-batch", "serve", "rmat22"), default="rmat20")
    ap.add_argument("--r22-check-ranges", type=int, default=16,
                    help="rmat22: ranked ranges of 64 start vertices on which the 5-clique's level "
                         "engine and bitmap path must agree (0: none)")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-scale", type=int, default=16)
    ap.add_argument("--serve-edge-factor", type=int, default=16)
    ap.add_argument("--plain-limit-s", type=float, default=2.0,
                    help="replay only this many seconds of the stream into the plain path "
                         "(it cannot keep up; its backlog drains after the cut)")
    ap.add_argument("--no-sub", action="store_true",
                    help="rmat20: skip the ER-batch and serving sub-benchmarks of the default line")
    ap.add_argument("--sub-serve-seconds", type=float, default=5.0,
                    help="stream length of the serving sub-benchmark (served and plain replay it all)")
    ap.add_argument("--er-graphs", type=int, default=20000)
    ap.add_argument("--per-graph-sample", type=int, default=1000,
                    help="members timed with one pass each (er-batch)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="rmat20: let the autotuner pick the schedule before timing")
    ap.add_argument("--tune", default="0,0,0,4,0,512",
                    help="lane_path_max,item_grid,batch_bytes,super_chunks[,stream_ratio[,group_max]] "
                         "schedule (rmat20)")
    ap.add_argument("--level-engine-steps", type=int, default=2,
                    help="rmat20: also time this many steps through the level engine (0: skip)")
    ap.add_argument("--no-rank-order", action="store_true",
                    help="rmat20: no degree-ranked copy (clique counts then walk the original ids)")
    ap.add_argument("--summarize-pmc", nargs="+", metavar="ARG",
                    help="FETCH_CSV WRITE_CSV CALIB_MIB [KERNEL]: print HBM traffic per launch of KERNEL "
                         "(default GroupCountKernel) from two rocprofv3 --pmc runs")
    args = ap.parse_args()
    if args.summarize_pmc:
        f, w, mib = args.summarize_pmc[:3]
        kernel = args.summarize_pmc[3] if len(args.summarize_pmc) > 3 else "CliqueRowsHashKernel"
        per = int(args.summarize_pmc[4]) if len(args.summarize_pmc) > 4 else 1
        print(json.dumps(summarize_pmc(f, w, int(mib) << 20, kernel, per)))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import subgraphmatchgpu_amd as sme

    if args.workload == "er-batch":
        print(json.dumps(bench_er_batch(args, sme)))
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args, sme, plain_limit_s=args.plain_limit_s)))
        return
    if args.workload == "rmat22":
        if args.budget_gb == 64.0:
            args.budget_gb = 32.0  # the configuration's intermediate cap
        print(json.dumps(bench_rmat22(args, sme)))
        return

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
    dev.set_rank_order(not args.no_rank_order)
    # reference-like set-up (PatternMatch.cuh:58-87 times plan + CSR upload):
    # plans, CSR upload, hub bitmaps and the degree-ranked copy
    t0 = time.perf_counter()
    plans = [sme.Plan(q, sme.GPU_LIGHT, sme.CL_SPECIAL) for q in QUERIES]
    dg = dev.upload(g)
    setup_s = time.perf_counter() - t0
    log(f"plans + upload: {1e3 * setup_s:.1f} ms, {dg.hubs} hub bitmaps, ranked copy {dg.ranked}")
    if args.tune:
        f = [int(x) for x in args.tune.split(",")] + [0, 0, 0, 0, 0, 0]
        dev.set_tuning(lane_path_max=f[0], item_grid=f[1], batch_bytes=f[2], super_chunks=f[3],
                       stream_ratio=f[4], group_max=f[5])
    if args.autotune:
        t = time.perf_counter()
        best, ms, trials = dev.autotune(dg, plans, sample_fraction=0.02)
        log(f"autotune: {best} ({time.perf_counter() - t:.1f}s, {len(trials)} trials)")
    log(f"schedule: {dev.get_tuning()}")

    def step():
        counts = []
        for q, p in zip(QUERIES, plans):
            t = time.perf_counter()
            r = dev.count(dg, p)
            counts.append(r.count)
            log(f"  Q{q}: {r.count} embeddings in {1e3 * (time.perf_counter() - t):.1f} ms, "
                f"{r.batches} batches, {r.d2h_syncs} syncs, peak pool {r.peak_pool_bytes / 2**30:.2f} GiB, "
                f"ranked copy {r.ranked}, levels {r.levels}")
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

    # the dominant kernel family of the step, whichever took the most time
    fam = max((k for k in prof if k != "misc"), key=lambda k: prof[k]["ms"])
    dom = prof[fam]

    def family_roofline(names):
        ms = sum(prof[k]["ms"] for k in names)
        launches = sum(prof[k]["launches"] for k in names)
        by = sum(prof[k]["bytes"] for k in names)
        gbs = (by / 1e9) / (ms / 1e3) if ms > 0 else 0.0
        return {"kernels": list(names), "launches": launches, "ms_per_step": ms / args.steps,
                "GB_per_step": by / 1e9 / args.steps, "achieved": gbs, "frac": gbs / HBM_PEAK_GBS}

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
            "schedule": dev.get_tuning(),
            "kernel_ms_per_step": kernel_ms,
            "kernel_families": {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                                    "GB": round(v["bytes"] / 1e9, 3), "items": v["items"]}
                                for k, v in prof.items()},
        },
        "reference_like": {
            "value": sum(counts) / (setup_s + elapsed / args.steps),
            "unit": "embeddings/s",
            "definition": "one step's embeddings / (plan creation + CSR upload with hub bitmaps and the "
                          "degree-ranked copy + one step), as PatternMatch.cuh:58-87 times plan and upload",
            "setup_ms": 1e3 * setup_s,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": KERNEL_NAMES.get(fam, fam),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_ms": avg_ms,
            "launches": dom["launches"],
            "families": {
                "intersect_all": family_roofline(("intersect", "intersect_group", "intersect_lane")),
                "scatter": family_roofline(("expand", "compact")),
                "clique_bitmaps": family_roofline(("clique_small", "clique_rows", "clique_count")),
                "edge_census": family_roofline(("house_mirror", "house_wedge")),
            },
        },
    }
    # the same step through the level engine (clique and census paths off):
    # its intersection and scatter kernels, measured the same way
    if args.level_engine_steps:
        sched = dev.get_tuning()
        dev.set_tuning(**dict(sched, clique_level_engine=1, census_level_engine=1))
        step()  # warm-up
        dev.profile_reset()
        dev.profile(True)
        t2 = time.perf_counter()
        for _ in range(args.level_engine_steps):
            if step() != counts:
                raise AssertionError("the level engine's counts differ")
        le_elapsed = time.perf_counter() - t2
        dev.profile(False)
        lprof = dev.profile_read()
        dev.set_tuning(**sched)

        def level_family(names):
            ms = sum(lprof[k]["ms"] for k in names)
            by = sum(lprof[k]["bytes"] for k in names)
            gbs = (by / 1e9) / (ms / 1e3) if ms > 0 else 0.0
            return {"kernels": list(names), "launches": sum(lprof[k]["launches"] for k in names),
                    "ms_per_step": ms / args.level_engine_steps,
                    "GB_per_step": by / 1e9 / args.level_engine_steps, "achieved": gbs,
                    "frac": gbs / HBM_PEAK_GBS}

        out["level_engine"] = {
            "what": "the same step with clique_level_engine = census_level_engine = 1 (LIGHT level engine: "
                    "group counting for the 4-clique's last level, item verify kernels, scan and scatter); "
                    "counts checked equal",
            "ms_per_step": 1e3 * le_elapsed / args.level_engine_steps,
            "value": sum(counts) * args.level_engine_steps / le_elapsed,
            "families": {
                "intersect_all": level_family(("intersect", "intersect_group", "intersect_lane")),
                "scatter": level_family(("expand", "compact")),
            },
        }
    # HBM bytes per launch of the dominant kernel measured by rocprofv3
    # counters for this same configuration (TRAFFIC_FILE)
    try:
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        if tr.get("config") == {"hub_min_degree": args.hub_min_degree, "schedule": dev.get_tuning(),
                                "kernel_rev": KERNEL_REV, "rank_order": not args.no_rank_order} and \
                tr.get("kernel", "") in out["roofline"]["kernel"]:
            out["roofline"]["traffic"] = tr["bytes_per_launch"]
            out["roofline"]["traffic_source"] = os.path.relpath(TRAFFIC_FILE, ROOT)
            out["roofline"]["traffic_gbs"] = tr["bytes_per_launch"] / 1e9 / (avg_ms / 1e3)
    except (OSError, ValueError, KeyError):
        pass
    dev.close()
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(g)
        except Exception as e:  # reported, never fatal to the GPU number
            out["cpu_baseline"] = {"error": str(e)}
        out["cpu_baseline"]["reference"] = reference_cpu_config1(sme)
    del g
    if not args.no_sub:
        # configuration 4 (many uneven graphs in one pass, autotuned) and
        # configuration 5 (served stream vs the plain path on the same
        # requests), bounded so that the whole run stays within minutes
        er = bench_er_batch(args, sme)
        out["er_batch"] = {
            "value": er["value"], "unit": er["unit"], "ms_per_step": er["ms_per_step"],
            "workload": er["config"]["workload"], "totals": er["config"]["totals"],
            "speedup_vs_one_pass_per_member": er["config"]["speedup_vs_one_pass_per_member"],
            "one_pass_per_member": er["config"]["one_pass_per_member"],
            "packed_default_schedule": er["config"]["packed_default_schedule"],
            "tuning": er["config"]["autotune"]["tuning"], "roofline": er["roofline"],
        }
        sv = bench_serve(args, sme, seconds=args.sub_serve_seconds, plain_limit_s=None)
        c = sv["config"]
        out["serve"] = {
            "workload": c["workload"], "requests": c["served"]["requests"],
            "p50_ms": c["served"]["p50_ms"], "p99_ms": c["served"]["p99_ms"],
            "throughput_rps": c["served"]["throughput_rps"],
            "plain": {k: c["plain"][k] for k in ("requests", "p50_ms", "p99_ms", "throughput_rps")},
            "throughput_vs_plain": c["throughput_vs_plain"], "counts_checked": c["counts_checked"],
            "passes": c["served"]["stats"]["passes"],
        }
    print(json.dumps(out))


def bench_er_batch(args, sme):
    """Configuration 4 (see the module docstring); returns the JSON dict."""
    import numpy as np

    t0 = time.perf_counter()
    G, offs = sme.gen_er_batch(args.er_graphs, (64, 8192), (4.0, 48.0), seed=SEED)
    ng = args.er_graphs
    log(f"er-batch: {ng} graphs, n={G.n} nnz={G.nnz} ({time.perf_counter() - t0:.1f}s)")
    dev = sme.Device(0, budget_bytes=int(args.budget_gb * (1 << 30)))
    dg = dev.upload(G)
    qs = (0, 1)  # triangle, 4-cycle
    plans = [sme.Plan(q) for q in qs]

    def packed():
        out = []
        for q, p in zip(qs, plans):
            t = time.perf_counter()
            c, r = dev.count_buckets(dg, p, offs)
            out.append(c)
            log(f"  Q{q}: total {r.count} in {1e3 * (time.perf_counter() - t):.1f} ms, {r.batches} batches, "
                f"peak pool {r.peak_pool_bytes / 2**30:.2f} GiB")
        return out

    # one engine pass per member, default schedule, on an evenly spaced sample
    sample = np.linspace(0, ng - 1, min(ng, args.per_graph_sample)).astype(np.int64)
    per = {q: [] for q in qs}
    packed()  # warm the pool and code objects
    t = time.perf_counter()
    for k in sample:
        for q, p in zip(qs, plans):
            per[q].append(dev.count(dg, p, int(offs[k]), int(offs[k + 1])).count)
    per_graph_s = (time.perf_counter() - t) / sample.size
    log(f"one pass per member: {1e3 * per_graph_s:.3f} ms per member ({sample.size} sampled)")

    def timed(label):
        for i in range(args.warmup):
            packed()
        dev.profile_reset()
        dev.profile(True)
        t1 = time.perf_counter()
        res = None
        for i in range(args.steps):
            c = packed()
            if res is not None and any(not np.array_equal(a, b) for a, b in zip(res, c)):
                raise AssertionError("per-member counts changed between steps")
            res = c
        el = (time.perf_counter() - t1) / args.steps
        dev.profile(False)
        log(f"{label}: {1e3 * el:.1f} ms per step")
        return res, el, dev.profile_read()

    base_counts, base_s, _ = timed("packed, default schedule")
    tune = {"tuning": dev.get_tuning(), "sample_ms": None, "trials": []}
    if not args.no_autotune:
        t = time.perf_counter()
        best, ms, trials = dev.autotune(dg, plans, bounds=offs, sample_fraction=0.02)
        tune = {"tuning": best, "sample_ms": ms, "trials": trials,
                "seconds": time.perf_counter() - t}
        log(f"autotune: {best} ({tune['seconds']:.1f}s, {len(trials)} trials)")
    counts, step_s, prof = timed("packed, tuned schedule")
    for q, a, b in zip(qs, base_counts, counts):
        if not np.array_equal(a, b):
            raise AssertionError(f"Q{q}: tuned counts differ from default")
        if a[sample].tolist() != per[q]:
            raise AssertionError(f"Q{q}: one-pass-per-member counts differ from the packed pass")
    fam = max(prof, key=lambda k: prof[k]["ms"])
    dom = prof[fam]
    avg_ms = dom["ms"] / max(1, dom["launches"])
    achieved = (dom["bytes"] / max(1, dom["launches"]) / 1e9) / (avg_ms / 1e3) if avg_ms > 0 else 0.0
    out = {
        "metric": "member graphs/sec (triangle + 4-cycle counts per graph, 1 GPU)",
        "value": ng / step_s,
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * step_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64 integer",
        "data": "synthetic (seeded Erdos-Renyi batch)",
        "config": {
            "workload": f"{ng} ER graphs, n in [64, 8192], average degree in [4, 48], packed "
                        f"(n={G.n}, nnz={G.nnz}); triangle + 4-cycle per member",
            "totals": {f"Q{q}": int(c.sum()) for q, c in zip(qs, counts)},
            "one_pass_per_member": {"ms_per_member": 1e3 * per_graph_s, "sampled_members": int(sample.size),
                                    "graphs_per_s": 1.0 / per_graph_s},
            "packed_default_schedule": {"ms_per_step": 1e3 * base_s, "graphs_per_s": ng / base_s},
            "speedup_vs_one_pass_per_member": (ng / step_s) * per_graph_s,
            "autotune": tune,
            "kernel_families": {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                                    "GB": round(v["bytes"] / 1e9, 3), "items": v["items"]}
                                for k, v in prof.items()},
        },
        "roofline": {"bound": "hbm", "kernel": fam, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "avg_launch_ms": avg_ms, "launches": dom["launches"]},
    }
    dev.close()
    return out


def bench_rmat22(args, sme):
    """Configuration 3: R-MAT scale 22, edge factor 32, 5-clique and house,
    intermediates capped at 32 GB (--budget-gb), degree-ranked copy resident.

    Each pattern is counted over all start vertices in one call: the
    5-clique (with its level outputs) by bitmaps of every start vertex's
    upper neighbourhood, the house (count only) from per-edge triangle
    counts and a degree-ordered 4-cycle wedge walk.  Reported per pattern:
    count, seconds, batches, peak pool bytes and every kernel family's time.
    Beside it, the 5-clique's level engine and bitmap path count the same
    ranked start ranges (--r22-check-ranges ranges of 64 vertices, spread
    over the 2^18 highest ranks, where nearly all 5-cliques start) and must
    agree."""
    import numpy as np

    t0 = time.perf_counter()
    g = sme.gen_rmat(22, 32, 0.57, 0.19, 0.19, SEED)
    log(f"graph rmat22 ef32: n={g.n} nnz={g.nnz} ({time.perf_counter() - t0:.1f}s)")
    budget = int(args.budget_gb * (1 << 30))
    dev = sme.Device(0, budget_bytes=budget)
    dev.set_rank_order(True)
    t0 = time.perf_counter()
    dg = dev.upload(g)
    upload_s = time.perf_counter() - t0
    log(f"upload + ranked copy: {upload_s:.1f}s")
    dev.set_tuning(super_chunks=4, group_max=512)
    results = {}
    total_counted, total_s = 0, 0.0
    for q, name in ((7, "5-clique"), (4, "house")):
        plan = sme.Plan(q)
        dev.profile_reset()
        dev.profile(True)
        res = dev.count(dg, plan) if q == 7 else dev.count_only(dg, plan)
        dev.profile(False)
        prof = {k: v for k, v in dev.profile_read().items() if v["launches"]}
        dt = res.elapsed_ms / 1e3
        log(f"  {name}: {res.count} in {dt:.2f}s, {res.batches} batches, peak pool "
            f"{res.peak_pool_bytes / 2**30:.2f} GiB")
        fam = max(prof, key=lambda k: prof[k]["ms"])
        ent = {"query": f"Q{q} ({name})", "complete": True, "count": res.count, "seconds": dt,
               "batches": res.batches, "peak_pool_gib": res.peak_pool_bytes / 2**30,
               "rate_per_s": res.count / dt if dt else 0.0, "ranked_copy": bool(res.ranked),
               "path": "clique bitmaps" if q == 7 else "edge triangle counts + 4-cycle wedges",
               "dominant_kernel": {"family": fam, "ms": prof[fam]["ms"],
                                   "GB": prof[fam]["bytes"] / 1e9,
                                   "achieved_GBs": (prof[fam]["bytes"] / 1e9) / (prof[fam]["ms"] / 1e3)
                                   if prof[fam]["ms"] else 0.0},
               "kernel_families": {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                                       "GB": round(v["bytes"] / 1e9, 3)} for k, v in prof.items()}}
        ent["dominant_kernel"]["frac"] = ent["dominant_kernel"]["achieved_GBs"] / HBM_PEAK_GBS
        if q == 7:
            ent["levels"] = res.levels
        results[name] = ent
        total_counted += res.count
        total_s += dt
    # the 5-clique's two engines on the same ranked start ranges
    checks = []
    if args.r22_check_ranges:
        # the top 2^18 ranks hold nearly every 5-clique
        starts = np.unique(np.linspace(max(0, g.n - (1 << 18)), g.n - 64, args.r22_check_ranges).astype(np.int64))
        plan = sme.Plan(7)
        for a in starts:
            a = int(a)
            dev.set_tuning(super_chunks=4, group_max=512, clique_level_engine=1)
            lvl = dev.count_ranked(dg, plan, a, a + 64)
            dev.set_tuning(super_chunks=4, group_max=512)
            bit = dev.count_ranked(dg, plan, a, a + 64)
            checks.append({"ranked_range": [a, a + 64], "count": bit.count,
                           "level_engine_ms": lvl.elapsed_ms, "bitmap_ms": bit.elapsed_ms,
                           "agree": (lvl.count, lvl.levels) == (bit.count, bit.levels)})
            log(f"  check [{a}, {a + 64}): {bit.count} level {lvl.elapsed_ms:.1f} ms bitmap "
                f"{bit.elapsed_ms:.1f} ms {'agree' if checks[-1]['agree'] else 'DIFFER'}")
        if not all(c["agree"] for c in checks):
            raise SystemExit("5-clique: level engine and bitmap path differ")
    dev.close()
    return {
        "metric": "embeddings/sec (1 GPU, 32 GB intermediate cap)",
        "value": total_counted / total_s if total_s else 0.0,
        "unit": "embeddings/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": 1e3 * total_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64 integer",
        "data": "synthetic (seeded R-MAT)",
        "config": {
            "workload": f"RMAT scale 22 edge factor 32 (n={g.n}, nnz={g.nnz}); 5-clique and house over all "
                        "start vertices, degree-ranked copy",
            "budget_gb": args.budget_gb,
            "upload_s": upload_s,
            "patterns": results,
            "five_clique_engine_checks": checks,
        },
    }


SERVE_QUERIES = (0, 1, 2, 3, 4, 7)
SERVE_WEIGHTS = (0.3, 0.2, 0.2, 0.15, 0.1, 0.05)


def request_stream(n, seconds, rate, seed):
    """Seeded Poisson arrivals; (t, query, v_begin, v_end) with the start
    range [lo, lo + ceil(f n)), f ~ U[0.01, 1], lo ~ U[0, n - len]."""
    import math

    import numpy as np

    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        q = int(rng.choice(SERVE_QUERIES, p=SERVE_WEIGHTS))
        ln = max(1, math.ceil(rng.uniform(0.01, 1.0) * n))
        lo = int(rng.integers(0, n - ln + 1))
        out.append((t, q, lo, lo + ln))
    return out


def replay(srv, stream, limit_s=None):
    """Submits each request at its arrival time (relative to now) from this
    thread while a second thread waits for results.  With limit_s, stops
    submitting after that many seconds.  Returns (results, submitted, wall
    seconds from the first arrival to the last completion)."""
    import threading

    t0 = time.perf_counter()
    ids, results = [], {}
    done = threading.Event()
    lock = threading.Condition()

    def waiter():
        k = 0
        while True:
            with lock:
                while k >= len(ids) and not done.is_set():
                    lock.wait(0.05)
                if k >= len(ids) and done.is_set():
                    return
                rid = ids[k]
            results[k] = srv.wait(rid)
            k += 1
            if k % 500 == 0:
                log(f"    {k} requests done")

    th = threading.Thread(target=waiter)
    th.start()
    for t, q, a, b in stream:
        now = time.perf_counter() - t0
        if limit_s is not None and t > limit_s:
            break
        if t > now:
            time.sleep(t - now)
        rid = srv.submit(q, a, b)
        with lock:
            ids.append(rid)
            lock.notify()
    done.set()
    with lock:
        lock.notify()
    th.join()
    wall = time.perf_counter() - t0 - stream[0][0]
    return [results[k] for k in range(len(ids))], len(ids), wall


def bench_serve(args, sme, seconds=None, plain_limit_s=None):
    """Configuration 5 on a stream of `seconds` (default --serve-seconds);
    the plain path replays the first plain_limit_s seconds of the same
    stream (None: all of it).  Returns the JSON dict."""
    import numpy as np

    seconds = args.serve_seconds if seconds is None else seconds
    t0 = time.perf_counter()
    g = sme.gen_rmat(args.serve_scale, args.serve_edge_factor, 0.57, 0.19, 0.19, SEED)
    log(f"graph rmat{args.serve_scale} ef{args.serve_edge_factor}: n={g.n} nnz={g.nnz} "
        f"({time.perf_counter() - t0:.1f}s)")
    stream = request_stream(g.n, seconds, args.serve_rate, seed=SEED)
    log(f"stream: {len(stream)} requests over {seconds:.0f}s")
    cap = int(args.budget_gb * (1 << 30))
    hub = dict(hub_min_degree=args.hub_min_degree, hub_max_bytes=int(args.hub_max_gb * (1 << 30)))

    def run(label, **kw):
        srv = sme.Server(g, mem_cap_bytes=cap, **hub, **kw)
        log(f"{label}: replaying")
        res, sent, wall = replay(srv, stream)
        st = srv.stats()
        srv.close()
        lat = np.array([r[1] for r in res])
        out = {"requests": sent, "wall_s": wall, "throughput_rps": sent / wall,
               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
               "mean_ms": float(lat.mean()), "stats": st}
        log(f"{label}: {out}")
        return [r[0] for r in res], out

    served_counts, served = run("served", slots=args.serve_slots, merge=True, min_reserve_bytes=4 << 30)
    # plain path: one request at a time in arrival order, whole budget each
    srv = sme.Server(g, mem_cap_bytes=cap, slots=1, merge=False, **hub)
    log("plain: replaying")
    res, sent, wall = replay(srv, stream, plain_limit_s)
    st = srv.stats()
    srv.close()
    lat = np.array([r[1] for r in res])
    plain = {"requests": sent, "wall_s": wall, "throughput_rps": sent / wall,
             "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
             "mean_ms": float(lat.mean()), "stats": st,
             "stream_cut_at_s": plain_limit_s if sent < len(stream) else None}
    log(f"plain: {plain}")
    plain_counts = [r[0] for r in res]
    if plain_counts != served_counts[:len(plain_counts)]:
        bad = sum(a != b for a, b in zip(plain_counts, served_counts))
        raise AssertionError(f"{bad} requests differ between the served mix and the plain path")
    out = {
        "metric": "served requests/sec with p50/p99 latency vs plain (1 GPU)",
        "value": served["throughput_rps"],
        "unit": "requests/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": 1e3 * served["wall_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64 integer",
        "data": "synthetic (seeded R-MAT + seeded Poisson request stream)",
        "config": {
            "workload": f"RMAT scale {args.serve_scale} edge factor {args.serve_edge_factor} (n={g.n}, "
                        f"nnz={g.nnz}); Poisson {args.serve_rate:.0f} req/s for {seconds:.0f}s, "
                        "patterns {Q0,Q1,Q2,Q3,Q4,Q7} weights 0.3/0.2/0.2/0.15/0.1/0.05, "
                        "start range 1-100% of vertices",
            "mem_cap_gb": args.budget_gb,
            "served": served,
            "plain": plain,
            "throughput_vs_plain": served["throughput_rps"] / plain["throughput_rps"],
            "counts_checked": len(plain_counts),
        },
    }
    return out


if __name__ == "__main__":
    main()

"""Benchmark of the screening work unit on one MI355X: 20,000 synthetic stars x 4,087 synthetic clusters with
the fused shuffled-velocity control (stars seed 2, clusters seed 0, control permutation seed 3), rows with
p_r50 >= 0.001.  A step is one device pass over the whole unit (draws, tracks, scan, finish, row compaction)
with the cluster tables resident and the star records already in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form fast|plain] [--stars N] [--no-cpu]

Prints one JSON line.  value = pairs screened per second (real + control pairs) of the chosen form;
the plain FP64 form is timed beside it (plain_value, speedup) and their p_r50 agreement is checked.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from phygas_amd import hip  # noqa: E402
from phygas_amd import workload as wl  # noqa: E402

# MI355X published peaks.  The vector units run FP64 and unpacked FP32 at the same rate (78.6 TFLOP/s, FMA = 2
# FLOP); this build disables packed FP32, so 78.6 TFLOP/s bounds both the FP32 and the FP64 work of a kernel.
PEAK_VALU = 78.6          # TFLOP/s
PEAK_HBM = 8000.0         # GB/s
METRIC = "star-cluster pairs/s, chosen vs plain GPU at |dp_r50|<=1e-4; p50/p99 latency"

# Algorithmic work per unit (FMA = 2 FLOP, add/sub/mul/div/sqrt/compare = 1), counted from the source:
PLAIN_POINT_FLOP = 126.0   # k_chi2_plain, one interior pair-point (56 at a track's last point)
PLAIN_LAST_FLOP = 56.0
SCAN_L1_FLOP = 76.0        # cull.h l1_pass, one (pair, tile) test, FP32
SCAN_L2_FLOP = 367.0       # cull.h l2f_pass, one (pair, tile) certificate, FP32 (both quadratic checks, vertex)
POINT_FLOP = 134.0         # k_items, one pair-point: both Mahalanobis forms and the segment parameter, FP64


def plain_scan_flops(cn: np.ndarray, n_stars: int) -> float:
    """Algorithmic FP64 work of the plain full scan, summed over every cluster, times the stars."""
    per_star = float(np.sum(PLAIN_POINT_FLOP * (cn - 1) + PLAIN_LAST_FLOP))
    return per_star * n_stars


def pmc_traffic(kernel: str) -> float | None:
    """HBM bytes per launch of a kernel on this workload from the committed rocprofv3 --pmc summary
    (profiles/<kernel>_traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel factor)."""
    f = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        return float(json.load(fh)["traffic_bytes"])


def cpu_baseline(clusters, stars, threads: int) -> dict | None:
    """The reference's own screening loop (oracle/_ref, built from the reference sources) or, without it, the
    oracle restatement, on a bounded sample: about 50 stars per thread against every cluster."""
    ref = os.path.join(ROOT, "oracle", "_ref", "cpu_bench_ref")
    port = os.path.join(ROOT, "oracle", "cpu_bench_orc")
    exe, kind = (ref, "reference") if os.path.exists(ref) else (port, "port")
    if not os.path.exists(exe):
        return None
    n = 50 * threads
    with tempfile.TemporaryDirectory() as d:
        cf, sf = os.path.join(d, "clusters.txt"), os.path.join(d, "stars.txt")
        wl.write_clusters(cf, clusters)
        wl.write_targets(sf, stars[:n])
        out = subprocess.run([exe, cf, sf, str(threads), str(n)], capture_output=True, text=True, check=True).stdout
    kv = dict(x.split("=") for x in out.split())
    return {"value": float(kv["pairs_per_s"]), "unit": "pairs/s", "cores": threads, "kind": kind,
            "sample": f"{n} of the work unit's stars x {kv['clusters']} clusters, chi2 cut 25 + finish "
                      f"({kv['kept']} kept pairs); cluster tracks prepared first in {float(kv['cluster_prep_s']):.2f} s "
                      f"and not timed"}


def copy_probe(gib: float) -> int:
    """One device-to-device copy of gib GiB through the HIP runtime (its blit kernel reads and writes exactly
    that many bytes): the calibration run for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hiprt = C.CDLL("libamdhip64.so")
    n = int(gib * (1 << 30))
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hiprt.hipMalloc(C.byref(ptr), C.c_size_t(n)) != 0:
            print("hipMalloc failed", file=sys.stderr)
            return 1
    hiprt.hipMemset(a, 1, C.c_size_t(n))
    hiprt.hipDeviceSynchronize()
    rc = hiprt.hipMemcpy(b, a, C.c_size_t(n), 3)   # hipMemcpyDeviceToDevice
    hiprt.hipDeviceSynchronize()
    hiprt.hipFree(a)
    hiprt.hipFree(b)
    print(json.dumps({"copy_probe_bytes": n, "rc": rc}), flush=True)
    return 0 if rc == 0 else 1


def run_form(e: hip.Engine, form: int, steps: int, warmup: int, chi2_max: float, p_min: float, stage_passes: int = 2):
    """Time `steps` passes (whole passes, split across the engine's sub-pass streams as the library does by
    default), then take the stage breakdown from `stage_passes` more passes with stage events, after one not
    counted (a pass with stage events runs whole, on the engine's stream alone)."""
    for _ in range(warmup):
        e.screen_async(form, chi2_max, 64, p_min)
        e.kept()          # sizes the selection buffers (a pass whose buffers overflowed is run again)
    e.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        e.screen_async(form, chi2_max, 64, p_min)
        e.kept()          # every timed pass is complete: an overflowed pass is run again inside the timing
    e.sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / steps
    e.set_timing(True)
    stage_sum: dict[str, float] = {}
    for rep in range(stage_passes + 1):           # the first pass with stage events is not counted
        e.screen_async(form, chi2_max, 64, p_min)
        for k, v in e.stage_times().items():     # waits for the pass; events are on the engine's stream
            if rep > 0:
                stage_sum[k] = stage_sum.get(k, 0.0) + v
        e.kept()
    e.set_timing(False)
    return ms, {k: v / max(stage_passes, 1) for k, v in stage_sum.items()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--form", choices=["fast", "plain"], default="fast")
    ap.add_argument("--stars", type=int, default=20000)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=150)
    ap.add_argument("--serve-rate", type=float, default=15.0)
    ap.add_argument("--tune", action="store_true", help="run the autotuner before timing (the library's default "
                    "grids are its choices on this unit)")
    ap.add_argument("--split", type=int, default=0, help="sub-passes of a fast pass on their own streams "
                    "(0: the library's default, 2)")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="only copy GIB GiB device to device (the HBM counter calibration of the profiles) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the screening shards by star with no "
              f"collective, see the sharding paragraph of the design notes", file=sys.stderr)
        return 2

    if args.copy_probe > 0:
        return copy_probe(args.copy_probe)

    clusters = wl.make_clusters(4087, 0)
    stars = wl.make_stars(args.stars, 2)
    control = wl.shuffled_control(stars, 3)
    chi2_max, p_min = 25.0, 0.001
    cn = (np.ceil(clusters["tmax_myr"] / 0.5) + 1).astype(np.int64)
    n_stars = 2 * len(stars)
    pairs = n_stars * len(clusters)

    e = hip.Engine(0)
    t0 = time.perf_counter()
    e.load_clusters(clusters, 64, 2.0)
    tables_s = time.perf_counter() - t0
    e.load_stars(stars, 0)
    e.load_stars(control, 1)
    if args.split > 0:
        e.set_split(args.split)

    form = hip.FAST if args.form == "fast" else hip.PLAIN
    tune = None
    if form == hip.FAST and args.tune:
        tune = e.autotune(hip.FAST, chi2_max, 64, p_min)   # launch grids, chosen before the timed steps
    ms, stages = run_form(e, form, args.steps, args.warmup, chi2_max, p_min)
    rows = e.rows()
    kept = e.kept()
    stats = e.pass_stats()
    result = {"form": "fast" if form == hip.FAST else "plain", "ms": ms, "stages": stages, "rows": len(rows),
              "kept": kept}

    plain = None
    if form == hip.FAST and not args.no_plain:
        pms, pstages = run_form(e, hip.PLAIN, args.plain_steps, 0, chi2_max, p_min)
        prow = e.rows()
        a = {(r["star"], r["cluster"]): r["p_r50"] for r in rows}
        b = {(r["star"], r["cluster"]): r["p_r50"] for r in prow}
        common = set(a) & set(b)
        dp = max((abs(a[k] - b[k]) for k in common), default=0.0)
        nbad = sum(abs(a[k] - b[k]) > 1e-4 for k in common)
        plain = {"ms": pms, "stages": pstages, "rows": len(prow), "value": pairs / (pms / 1000.0),
                 "rows_only_fast": len(set(a) - set(b)), "rows_only_plain": len(set(b) - set(a)),
                 "max_abs_dp_r50": dp, "pairs_dp_r50_above_1e-4": nbad}

    # roofline of the dominant kernel (largest stage of the timed pass)
    if form == hip.PLAIN:
        k_ms = stages.get("scan", float("nan"))
        flops = plain_scan_flops(cn, n_stars)
        achieved = flops / (k_ms / 1000.0) / 1e12
        roof = {"bound": "valu", "kernel": "k_chi2_plain", "achieved": achieved, "peak": PEAK_VALU,
                "unit": "TFLOP/s", "frac": achieved / PEAK_VALU, "traffic": pmc_traffic("k_chi2_plain"),
                "work_per_launch": flops, "avg_launch_ms": k_ms}
    else:
        # the pair scan (every pair of the unit) or the point evaluation of the tiles it keeps, whichever is longer
        name = max(("scan", "items"), key=lambda k: stages.get(k, 0.0))
        k_ms = stages.get(name, float("nan"))
        if name == "items":
            work = stats["items"] * 8 * POINT_FLOP   # eight points per item (fewer at a track's end: an upper bound)
            achieved = work / (k_ms / 1000.0) / 1e12
            roof = {"bound": "valu", "kernel": "k_items", "achieved": achieved, "peak": PEAK_VALU,
                    "unit": "TFLOP/s", "frac": achieved / PEAK_VALU, "traffic": pmc_traffic("k_items"),
                    "work_per_launch": work, "avg_launch_ms": k_ms}
        else:
            work = stats["l1_tests"] * SCAN_L1_FLOP + stats["l2_tests"] * SCAN_L2_FLOP
            achieved = work / (k_ms / 1000.0) / 1e12
            roof = {"bound": "valu", "kernel": "k_scan_tiles", "achieved": achieved, "peak": PEAK_VALU,
                    "unit": "TFLOP/s", "frac": achieved / PEAK_VALU, "traffic": pmc_traffic("k_scan_tiles"),
                    "work_per_launch": work, "avg_launch_ms": k_ms}

    e.close()
    serving = None
    if args.serve_requests > 0:
        from phygas_amd import serve
        mix = serve.request_mix(args.serve_requests, args.serve_rate, seed=5, n_clusters=len(clusters))
        serving = {"rate_hz": args.serve_rate, "requests": args.serve_requests,
                   "mix": "70% screening chunks of 256..20,000 stars (log-uniform), 30% pulsar jobs of 100..600 "
                          "clusters x 1,000 clones; Poisson arrivals, seed 5"}
        for name, f in (("fast", hip.FAST), ("plain", hip.PLAIN)):
            if f == hip.FAST and form != hip.FAST:
                continue
            with serve.Server(clusters, form=f, screen_slots=2, pair_slots=1, p_min=p_min) as srv:
                serving[name] = serve.run_mix(srv, mix)

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(clusters, stars, threads)

    value = pairs / (ms / 1000.0)
    line = {
        "metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if form == hip.PLAIN else "f32+f64", "data": "synthetic",
        "config": {"workload": "screen_20k_x_4087_with_shuffled_control", "stars": len(stars), "control_stars":
                   len(control), "clusters": len(clusters), "pairs_per_step": pairs, "chi2_max": chi2_max,
                   "n_samp": 64, "r_kernel": 2.0, "p_min": p_min, "form": result["form"],
                   "sub_passes": args.split if args.split > 0 else 3, "cluster_tables_s": tables_s},
        "stages_ms": stages, "rows": result["rows"], "kept_pairs": kept, "pass_stats": stats, "autotune": tune,
        "roofline": roof, "cpu_baseline": cpu,
    }
    if plain:
        line["plain"] = plain
        line["speedup_vs_plain"] = plain["ms"] / ms
    if serving:
        line["serving"] = serving
    print(json.dumps(line, default=lambda o: o.item() if hasattr(o, "item") else str(o)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

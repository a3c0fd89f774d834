"""Benchmark of the trace-back path on one MI355X.

Headline (value): the screening work unit, 20,000 synthetic stars x 4,087 synthetic clusters with the fused
shuffled-velocity control (stars seed 2, clusters seed 0, control permutation seed 3), rows with
p_r50 >= 0.001.  A step is one device pass over the whole unit (draws, tracks, scan, finish, row compaction)
with the cluster tables resident and the star records already in HBM.  The plain FP64 form is timed beside it
and the rows of the two compared.

Further sections of the same JSON line:
  roofline_items  the point evaluation (k_items): its fabric rate (bytes from the profiled traffic record) against
               the HBM peak and its FP64 rate against the vector peak; latency-bound, neither roof binds
  monte_carlo  1 pulsar x 300 prefiltered clusters x 200 clones, plain FP64 and FP32 clones, device draws
  two_stage    a 100,000-star shard (seed 4) with its control: the fast screening pass, then the Monte Carlo
               (1,000 clones, device draws, graph replay) of every real row with p_r50 >= 0.01
  serving      the request mix at swept Poisson rates, fast and plain forms (p50/p95/p99, pairs/s)
  unit_sizes   the fast and the plain form at 5,000, 10,000 and 40,000 stars (+ control), each unit autotuned
  cpu_baseline the reference's own code on the host: 1 thread and N threads, interleaved runs, medians

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form fast|plain] [--stars N] [--no-tune] [--split N]
                    [--no-cpu] [--no-plain] [--no-mc] [--no-two-stage] [--no-serve] [--no-sizes]

Before the timed steps the library's autotuner chooses the fast pass's settings on the unit (rows do not depend
on them); --no-tune times the defaults, --split N fixes the sub-passes and skips the tuner.

Prints one JSON line.
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
# Monte Carlo, per clone-step (two bodies, one kick-drift-kick each, then the step segment's closest approach):
PLAIN_CLONE_STEP_FLOP = 192.0   # FP64: 2 x (58 force (each pow counted once) + 20 kick-drift-kick) + 36 segment
FAST_CLONE_STEP_FLOP = 154.0    # 2 x 41 FP32 force + 2 x 18 FP64 kick-drift-kick + 36 FP64 segment


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


def _cpu_exe():
    ref = os.path.join(ROOT, "oracle", "_ref", "cpu_bench_ref")
    port = os.path.join(ROOT, "oracle", "cpu_bench_orc")
    if os.path.exists(ref):
        return ref, "reference"
    return (port, "port") if os.path.exists(port) else (None, None)


def _kv(line: str) -> dict:
    return {k: (float(v) if any(ch.isdigit() for ch in v) and k != "mode" else v)
            for k, v in (x.split("=") for x in line.split())}


def cpu_baseline(clusters, stars, pulsar, threads: int, reps: int = 3) -> dict | None:
    """The reference's own code (oracle/_ref, built from the reference sources) or, without it, the oracle
    restatement, on the host cores of the machine the bench runs on: the screening loop of phygas_screen on a bounded sample (25 stars
    per thread against every cluster, the chi2 cut and the finish) and ph_pair (5 pairs of 200 clones per
    thread), each at 1 thread and at `threads` threads, interleaved over `reps` rounds, medians reported."""
    exe, kind = _cpu_exe()
    if exe is None:
        return None
    n1, nN = 25, 25 * threads
    p1, pN = 5, 5 * threads
    with tempfile.TemporaryDirectory() as d:
        cf, sf, pf = (os.path.join(d, f) for f in ("clusters.txt", "stars.txt", "pulsar.txt"))
        wl.write_clusters(cf, clusters)
        wl.write_targets(sf, stars[:nN])
        wl.write_targets(pf, pulsar)
        scr = subprocess.run([exe, "screen", cf, sf, str(reps), f"1:{n1}", f"{threads}:{nN}"], capture_output=True,
                             text=True, check=True).stdout.splitlines()
        par = subprocess.run([exe, "pair", cf, pf, str(reps), "200", "0.05", f"1:{p1}", f"{threads}:{pN}"],
                             capture_output=True, text=True, check=True).stdout.splitlines()
    s1, sN = _kv(scr[0]), _kv(scr[1])
    m1, mN = _kv(par[0]), _kv(par[1])
    return {"value": sN["pairs_per_s"], "unit": "pairs/s", "cores": threads, "kind": kind,
            "sample": f"{nN} of the work unit's stars x {int(sN['clusters'])} clusters on {threads} threads, "
                      f"chi2 cut 25 + finish ({int(sN['kept'])} kept pairs); median of {reps} runs interleaved with "
                      f"the 1-thread runs; cluster tracks prepared first ({sN['cluster_prep_s']:.2f} s) and not timed",
            "screen_1_thread": {"pairs_per_s": s1["pairs_per_s"], "stars": int(s1["stars"]), "median_s": s1["median_s"],
                                "min_s": s1["min_s"], "max_s": s1["max_s"]},
            "screen_n_threads": {"threads": threads, "pairs_per_s": sN["pairs_per_s"], "stars": int(sN["stars"]),
                                 "median_s": sN["median_s"], "min_s": sN["min_s"], "max_s": sN["max_s"]},
            "pair_1_thread": {"clone_steps_per_s": m1["clone_steps_per_s"], "pairs": int(m1["pairs"]),
                              "median_s": m1["median_s"], "min_s": m1["min_s"], "max_s": m1["max_s"]},
            "pair_n_threads": {"threads": threads, "clone_steps_per_s": mN["clone_steps_per_s"],
                               "pairs": int(mN["pairs"]), "median_s": mN["median_s"], "min_s": mN["min_s"],
                               "max_s": mN["max_s"]},
            "runs": reps}


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


def _qrel(a, b) -> float:
    fs = ("d05", "d50", "d95", "t05", "t50", "t95", "v05", "v50", "v95")
    return max(float(np.max(np.abs(a[f] - b[f]) / np.maximum(np.abs(b[f]), 1e-300))) for f in fs)


def _median_ms(fn, reps: int) -> float:
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        t.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(t))


def monte_carlo(e: hip.Engine, clusters, pulsar, reps: int = 5) -> dict:
    """1 pulsar x 300 prefiltered clusters x 200 clones: the plain FP64 and the FP32 clone forms, clones drawn on
    the device (every result checked against the plain form), timed per call with the host round trip, and the
    clone kernel alone from stage events."""
    mask = e.prefilter(pulsar, clusters)[0]
    ci = np.nonzero(mask)[0]
    if len(ci) < 300:
        ci = np.concatenate([ci, np.setdiff1d(np.arange(len(clusters)), ci)[:300 - len(ci)]])
    ci = np.sort(ci[:300]).astype(np.int32)
    ti = np.zeros(len(ci), np.int32)
    n_clones, dt = 200, 0.05
    steps = float(np.sum(np.ceil(clusters["tmax_myr"][ci] / dt))) * n_clones
    out = {"workload": "1 pulsar x 300 prefiltered clusters x 200 clones, dt 0.05 Myr, device draws",
           "clone_steps": steps}
    res = {}
    for name, form, flop in (("plain", hip.PLAIN, PLAIN_CLONE_STEP_FLOP), ("fast", hip.FAST, FAST_CLONE_STEP_FLOP)):
        res[name] = e.pair(pulsar, clusters, ti, ci, n_clones, dt, form=form, device_draws=True)
        ms = _median_ms(lambda: e.pair(pulsar, clusters, ti, ci, n_clones, dt, form=form, device_draws=True), reps)
        lanes = e.pair_lanes(len(ci) * n_clones, form)
        e.set_timing(True)
        by_lanes = {}
        for setting in (1, 2, 0):   # the clone kernel alone at each lane count, then as the library chooses
            e.set_pair_lanes(setting)
            kt = []
            for _ in range(3):
                assert e.pair(pulsar, clusters, ti, ci, n_clones, dt, form=form, device_draws=True).tobytes() == \
                    res[name].tobytes()
                kt.append(e.stage_times())
            if setting:
                by_lanes[str(setting)] = float(np.median([t.get("clones", float("nan")) for t in kt]))
        st = kt[len(kt) // 2]
        e.set_timing(False)
        kms = float(np.median([t.get("clones", float("nan")) for t in kt]))
        achieved = steps * flop / (kms / 1e3) / 1e12
        out[name] = {"ms_per_job": ms, "clone_steps_per_s": steps / (ms / 1e3), "stages_ms": st,
                     "clone_kernel_clone_steps_per_s": steps / (kms / 1e3), "lanes_per_clone": lanes,
                     "clone_kernel_ms_by_lanes": by_lanes,
                     "roofline": {"bound": "valu", "kernel": "k_clones<%d, %d>" % (int(form == hip.FAST), lanes),
                                  "achieved": achieved, "peak": PEAK_VALU, "unit": "TFLOP/s",
                                  "frac": achieved / PEAK_VALU, "work_per_launch": steps * flop,
                                  "flop_per_clone_step": flop, "avg_launch_ms": kms}}
    host = e.pair(pulsar, clusters, ti, ci, n_clones, dt, form=hip.PLAIN, device_draws=False)
    out["fast_vs_plain_quantile_rel_max"] = _qrel(res["fast"], res["plain"])
    out["fast_vs_plain_f_max_abs"] = float(max(np.max(np.abs(res["fast"]["f_r50"] - res["plain"]["f_r50"])),
                                               np.max(np.abs(res["fast"]["f_rtot"] - res["plain"]["f_rtot"]))))
    out["device_vs_host_draws_quantile_rel_max"] = _qrel(res["plain"], host)
    out["speedup_fast_vs_plain"] = out["plain"]["ms_per_job"] / out["fast"]["ms_per_job"]
    return out


def unit_sizes(clusters, no_tune: bool, sizes=(5000, 10000, 40000), steps: int = 5) -> list:
    """The fast and the plain form at other unit sizes (N stars of seed 2 + their shuffled control x the 4,087
    clusters): the autotuner run on each unit first (unless --no-tune), then the pass time of each form and the
    rows of the two compared."""
    out = []
    for n in sizes:
        stars = wl.make_stars(n, 2)
        control = wl.shuffled_control(stars, 3)
        with hip.Engine(0) as e:
            e.load_clusters(clusters, 64, 2.0)
            e.load_stars(stars, 0)
            e.load_stars(control, 1)
            tune = None if no_tune else e.autotune_ex(hip.FAST, 25.0, 64, 0.001)
            fms, _ = run_form(e, hip.FAST, steps, 1, 25.0, 0.001, stage_passes=0)
            frows = e.rows()
            pms, _ = run_form(e, hip.PLAIN, 1, 1, 25.0, 0.001, stage_passes=0)
            prows = e.rows()
        a = {(r["star"], r["cluster"]): r["p_r50"] for r in frows}
        b = {(r["star"], r["cluster"]): r["p_r50"] for r in prows}
        common = set(a) & set(b)
        pairs = 2 * n * len(clusters)
        out.append({"stars": n, "control_stars": n, "pairs": pairs, "fast_ms": fms, "fast_pairs_per_s": pairs / (fms / 1e3),
                    "plain_ms": pms, "speedup_vs_plain": pms / fms, "rows": len(frows),
                    "rows_one_sided": len(set(a) ^ set(b)),
                    "max_abs_dp_r50": max((abs(a[k] - b[k]) for k in common), default=0.0),
                    "autotune": tune})
    return out


def two_stage(clusters, n_stars: int = 100000, n_clones: int = 1000, p_stage2: float = 0.01) -> dict:
    """A 100,000-star shard (seed 4) with its shuffled control: stage 1 the fast screening pass, stage 2 the
    Monte Carlo of every real row with p_r50 >= 0.01 (device draws, FP32 clones, the batch's launches replayed
    from a hipGraph), the plain FP64 clones timed beside it."""
    stars = wl.make_stars(n_stars, 4)
    control = wl.shuffled_control(stars, 3)
    out = {"workload": f"{n_stars:,} stars (seed 4) + shuffled control x {len(clusters)} clusters; stage 2: real rows "
                       f"with p_r50 >= {p_stage2}, {n_clones} clones each, dt 0.05 Myr"}
    with hip.Engine(0) as e:
        e.load_clusters(clusters, 64, 2.0)
        e.load_stars(stars, 0)
        e.load_stars(control, 1)
        e.screen_async(hip.FAST, 25.0, 64, 0.001)
        e.kept()
        rows = e.rows()
        t0 = time.perf_counter()
        e.screen_async(hip.FAST, 25.0, 64, 0.001)
        e.kept()
        rows = e.rows()
        s1 = (time.perf_counter() - t0) * 1e3
        sel = rows[(rows["flags"] == 0) & (rows["p_r50"] >= p_stage2)]
        si, ci = sel["star"].astype(np.int32), sel["cluster"].astype(np.int32)
        steps = float(np.sum(np.ceil(clusters["tmax_myr"][ci] / 0.05))) * n_clones
        e.set_graphs(True)
        res = {}
        for name, form in (("fast", hip.FAST), ("plain", hip.PLAIN)):
            res[name] = e.pair_loaded(si, ci, n_clones, 0.05, form)
            ms = _median_ms(lambda: e.pair_loaded(si, ci, n_clones, 0.05, form), 3)
            out[f"stage2_{name}_ms"] = ms
            out[f"stage2_{name}_clone_steps_per_s"] = steps / (ms / 1e3)
        cap, rep = e.pair_graph_stats()
        out["stage2_lanes_per_clone"] = {"fast": e.pair_lanes(len(si) * n_clones, hip.FAST),
                                         "plain": e.pair_lanes(len(si) * n_clones, hip.PLAIN)}
        # stage 2 at the config's expected scale (~60,000 pairs): the real rows with the highest p_r50
        big = rows[rows["flags"] == 0]
        big = big[np.argsort(-big["p_r50"], kind="stable")][:60000]
        bsi, bci = big["star"].astype(np.int32), big["cluster"].astype(np.int32)
        bsteps = float(np.sum(np.ceil(clusters["tmax_myr"][bci] / 0.05))) * n_clones
        e.pair_loaded(bsi, bci, n_clones, 0.05, hip.FAST)
        bms = _median_ms(lambda: e.pair_loaded(bsi, bci, n_clones, 0.05, hip.FAST), 2)
        out["stage2_at_60k_pairs"] = {"pairs": int(len(bsi)), "p_r50_min": float(big["p_r50"].min()) if len(big) else None,
                                      "clones": int(len(bsi)) * n_clones, "clone_steps": bsteps, "fast_ms": bms,
                                      "fast_clone_steps_per_s": bsteps / (bms / 1e3)}
    out.update({"stage1_ms": s1, "stage1_pairs": 2 * n_stars * len(clusters),
                "stage1_pairs_per_s": 2 * n_stars * len(clusters) / (s1 / 1e3), "rows": int(len(rows)),
                "stage2_pairs": int(len(si)), "stage2_clone_steps": steps,
                "total_ms_fast": s1 + out["stage2_fast_ms"], "graph_captures": cap, "graph_replays": rep,
                "stage2_fast_vs_plain_quantile_rel_max": _qrel(res["fast"], res["plain"]) if len(si) else 0.0})
    return out


def serving_sweep(clusters, p_min: float, rates_fast, rates_plain, n_req: int) -> dict:
    """The serving mix (70% screening chunks of 256..20,000 stars, log-uniform; 30% pulsar jobs of 100..600
    clusters x 1,000 clones; seed 5) at swept Poisson rates, for the fast and the plain form: latency
    percentiles, those of the large chunks (>= 10,000 stars), pairs/s and clones/s per rate."""
    from phygas_amd import serve
    base = serve.request_mix(n_req, 1.0, seed=5, n_clusters=len(clusters))   # arrivals at 1 Hz, rescaled per rate
    out = {"requests_per_rate": n_req, "mix": "70% screening chunks of 256..20,000 stars (log-uniform), 30% pulsar "
           "jobs of 100..600 clusters x 1,000 clones; Poisson arrivals, seed 5", "fast": [], "plain": []}
    for name, form, rates in (("fast", hip.FAST, rates_fast), ("plain", hip.PLAIN, rates_plain)):
        with serve.Server(clusters, form=form, screen_slots=2, pair_slots=1, p_min=p_min) as srv:
            for rate in rates:
                mix = [(t / rate, kind, payload) for t, kind, payload in base]
                r = serve.run_mix(srv, mix)
                r["rate_hz"] = rate
                out[name].append(r)
        ok = [r["rate_hz"] for r in out[name] if r.get("large_screen_p99_s") is not None and r["large_screen_p99_s"] <= 2.0]
        out[f"{name}_max_rate_large_chunk_p99_le_2s"] = max(ok) if ok else None
    return out


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
    ap.add_argument("--no-mc", action="store_true")
    ap.add_argument("--no-two-stage", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=100)
    ap.add_argument("--serve-rates-fast", default="15,60,240,960")
    ap.add_argument("--serve-rates-plain", default="15,30,60")
    ap.add_argument("--no-tune", action="store_true", help="time the library's default settings; otherwise the "
                    "autotuner (sub-passes and the last one's share, the tracks' offsets per lane, the scan's "
                    "first-level time tile, stars per task, launch grids) runs on the unit before the timed steps, "
                    "as a long run would")
    ap.add_argument("--tune", action="store_true", help=argparse.SUPPRESS)   # the default; kept for old commands
    ap.add_argument("--split", type=int, default=0, help="sub-passes of a fast pass on their own streams "
                    "(0: chosen by the autotuner, else fixed and the autotuner not run)")
    ap.add_argument("--no-sizes", action="store_true", help="skip the fast/plain comparison at other unit sizes")
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
    pulsar = wl.make_pulsars(1, 1)
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
    if form == hip.FAST and not args.no_tune and args.split <= 0:
        tune = e.autotune_ex(hip.FAST, chi2_max, 64, p_min)   # every setting, chosen before the timed steps
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
    roof_items = None
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
            work = stats["points"] * POINT_FLOP   # the points the evaluation takes (counted on the device)
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

        # the point evaluation beside it: the reference's FP64 arithmetic on the points of the tiles the scan
        # keeps.  It fetches ~720 bytes per item from beyond L2 at ~0.8 of the HBM peak, but those bytes do not
        # bound it (items ordered by XCD and tile index cut them 15x and its time 4%): it waits on each lane's
        # twenty tile loads and the chains of its unfused FP64 arithmetic
        i_ms = stages.get("items", float("nan"))
        i_work = stats["points"] * POINT_FLOP
        i_bytes = pmc_traffic("k_items")
        roof_items = {"bound": "latency", "kernel": "k_items", "avg_launch_ms": i_ms, "points": stats["points"],
                      "work_per_launch": i_work, "valu_tflops": i_work / (i_ms / 1000.0) / 1e12,
                      "valu_frac": i_work / (i_ms / 1000.0) / 1e12 / PEAK_VALU,
                      "traffic": i_bytes, "unit": "GB/s", "peak": PEAK_HBM,
                      "achieved": None if i_bytes is None else i_bytes / (i_ms / 1000.0) / 1e9}
        roof_items["frac"] = None if i_bytes is None else roof_items["achieved"] / PEAK_HBM

    sub_passes = e.tuning()["sub_passes"] if args.split <= 0 else args.split
    mc = None if args.no_mc else monte_carlo(e, clusters, pulsar)
    e.close()
    sizes = None if args.no_sizes or form != hip.FAST or args.no_plain else unit_sizes(clusters, args.no_tune)
    shard = None if args.no_two_stage else two_stage(clusters)
    serving = None
    if not args.no_serve and args.serve_requests > 0:
        rates_f = [float(x) for x in args.serve_rates_fast.split(",") if x]
        rates_p = [float(x) for x in args.serve_rates_plain.split(",") if x]
        serving = serving_sweep(clusters, p_min, rates_f if form == hip.FAST else [], rates_p, args.serve_requests)

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(clusters, stars, pulsar, threads)

    value = pairs / (ms / 1000.0)
    line = {
        "metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if form == hip.PLAIN else "f32+f64", "data": "synthetic",
        "config": {"workload": "screen_20k_x_4087_with_shuffled_control", "stars": len(stars), "control_stars":
                   len(control), "clusters": len(clusters), "pairs_per_step": pairs, "chi2_max": chi2_max,
                   "n_samp": 64, "r_kernel": 2.0, "p_min": p_min, "form": result["form"],
                   "sub_passes": sub_passes, "cluster_tables_s": tables_s},
        "stages_ms": stages, "rows": result["rows"], "kept_pairs": kept, "pass_stats": stats, "autotune": tune,
        "roofline": roof, "cpu_baseline": cpu,
    }
    if roof_items:
        line["roofline_items"] = roof_items
    if plain:
        line["plain"] = plain
        line["speedup_vs_plain"] = plain["ms"] / ms
    if sizes:
        line["unit_sizes"] = sizes
    if mc:
        line["monte_carlo"] = mc
    if shard:
        line["two_stage"] = shard
    if serving:
        line["serving"] = serving
    print(json.dumps(line, default=lambda o: o.item() if hasattr(o, "item") else str(o)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

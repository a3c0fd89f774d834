"""GPU parity of the fast forms.

* FAST64 (plain FP64 tracks + culled scan + screened finish) must give rows identical to PLAIN, byte for
  byte: this pins the exactness of the tile/point culling and of the draw screening of the finish.
* FAST (FP32-offset tracks) against the oracle: positions within 1e-9 relative, covariances within 2e-5
  (relative to sqrt(S_ii S_jj)), chi2 within 1e-4 relative; p_r50 / p_rtot counts identical for every kept
  pair on these inputs (the bar is |dp_r50| <= 1e-4, i.e. the same count); kept sets equal except
  pairs within 1e-4 of the chi2 cut.  Measured errors go to parity_fast.json in the report directory.
"""
import os

import numpy as np
import pytest

from phygas_amd import hip
from phygas_amd import io as pio

import parity_report

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INP = os.path.join(ROOT, "tests", "golden", "inputs")
REPORT = {}


def _report(key, value):
    REPORT[key] = float(value)
    parity_report.write("parity_fast.json", REPORT)


@pytest.fixture(scope="module")
def small():
    return pio.read_clusters(os.path.join(INP, "clusters400.txt")), pio.read_targets(os.path.join(INP, "stars20.txt"))


@pytest.fixture(scope="module")
def unit(workload):
    """A bigger sample: 4,087 clusters x 300 stars with their shuffled control."""
    cl = workload.make_clusters(4087, 0)
    st = workload.make_stars(300, 2)
    return cl, st, workload.shuffled_control(st, 3)


def test_finish_screen_margin():
    with hip.Engine(0) as e:
        worst = e.check_finish_screen(seed=20261203, n=1 << 23)
    _report("z0_estimate_max_abs_err", worst)
    assert worst < 1e-4      # the screen keeps draws within 1e-3 of the counting interval


@pytest.mark.parametrize("p_min", [0.0, 0.001])
def test_fast64_rows_identical_to_plain(unit, p_min):
    cl, st, ctl = unit
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st, 0)
        e.load_stars(ctl, 1)
        plain = e.screen(hip.PLAIN, chi2_max=25.0, p_min=p_min)
        kp = e.kept()
        fast = e.screen(hip.FAST64, chi2_max=25.0, p_min=p_min)
        kf = e.kept()
    _report(f"fast64_rows_p{p_min}", len(fast))
    assert kp == kf
    assert len(plain) == len(fast)
    assert plain.tobytes() == fast.tobytes()


def test_fast64_small_chi2_cut_identical(small):
    cl, st = small
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st)
        for cut in (1.0, 9.0, 100.0):
            a = e.screen(hip.PLAIN, chi2_max=cut, p_min=0.0)
            b = e.screen(hip.FAST64, chi2_max=cut, p_min=0.0)
            assert a.tobytes() == b.tobytes(), cut


def test_fast_tracks_match_oracle(small, oracle):
    cl, st = small
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st)
        e.screen(hip.FAST, chi2_max=25.0)
        trk = e.star_tracks()
    n, off, pos, vel, S = oracle.lin_targets(st, float(cl["tmax_myr"].max()))
    ep = es = 0.0
    for j in range(len(st)):
        sl = slice(off[j], off[j] + n[j])
        ep = max(ep, np.max(np.abs(trk[j, :, :3] - pos[sl]) / np.linalg.norm(pos[sl], axis=1, keepdims=True)))
        R = S[sl]
        d = np.sqrt(np.abs(R[:, [0, 0, 0, 3, 3, 5]] * R[:, [0, 3, 5, 3, 5, 5]])) + 1e-300
        es = max(es, np.max(np.abs(trk[j, :, 6:] - R) / d))
    _report("fast_track_pos_rel", ep)
    _report("fast_track_cov_rel", es)
    assert ep < 1e-9
    assert es < 2e-5


def test_fast_rows_match_oracle(small, oracle):
    cl, st = small
    want = oracle.screen(cl, st, chi2_max=25.0)
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st)
        got = e.screen(hip.FAST, chi2_max=25.0, p_min=0.0)
    a = {(r["star"], r["cluster"]): r for r in want}
    b = {(r["star"], r["cluster"]): r for r in got}
    for k in set(a) ^ set(b):     # only pairs at the cut may differ
        r = a.get(k, b.get(k))
        assert abs(r["chi2"] - 25.0) < 25.0 * 1e-4, (k, r["chi2"])
    common = sorted(set(a) & set(b))
    c2 = np.array([abs(a[k]["chi2"] - b[k]["chi2"]) / a[k]["chi2"] for k in common])
    dp50 = np.array([abs(a[k]["p_r50"] - b[k]["p_r50"]) for k in common])
    dpt = np.array([abs(a[k]["p_rtot"] - b[k]["p_rtot"]) for k in common])
    _report("fast_small_chi2_rel_max", c2.max())
    _report("fast_small_pairs", len(common))
    _report("fast_small_dp_r50_max", dp50.max())
    assert c2.max() < 1e-4
    assert dp50.max() <= 1e-4 and dpt.max() <= 1e-4


def test_fast_vs_plain_on_work_unit_sample(unit):
    """The chosen form against the plain FP64 GPU form on 300 stars + control x 4,087 clusters."""
    cl, st, ctl = unit
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st, 0)
        e.load_stars(ctl, 1)
        plain = e.screen(hip.PLAIN, chi2_max=25.0, p_min=0.0)
        fast = e.screen(hip.FAST, chi2_max=25.0, p_min=0.0)
    a = {(r["star"], r["cluster"]): r for r in plain}
    b = {(r["star"], r["cluster"]): r for r in fast}
    diff = set(a) ^ set(b)
    for k in diff:
        r = a.get(k, b.get(k))
        assert abs(r["chi2"] - 25.0) < 25.0 * 1e-4
    common = sorted(set(a) & set(b))
    dp = np.array([abs(a[k]["p_r50"] - b[k]["p_r50"]) for k in common])
    dpt = np.array([abs(a[k]["p_rtot"] - b[k]["p_rtot"]) for k in common])
    c2 = np.array([abs(a[k]["chi2"] - b[k]["chi2"]) / a[k]["chi2"] for k in common])
    _report("fast_vs_plain_pairs", len(common))
    _report("fast_vs_plain_only_one_side", len(diff))
    _report("fast_vs_plain_dp_r50_max", dp.max())
    _report("fast_vs_plain_frac_dp_r50_above_1e-4", np.mean(dp > 1e-4))
    _report("fast_vs_plain_frac_dp_rtot_above_1e-4", np.mean(dpt > 1e-4))
    _report("fast_vs_plain_chi2_rel_max", c2.max())
    assert dp.max() <= 1e-4 and dpt.max() <= 1e-4
    assert c2.max() < 1e-4


def _compare(want, got, chi2_rel, what):
    """Rows of one form against the reference rows: the same pairs except where chi2 lies within chi2_rel of
    the cut (25) on one side; every common pair with |dp_r50|, |dp_rtot| <= 1e-4 and chi2 within chi2_rel."""
    a = {(int(r["star"]), int(r["cluster"])): r for r in want}
    b = {(int(r["star"]), int(r["cluster"])): r for r in got}
    one_sided = set(a) ^ set(b)
    for k in one_sided:
        r = a.get(k, b.get(k))
        assert abs(r["chi2"] - 25.0) < 25.0 * chi2_rel, (what, k, r["chi2"])
    common = sorted(set(a) & set(b))
    dp = np.array([abs(a[k]["p_r50"] - b[k]["p_r50"]) for k in common])
    dpt = np.array([abs(a[k]["p_rtot"] - b[k]["p_rtot"]) for k in common])
    c2 = np.array([abs(a[k]["chi2"] - b[k]["chi2"]) / a[k]["chi2"] for k in common])
    _report(f"{what}_pairs", len(common))
    _report(f"{what}_one_sided", len(one_sided))
    _report(f"{what}_dp_r50_max", dp.max())
    _report(f"{what}_dp_rtot_max", dpt.max())
    _report(f"{what}_chi2_rel_max", c2.max())
    assert len(common) > 50
    assert dp.max() <= 1e-4 and dpt.max() <= 1e-4, what
    assert c2.max() < chi2_rel, (what, c2.max())
    return a, b, common


def test_work_unit_inputs_against_the_oracle(unit, oracle):
    """The work unit's own inputs (stars seed 2, shuffled-velocity control seed 3, 4,087 clusters) on a 300-star
    sample: the oracle's rows (p_r50 >= 0.001, the app's cut) against the plain FP64 form and the chosen form."""
    cl, st, ctl = unit
    nth = min(16, os.cpu_count() or 1)
    w_real = oracle.screen(cl, st, chi2_max=25.0, nthreads=nth)
    w_ctl = oracle.screen(cl, ctl, chi2_max=25.0, nthreads=nth)
    w_ctl["star"] += len(st)
    want = np.concatenate([w_real, w_ctl])
    want = want[want["p_r50"] >= 0.001]
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st, 0)
        e.load_stars(ctl, 1)
        plain = e.screen(hip.PLAIN, chi2_max=25.0, p_min=0.001)
        fast = e.screen(hip.FAST, chi2_max=25.0, p_min=0.001)
    a, b, common = _compare(want, plain, 1e-9, "unit300_plain_vs_oracle")
    for k in common:
        assert a[k]["kb"] == b[k]["kb"] and a[k]["sb"] == b[k]["sb"]
        for f in ("t", "d_nom_pc", "vrel"):
            assert abs(a[k][f] - b[k][f]) <= 1e-9 * max(abs(a[k][f]), 1e-12), (k, f)
    assert np.array_equal(plain["flags"], (plain["star"] >= len(st)).astype(plain["flags"].dtype))
    _compare(want, fast, 1e-4, "unit300_fast_vs_oracle")


def test_work_unit_slice_against_the_oracle(workload, oracle):
    """The first 2,000 stars of the 20,000-star work unit and their shuffled-velocity controls (the permutation of
    the whole unit's triplets, seed 3) against the oracle: rows with p_r50 >= 0.001 of the plain FP64 form (the
    1e-9 bar, kb and sb identical), of FAST64 (byte-identical to plain) and of the chosen form (the 1e-4 bar)."""
    cl = workload.make_clusters(4087, 0)
    st_all = workload.make_stars(20000, 2)
    st, ctl = st_all[:2000], workload.shuffled_control(st_all, 3)[:2000]
    nth = min(16, os.cpu_count() or 1)
    w_real = oracle.screen(cl, st, chi2_max=25.0, nthreads=nth)
    w_ctl = oracle.screen(cl, ctl, chi2_max=25.0, nthreads=nth)
    w_ctl["star"] += len(st)
    want = np.concatenate([w_real, w_ctl])
    want = want[want["p_r50"] >= 0.001]
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st, 0)
        e.load_stars(ctl, 1)
        plain = e.screen(hip.PLAIN, chi2_max=25.0, p_min=0.001)
        fast64 = e.screen(hip.FAST64, chi2_max=25.0, p_min=0.001)
        fast = e.screen(hip.FAST, chi2_max=25.0, p_min=0.001)
    assert len(want) > 1000
    a, b, common = _compare(want, plain, 1e-9, "unit2000_plain_vs_oracle")
    for k in common:
        assert a[k]["kb"] == b[k]["kb"] and a[k]["sb"] == b[k]["sb"], k
    assert plain.tobytes() == fast64.tobytes()
    _compare(want, fast, 1e-4, "unit2000_fast_vs_oracle")


def test_full_work_unit_fast_and_fast64_against_plain(workload):
    """The whole work unit, 20,000 stars + control x 4,087 clusters: FAST64 rows byte-identical to the plain FP64
    form's, the chosen FAST form's within the bar of every common pair."""
    cl = workload.make_clusters(4087, 0)
    st = workload.make_stars(20000, 2)
    ctl = workload.shuffled_control(st, 3)
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st, 0)
        e.load_stars(ctl, 1)
        plain = e.screen(hip.PLAIN, chi2_max=25.0, p_min=0.001)
        fast64 = e.screen(hip.FAST64, chi2_max=25.0, p_min=0.001)
        fast = e.screen(hip.FAST, chi2_max=25.0, p_min=0.001)
        head = e.rows(cap=1000)   # a truncated copy is the first rows in (star, cluster) order
    assert len(plain) > 10000
    assert plain.tobytes() == fast64.tobytes()
    _compare(plain, fast, 1e-4, "unit20k_fast_vs_plain")
    assert head.tobytes() == fast[:1000].tobytes()


def test_autotune_keeps_rows_and_stats_consistent(unit):
    """Settings chosen by the autotuner (the scan's first-level time tile, stars per task, launch grids) change no
    row; every combination of the scan settings gives the same rows, and the pass's work counts follow the
    levels it ran."""
    cl, st, ctl = unit
    cn = np.ceil(cl["tmax_myr"] / 0.5).astype(np.int64) + 1
    fine_tiles = int(np.sum((cn + 7) // 8))
    coarse_tiles = int(np.sum(((cn + 7) // 8 + 1) // 2))
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st, 0)
        e.load_stars(ctl, 1)
        before = e.screen(hip.FAST, chi2_max=25.0, p_min=0.0)
        plain = e.screen(hip.PLAIN, chi2_max=25.0, p_min=0.0)
        tune = e.autotune_ex(hip.FAST, 25.0, 64, 0.0)
        after = e.screen(hip.FAST, chi2_max=25.0, p_min=0.0)
        assert before.tobytes() == after.tobytes()
        for k in ("scan", "items", "reduce", "finish"):
            assert tune[f"{k}_blocks"] > 0
            _report(f"autotune_{k}_blocks", tune[f"{k}_blocks"])
        _report("autotune_stars_per_task", tune["stars_per_task"])
        _report("autotune_coarse_tiles", tune["coarse_tiles"])
        assert tune["stars_per_task"] in (8, 16) and tune["coarse_tiles"] in (0, 1) and tune["pass_ms"] > 0
        assert 1 <= tune["sub_passes"] <= 4 and 30 <= tune["last_share_pct"] <= 50   # the tuner's options
        seen = {}
        for coarse in (1, 0):
            for spw in (16, 8):
                e.set_tuning(stars_per_task=spw, coarse_tiles=coarse)
                rows = e.screen(hip.FAST, chi2_max=25.0, p_min=0.0)
                stats = e.pass_stats()
                assert rows.tobytes() == before.tobytes(), (coarse, spw)
                assert e.screen(hip.FAST64, chi2_max=25.0, p_min=0.0).tobytes() == plain.tobytes(), (coarse, spw)
                n = len(st) + len(ctl)
                # level 1 on coarse tiles (two tiles each) or on every tile; level 2 on its survivors (and, with the
                # coarse level, on the tiles of the coarse ones left); at most 8 points per item
                assert stats["l1_tests"] == n * (coarse_tiles if coarse else fine_tiles)
                assert stats["candidates"] <= stats["items"] <= stats["l2_tests"] <= 3 * stats["l1_tests"]
                assert stats["items"] < stats["points"] <= 8 * stats["items"]
                assert stats["rows"] == len(before) and stats["kept"] >= stats["rows"]
                if coarse in seen:   # stars per task change no test
                    assert stats == seen[coarse]
                seen[coarse] = stats
                _report(f"scan_l1_tests_coarse{coarse}", stats["l1_tests"])
                _report(f"scan_items_coarse{coarse}", stats["items"])


@pytest.mark.parametrize("ns", [0, 1, 17, 4100])
def test_fast64_edge_batches_identical_to_plain(workload, ns):
    """FAST64 rows equal PLAIN's byte for byte at the batch-size edges (empty, one star, a partial star block,
    a split pass), on clusters that include 1- to 4-point tracks, with a tight and a wide chi2 cut."""
    cl = workload.make_clusters(300, 0)
    cl["tmax_myr"][:4] = [0.0, 0.5, 1.0, 1.5]          # tracks of 1, 2, 3 and 4 points
    half = (ns + 1) // 2
    st = workload.make_stars(max(half, 1), 2)[:half]
    ctl = workload.shuffled_control(st, 3)[:ns - half] if ns > 1 else st[:0]
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st, 0)
        e.load_stars(ctl, 1)
        assert e.ns == len(st) + len(ctl)
        for chi2_max, p_min in ((4.0, 0.0), (25.0, 0.0), (100.0, 0.001)):
            plain = e.screen(hip.PLAIN, chi2_max=chi2_max, p_min=p_min)
            kp = e.kept()
            f64 = e.screen(hip.FAST64, chi2_max=chi2_max, p_min=p_min)
            kf = e.kept()
            assert kp == kf and plain.tobytes() == f64.tobytes(), (ns, chi2_max, p_min)
            if ns == 0:
                assert len(plain) == 0


def test_long_tracks_wide_tile_masks(workload, oracle):
    """Trace-backs of 120-250 Myr (up to 501 points, 63 tiles): the scan's 64-bit tile masks, whose LDS half
    moves to tiles 32-63 once the first half is done.  FAST64 rows equal PLAIN's byte for byte, and PLAIN's
    the oracle's on the first 40 stars."""
    cl = workload.make_clusters(400, 0)
    rng = np.random.default_rng(7)
    cl["tmax_myr"] = rng.uniform(120.0, 250.0, len(cl))   # 241-501 points
    cl["tmax_myr"][:3] = [250.0, 200.0, 100.0]
    st = workload.make_stars(300, 2)
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st, 0)
        assert e.npts == 501
        for chi2_max, p_min in ((25.0, 0.0), (100.0, 0.001)):
            plain = e.screen(hip.PLAIN, chi2_max=chi2_max, p_min=p_min)
            kp = e.kept()
            f64 = e.screen(hip.FAST64, chi2_max=chi2_max, p_min=p_min)
            kf = e.kept()
            assert kp == kf and plain.tobytes() == f64.tobytes(), (chi2_max, p_min)
            if chi2_max == 25.0:
                assert len(plain) > 100 and int(plain["kb"].max()) >= 8 * 32   # minima past tile 32
                sub = plain[plain["star"] < 40]
                want = oracle.screen(cl, st[:40], chi2_max=25.0, use_cut=True)
                assert len(sub) == len(want) > 10
                assert np.array_equal(sub["star"], want["star"]) and np.array_equal(sub["cluster"], want["cluster"])
                assert np.array_equal(sub["kb"], want["kb"]) and np.array_equal(sub["p_r50"], want["p_r50"])
                assert np.allclose(sub["chi2"], want["chi2"], rtol=1e-9, atol=1e-12)


def test_scan_test_counts_match_the_cpu_order(workload):
    """The GPU scan (level-2 entries collected across coarse tiles in rings) runs the tile tests of the CPU
    model of the same two-level order (tests/native/cull_harness.cpp cc_scan2, fixture made by
    tests/golden/make_scan_counts.py): the coarse level-1 tests exactly (they depend on the track lengths
    only), the level-2 tests and the pairs with a tile left within 0.1% (the CPU runs on the oracle's tracks,
    the GPU on its own FP64 tracks, equal to 1e-15, so a test at the edge may go the other way)."""
    import json
    with open(os.path.join(ROOT, "tests", "golden", "scan_counts.json")) as f:
        want = json.load(f)
    cl = workload.make_clusters(want["clusters"], want["cluster_seed"])
    st = workload.make_stars(want["stars"], want["star_seed"])
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st, 0)
        e.screen(hip.FAST64, chi2_max=want["chi2_max"], p_min=0.0)
        stats = e.pass_stats()
    l2 = want["coarse_l1_passes"] + want["fine_tests"]
    _report("scan_l2_tests_vs_cpu_rel", abs(stats["l2_tests"] - l2) / l2)
    _report("scan_candidates_vs_cpu", stats["candidates"] - want["pairs_with_a_tile_left"])
    assert stats["l1_tests"] == want["coarse_l1_tests"]
    assert abs(stats["l2_tests"] - l2) <= 1e-3 * l2, (stats["l2_tests"], l2)
    assert abs(stats["candidates"] - want["pairs_with_a_tile_left"]) <= 1e-3 * want["pairs_with_a_tile_left"]


def test_split_passes_give_the_same_rows(workload):
    """A pass split into 2, 3 or 4 sub-passes on their own streams, with the last one's share of the stars from 30
    to 60%, gives the whole pass's rows, work counts and tracks (also with hipGraph replays)."""
    cl = workload.make_clusters(4087, 0)
    st = workload.make_stars(2500, 2)
    ctl = workload.shuffled_control(st, 3)
    out = {}
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st, 0)
        e.load_stars(ctl, 1)
        npts = e.npts
        for parts in (1, 2, 3, 4):
            e.set_split(parts)
            for graphs in (False, True):
                e.set_graphs(graphs)
                rows = [e.screen(hip.FAST, chi2_max=25.0, p_min=p) for p in (0.0, 0.001)]
                out[(parts, graphs)] = (rows, e.pass_stats(), e.star_tracks(1240, 20))
        e.set_graphs(False)
        for parts, share in ((3, 30), (3, 50), (2, 60)):   # the last sub-pass's share of the stars
            e.set_tuning(sub_passes=parts, last_share_pct=share)
            rows = [e.screen(hip.FAST, chi2_max=25.0, p_min=p) for p in (0.0, 0.001)]
            out[(parts, share)] = (rows, e.pass_stats(), e.star_tracks(1240, 20))
        e.set_tuning(sub_passes=3, last_share_pct=35)
    ref_rows, ref_stats, ref_trk = out[(1, False)]
    assert len(ref_rows[0]) > 1000 and len(ref_rows[1]) > 50
    assert ref_trk.shape == (20, npts, 12)
    for key, (rows, stats, trk) in out.items():
        for a, b in zip(rows, ref_rows):
            assert a.tobytes() == b.tobytes(), key
        assert stats == ref_stats, key
        assert trk.tobytes() == ref_trk.tobytes(), key


@pytest.fixture(scope="module")
def pulsar_unit(workload):
    """1 synthetic pulsar x 300 clusters x 200 clones, dt 0.05 Myr (clones drawn on the host, seed per pair)."""
    return workload.make_pulsars(1, 1), workload.make_clusters(300, 1)


@pytest.fixture(scope="module")
def pulsar_oracle(pulsar_unit, oracle):
    t, cl = pulsar_unit
    ti, ci = np.zeros(len(cl), np.int32), np.arange(len(cl), dtype=np.int32)
    return ti, ci, oracle.pair_run(t, cl, ti, ci, n_clones=200, dt_myr=0.05, nthreads=16)


@pytest.mark.parametrize("form", ["plain", "fast"])
def test_pair_forms_match_oracle(pulsar_unit, pulsar_oracle, form):
    """Monte Carlo quantiles within 1e-3 relative (the Monte Carlo bar) and identical clone counts; the plain form
    within 1e-9.  Fractions may differ only by clones within rounding of r50 / rtot."""
    t, cl = pulsar_unit
    ti, ci, want = pulsar_oracle
    with hip.Engine(0) as e:
        got = e.pair(t, cl, ti, ci, n_clones=200, dt_myr=0.05, form=hip.PLAIN if form == "plain" else hip.FAST)
    assert np.array_equal(got["n"], want["n"])
    tol = 1e-9 if form == "plain" else 1e-3
    worst = 0.0
    for f in ("d05", "d50", "d95", "t05", "t50", "t95", "v05", "v50", "v95"):
        rel = np.abs(got[f] - want[f]) / np.maximum(np.abs(want[f]), 1e-6)
        worst = max(worst, rel.max())
        assert rel.max() < (tol if not f.startswith("t") else max(tol, 1e-6)), (f, rel.max())
    _report(f"pair_{form}_quantile_rel_max", worst)
    dfr = np.abs(got["f_r50"] - want["f_r50"]) * 200
    dft = np.abs(got["f_rtot"] - want["f_rtot"]) * 200
    _report(f"pair_{form}_fraction_clone_flips", float(dfr.sum() + dft.sum()))
    assert dfr.max() <= (0 if form == "plain" else 1) and dft.max() <= (0 if form == "plain" else 1)


def test_device_draws_match_host_reference(workload, oracle):
    """Clones drawn on the device (speculative prefix over the splitmix64 counter) give the reference's clone
    counts and, with the plain FP64 tracer, the same quantiles to 1e-9; the second target's parallax draws
    fail and retry often, exercising the irregular path."""
    tg = np.concatenate([workload.make_pulsars(1, 1), workload.make_stars(2, 7)])
    tg[2]["plx"], tg[2]["e_plx"] = 0.15, 0.3
    cl = workload.make_clusters(40, 2)
    ti = np.repeat(np.arange(3, dtype=np.int32), 40)
    ci = np.tile(np.arange(40, dtype=np.int32), 3)
    want = oracle.pair_run(tg, cl, ti, ci, n_clones=150, dt_myr=0.1, nthreads=16)
    with hip.Engine(0) as e:
        got = e.pair(tg, cl, ti, ci, n_clones=150, dt_myr=0.1, form=hip.PLAIN, device_draws=True)
        fast = e.pair(tg, cl, ti, ci, n_clones=150, dt_myr=0.1, form=hip.FAST, device_draws=True)
    assert np.array_equal(got["n"], want["n"]) and np.array_equal(fast["n"], want["n"])
    worst = 0.0
    for f in ("d05", "d50", "d95", "v05", "v50", "v95"):
        rel = np.abs(got[f] - want[f]) / np.maximum(np.abs(want[f]), 1e-6)
        worst = max(worst, rel.max())
        assert rel.max() < 1e-9, f
        assert (np.abs(fast[f] - want[f]) / np.maximum(np.abs(want[f]), 1e-6)).max() < 1e-3, f
    _report("device_draws_quantile_rel_max", worst)
    assert np.array_equal(got["f_r50"], want["f_r50"])


def test_track_offsets_give_the_same_bits(workload):
    """The fast tracks with 1, 2 or 3 offsets per lane (0 restores the default, 2): tracks, rows and work counts
    bit for bit the same.  171 stars + control leave the last workgroup of 64 stars partly empty."""
    cl = workload.make_clusters(4087, 0)
    st = workload.make_stars(171, 2)
    ctl = workload.shuffled_control(st, 3)
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st, 0)
        e.load_stars(ctl, 1)
        n = len(st) + len(ctl)
        ref = None
        for offsets in (0, 1, 2, 3, 0):
            e.set_track_offsets(offsets)
            rows = e.screen(hip.FAST, chi2_max=25.0, p_min=0.0)
            got = (rows.tobytes(), e.pass_stats(), e.star_tracks(0, n).tobytes())
            if ref is None:
                ref = got
                assert len(rows) > 100
            assert got == ref, offsets
        e.set_track_offsets(0)

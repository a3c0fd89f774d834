"""Work-unit runner on the GPU (phx_wu_run): the app's rows for a screening and a pair-mode work unit against
the oracle (rows formatted as the reference app prints them, src/main_phygas_boinc.cpp:165,256), and a run
interrupted after two batches, with a torn row left in `out`, resumed into the same bytes as an uninterrupted
run."""
import os

import numpy as np
import pytest

from phygas_amd import hip
from phygas_amd import io as pio

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INP = os.path.join(ROOT, "tests", "golden", "inputs")
CLUSTERS = os.path.join(INP, "clusters400.txt")


def _s(b):
    return b.decode() if isinstance(b, bytes) else str(b)


def _write_in(path, header, lines):
    with open(path, "w") as f:
        f.write(header + "\n")
        for ln in lines:
            f.write(ln.rstrip("\n") + "\n")


def _star_lines():
    with open(os.path.join(INP, "stars20.txt")) as f:
        return [ln for ln in f if ln.strip()]


def _screen_expected(oracle, tg, cl, p_min):
    rows = oracle.screen(cl, tg, chi2_max=25.0, n_samp=64, r_kernel=2.0)
    out = []
    for r in rows:
        if r["p_r50"] < p_min:
            continue
        t, c = tg[r["star"]], cl[r["cluster"]]
        out.append(f"{_s(t['id'])} {_s(c['name'])} {_s(c['kind'])} {r['chi2']:.3f} {r['t']:.3f} {r['d_nom_pc']:.2f} "
                   f"{r['sig_pc']:.2f} {r['p_r50']:.4f} {r['p_rtot']:.4f} {r['vrel']:.2f}\n")
    return out + [f"# done {len(tg)}\n"]


def _lines(path):
    with open(path) as f:
        return f.readlines()


def test_screen_work_unit_rows_match_oracle(oracle, tmp_path):
    inp, out, ck = str(tmp_path / "in"), str(tmp_path / "out"), str(tmp_path / "ckpt")
    _write_in(inp, "PHYGAS 1 screen 25 64 2 0", _star_lines())
    rc, done = hip.wu_run(inp, CLUSTERS, out, ck, form=hip.PLAIN, batch_targets=8)
    assert rc == 0 and done == 20
    tg, cl = pio.read_targets(os.path.join(INP, "stars20.txt")), pio.read_clusters(CLUSTERS)
    want, got = _screen_expected(oracle, tg, cl, 0.0), _lines(out)
    assert len(got) == len(want) and len(want) > 5
    assert got == want


def test_default_options_give_the_oracle_rows(oracle, tmp_path):
    """With phx_wu_defaults (FAST64 screening, FP64 clones) the app's rows are the reference's."""
    o = hip.wu_defaults()
    assert (o.form, o.pair_form) == (hip.FAST64, hip.PLAIN)
    inp, out, ck = str(tmp_path / "in"), str(tmp_path / "out"), str(tmp_path / "ckpt")
    _write_in(inp, "PHYGAS 1 screen 25 64 2 0", _star_lines())
    assert hip.wu_run(inp, CLUSTERS, out, ck, batch_targets=7) == (hip.WU_DONE, 20)
    tg, cl = pio.read_targets(os.path.join(INP, "stars20.txt")), pio.read_clusters(CLUSTERS)
    want = _screen_expected(oracle, tg, cl, 0.0)
    assert len(want) > 5 and _lines(out) == want
    with open(os.path.join(INP, "pulsar1.txt")) as f:
        pulsar = [ln for ln in f if ln.strip()][0]
    pin, pout, pck = str(tmp_path / "pin"), str(tmp_path / "pout"), str(tmp_path / "pckpt")
    _write_in(pin, "PHYGAS 1 pair 40 0.05 0 0", [pulsar])
    assert hip.wu_run(pin, CLUSTERS, pout, pck) == (hip.WU_DONE, 1)
    pt = pio.read_targets(os.path.join(INP, "pulsar1.txt"))[:1]
    mask = oracle.prefilter_mask(pt, cl)[0]
    ci = list(np.nonzero(mask)[0])
    res = oracle.pair_run(pt, cl, [0] * len(ci), ci, 40, 0.05)
    want = [f"{_s(pt[0]['id'])} {_s(cl[i]['name'])} {_s(cl[i]['kind'])} {r['n']} {r['d05']:.3f} {r['d50']:.3f} "
            f"{r['d95']:.3f} {r['t05']:.4f} {r['t50']:.4f} {r['t95']:.4f} {r['v05']:.2f} {r['v50']:.2f} "
            f"{r['v95']:.2f} {r['f_r50']:.4f} {r['f_rtot']:.4f}\n" for i, r in zip(ci, res)] + ["# done 1\n"]
    assert len(want) > 10 and _lines(pout) == want


def test_resumed_run_is_identical(tmp_path):
    inp = str(tmp_path / "in")
    _write_in(inp, "PHYGAS 1 screen 25 64 2 0", _star_lines())
    ref_out, ref_ck = str(tmp_path / "ref_out"), str(tmp_path / "ref_ckpt")
    assert hip.wu_run(inp, CLUSTERS, ref_out, ref_ck, form=hip.FAST, batch_targets=4) == (0, 20)
    out, ck = str(tmp_path / "out"), str(tmp_path / "ckpt")
    assert hip.wu_run(inp, CLUSTERS, out, ck, form=hip.FAST, batch_targets=4, max_batches=2) == (hip.WU_STOPPED, 8)
    with open(out, "a") as f:
        f.write("S0000009 C0001 o 12.3")              # a row torn by the interruption
    assert hip.wu_run(inp, CLUSTERS, out, ck, form=hip.FAST, batch_targets=4) == (0, 20)
    with open(out, "rb") as a, open(ref_out, "rb") as b:
        assert a.read() == b.read()
    # a finished task run again writes the same file
    assert hip.wu_run(inp, CLUSTERS, out, ck, form=hip.FAST, batch_targets=4) == (0, 20)
    with open(out, "rb") as a, open(ref_out, "rb") as b:
        assert a.read() == b.read()


def test_pair_work_unit_rows_match_oracle(oracle, tmp_path):
    cl = pio.read_clusters(CLUSTERS)
    names = [_s(c) for c in cl["name"][[5, 17, 300]]]
    with open(os.path.join(INP, "pulsar1.txt")) as f:
        pulsar = [ln for ln in f if ln.strip()][0]
    star = _star_lines()[3].rstrip("\n") + " clusters=" + ",".join(names)
    inp, out, ck = str(tmp_path / "in"), str(tmp_path / "out"), str(tmp_path / "ckpt")
    _write_in(inp, "PHYGAS 1 pair 40 0.05 0 0", [pulsar, star])
    rc, done = hip.wu_run(inp, CLUSTERS, out, ck, pair_form=hip.PLAIN, batch_targets=2)
    assert rc == 0 and done == 2
    tg = np.concatenate([pio.read_targets(os.path.join(INP, "pulsar1.txt"))[:1],
                         pio.read_targets(os.path.join(INP, "stars20.txt"))[3:4]])
    mask = oracle.prefilter_mask(tg[:1], cl)[0]
    ti = [0] * int(mask.sum()) + [1, 1, 1]
    ci = list(np.nonzero(mask)[0]) + [5, 17, 300]
    res = oracle.pair_run(tg, cl, ti, ci, 40, 0.05)
    want = []
    for j, i, r in zip(ti, ci, res):
        want.append(f"{_s(tg[j]['id'])} {_s(cl[i]['name'])} {_s(cl[i]['kind'])} {r['n']} {r['d05']:.3f} {r['d50']:.3f} "
                    f"{r['d95']:.3f} {r['t05']:.4f} {r['t50']:.4f} {r['t95']:.4f} {r['v05']:.2f} {r['v50']:.2f} "
                    f"{r['v95']:.2f} {r['f_r50']:.4f} {r['f_rtot']:.4f}\n")
    want.append("# done 2\n")
    got = _lines(out)
    assert len(got) == len(want) and len(want) > 10
    assert got == want

"""Arithmetic of the linearised star tracks, measured on the CPU (tests/native/precision_harness.cpp, a host build
of the fast form's FP32 force and difference force): each candidate's error against the oracle's FP64 tracks, on
config-2 stars (seed 2) against the 4,087-cluster table, through the oracle's own chi2 cut and finish.  chi2
errors are relative to max(chi2, 1).

* mode 0 (the reference's FP64 arithmetic, built from the product's headers) must reproduce the oracle's tracks
  bit for bit: that pins the harness.
* mode 4, the fast form's split (FP64 nominal, six offsets driven by the FP32 difference force): every count the
  same, chi2 within 1e-4 (the error comes from the covariances, not the nominal).
* mode 2, the FP32 nominal with Kahan-compensated (float-float) position and velocity and the same offsets:
  positions within 1e-7, but one draw of star 303 x cluster 443 changes side of the r_tot sphere
  (|dp_rtot| = 1/512, over the 1e-4 bar): the reason the fast form keeps the nominal in FP64.
* mode 3, the FP32 nominal without compensation: positions off by 1e-5, chi2 by 1e-3.
* mode 1, all seven orbits in FP32 with compensation, the matrix by differencing them: the 1e-5 offsets lose
  their FP32 force difference (covariances wrong by their own size), so the difference force is needed.
Measured figures go to precision_modes.json in the report directory."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import parity_report

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# nine stars of the config-2 generator and three (204, 207, 303 of its first 400) whose pairs show the FP32
# nominal's errors: 400 stars x 4,087 clusters measured 75,260 kept pairs, of which these held every chi2
# difference above 1e-4 and the one count that changed
STARS = list(range(9)) + [204, 207, 303]
CHI2 = 25.0


class _Lin(C.Structure):
    _fields_ = [("n", C.c_int), ("pos", C.c_void_p), ("vel", C.c_void_p), ("S", C.c_void_p)]


class _Screen(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("chi2", "t", "d_nom_pc", "sig_pc", "p_r50", "p_rtot", "vrel")]


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("prec") / "libprec.so")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Xclang",
           "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "phygas_amd", "csrc"),
           os.path.join(ROOT, "tests", "native", "precision_harness.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr[-2000:])
    lib = C.CDLL(out)
    lib.pr_tracks.restype = C.c_int
    lib.pr_tracks.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    return lib


def _aos(n, off, pos, vel, S, npts):
    out = np.zeros((len(n), npts, 12))
    for i in range(len(n)):
        sl = slice(off[i], off[i] + n[i])
        out[i, :n[i], 0:3] = pos[sl]
        out[i, :n[i], 3:6] = vel[sl]
        out[i, :n[i], 6:12] = S[sl]
    return out


@pytest.fixture(scope="module")
def setup(oracle, workload):
    cl = workload.make_clusters(4087, 0)
    st = workload.make_stars(400, 2)[STARS]
    tmax = float(cl["tmax_myr"].max())
    npts = int(np.ceil(tmax / 0.5)) + 1
    ps0, sig = oracle.lin_init_targets(st)
    ref = _aos(*oracle.lin_targets(st, tmax), npts)
    n, off, pos, vel, S = oracle.lin_clusters(cl)
    return dict(cl=cl, st=st, npts=npts, ps0=np.ascontiguousarray(ps0), sig=np.ascontiguousarray(sig), ref=ref,
                clu=(n, off, pos, vel, S))


def _rows(oracle, s, trk):
    """The oracle's chi2 cut and finish of every (star, cluster) pair on the given star tracks: {pair: (chi2,
    p_r50, p_rtot, kb)} for the kept pairs."""
    from oracle.oracle_py import _cl
    lib = oracle.lib
    lib.ph_screen_chi2_cut.restype = C.c_double
    lib.ph_screen_chi2_cut.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_int),
                                       C.POINTER(C.c_double)]
    lib.ph_screen_finish.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_double, C.c_int,
                                     C.c_double, C.c_void_p]
    cl, st = _cl(s["cl"]), s["st"]
    n, off, pos, vel, S = s["clu"]
    keep = []   # the arrays the cluster PhLin views point into
    clin = []
    for c in range(len(cl)):
        sl = slice(off[c], off[c] + n[c])
        a = [np.ascontiguousarray(pos[sl]), np.ascontiguousarray(vel[sl]), np.ascontiguousarray(S[sl])]
        keep.append(a)
        clin.append(_Lin(int(n[c]), a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data))
    out = {}
    kb, sb = C.c_int(), C.c_double()
    for j in range(len(st)):
        a = [np.ascontiguousarray(trk[j, :, 0:3]), np.ascontiguousarray(trk[j, :, 3:6]),
             np.ascontiguousarray(trk[j, :, 6:12])]
        tl = _Lin(s["npts"], a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data)
        for c in range(len(cl)):
            rec = cl[c:c + 1].ctypes.data
            chi2 = lib.ph_screen_chi2_cut(C.byref(tl), C.byref(clin[c]), rec, CHI2, C.byref(kb), C.byref(sb))
            if not chi2 < CHI2:
                continue
            r = _Screen()
            lib.ph_screen_finish(C.byref(tl), C.byref(clin[c]), rec, oracle.seed(st[j]["id"], cl[c]["name"]), chi2,
                                 kb.value, sb.value, C.byref(r))
            out[(j, c)] = (chi2, r.p_r50, r.p_rtot, kb.value)
    return out


@pytest.fixture(scope="module")
def ref_rows(oracle, setup):
    return _rows(oracle, setup, setup["ref"])


REPORT = {}


def _measure(harness, oracle, setup, ref_rows, mode):
    s = setup
    got = np.zeros((len(s["st"]), s["npts"], 12))
    assert harness.pr_tracks(mode, s["ps0"].ctypes.data, s["sig"].ctypes.data, len(s["st"]), s["npts"],
                             got.ctypes.data) == 0
    ref = s["ref"]
    dx = np.linalg.norm(got[..., 0:3] - ref[..., 0:3], axis=-1)
    pos_rel = float((dx / np.linalg.norm(ref[..., 0:3], axis=-1)).max())
    den = np.sqrt(ref[..., [6, 6, 6, 9, 9, 11]] * ref[..., [6, 9, 11, 9, 11, 11]])
    s_rel = float((np.abs(got[..., 6:12] - ref[..., 6:12]) / den).max())
    rows = _rows(oracle, s, got)
    # chi2 errors relative to max(chi2, 1): a minimum near 0 is a difference of nearly equal positions
    both = sorted(set(rows) & set(ref_rows))
    m = dict(pos_rel=pos_rel, S_rel=s_rel, rows=len(rows), ref_rows=len(ref_rows),
             one_sided=len(set(rows) ^ set(ref_rows)),
             chi2_rel=max(abs(rows[k][0] - ref_rows[k][0]) / max(ref_rows[k][0], 1.0) for k in both),
             dp_r50=max(abs(rows[k][1] - ref_rows[k][1]) for k in both),
             dp_rtot=max(abs(rows[k][2] - ref_rows[k][2]) for k in both),
             kb_differs=sum(rows[k][3] != ref_rows[k][3] for k in both))
    REPORT[f"mode{mode}"] = m
    parity_report.write("precision_modes.json", REPORT)
    return got, m


def test_harness_reproduces_the_oracle(harness, oracle, setup, ref_rows):
    got, m = _measure(harness, oracle, setup, ref_rows, 0)
    assert np.array_equal(got, setup["ref"])
    assert len(ref_rows) > 1000


def test_fast_form_split(harness, oracle, setup, ref_rows):
    _, m = _measure(harness, oracle, setup, ref_rows, 4)
    assert m["pos_rel"] == 0.0 and m["S_rel"] < 1e-4 and m["chi2_rel"] < 1e-4, m
    assert m["one_sided"] == 0 and m["dp_r50"] == 0.0 and m["dp_rtot"] == 0.0 and m["kb_differs"] == 0, m


def test_fp32_nominal_with_compensation(harness, oracle, setup, ref_rows):
    _, m = _measure(harness, oracle, setup, ref_rows, 2)
    assert 1e-9 < m["pos_rel"] < 1e-7 and m["S_rel"] < 1e-4 and m["one_sided"] == 0, m
    assert m["dp_rtot"] > 1e-4, m   # a count changes


def test_fp32_nominal_without_compensation(harness, oracle, setup, ref_rows):
    _, m = _measure(harness, oracle, setup, ref_rows, 3)
    assert m["pos_rel"] > 1e-6 and m["chi2_rel"] > 1e-3 and m["dp_rtot"] > 1e-4, m


def test_seven_fp32_orbits_differenced(harness, oracle, setup, ref_rows):
    _, m = _measure(harness, oracle, setup, ref_rows, 1)
    assert m["S_rel"] > 0.1 and m["chi2_rel"] > 0.1 and m["one_sided"] > 0, m

"""Tile culling of the fast scan (phygas_amd/csrc/cull.h), built for the host: no tile that the two tests
drop may hold a point under the chi2 cut.  Checked against the chi2 of every point of every (star, cluster)
pair computed directly from the oracle's FP64 tracks, and against the oracle's kept pairs (the tile of each
kept pair's minimum must survive).  Also bounds the pass rates, so that a vacuous test would fail."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CHI2 = 25.0


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cull") / "libcull.so")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Xclang",
           "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "phygas_amd", "csrc"), os.path.join(ROOT, "tests", "native", "cull_harness.cpp"),
           "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr[-2000:])
    lib = C.CDLL(out)
    lib.cc_scan.restype = C.c_int
    P, I, D = C.c_void_p, C.c_int, C.c_double
    lib.cc_scan.argtypes = [P, I, I, P, P, P, I, I, D, P, P, P]
    return lib


def _aos(n, off, pos, vel, S, npts):
    out = np.zeros((len(n), npts, 12))
    for i in range(len(n)):
        sl = slice(off[i], off[i] + n[i])
        out[i, :n[i], 0:3] = pos[sl]
        out[i, :n[i], 3:6] = vel[sl]
        out[i, :n[i], 6:12] = S[sl]
    return out


def _point_chi2(st, ct, cn, R2):
    """chi2 of every point of every pair, [ns][nc][npts] (inf past a cluster's track), computed directly:
    min(q(m0), q(ms)) with ms the point of the step's segment closest to the origin."""
    ns, npts = st.shape[0], ct.shape[1]
    m0 = st[:, None, :, 0:3] - ct[None, :, :, 0:3]
    m1 = np.concatenate([m0[:, :, 1:], m0[:, :, -1:]], axis=2)
    r = m1 - m0
    dd = (r * r).sum(-1)
    ad = (m0 * r).sum(-1)
    s = np.where(dd > 0, np.clip(-ad / np.where(dd > 0, dd, 1.0), 0.0, 1.0), 0.0)
    ms = m0 + r * s[..., None]
    S6 = st[:, None, :, 6:12] + ct[None, :, :, 6:12]
    M = np.empty(S6.shape[:-1] + (3, 3))
    idx = [(0, 0, 0), (0, 1, 1), (0, 2, 2), (1, 1, 3), (1, 2, 4), (2, 2, 5)]
    for a, b, q in idx:
        M[..., a, b] = S6[..., q]
        M[..., b, a] = S6[..., q]
    for a in range(3):
        M[..., a, a] += R2[None, :, None]
    valid = np.arange(npts)[None, None, :] < cn[None, :, None]
    M = np.where(valid[..., None, None], M, np.eye(3))
    Mi = np.linalg.inv(M)
    q0 = np.einsum("...i,...ij,...j->...", m0, Mi, m0)
    qs = np.einsum("...i,...ij,...j->...", ms, Mi, ms)
    last = np.arange(npts)[None, None, :] == (cn[None, :, None] - 1)
    q = np.where(last, q0, np.minimum(q0, qs))
    return np.where(valid, q, np.inf)


def test_dropped_tiles_hold_no_point_under_the_cut(harness, oracle, workload):
    cl = workload.make_clusters(4087, 0)
    st = workload.make_stars(6, 2)
    n, off, pos, vel, S = oracle.lin_clusters(cl)
    npts = int(np.ceil(float(cl["tmax_myr"].max()) / 0.5)) + 1
    tn, toff, tpos, tvel, tS = oracle.lin_targets(st, float(cl["tmax_myr"].max()))
    assert int(tn.max()) == npts and int(n.max()) <= npts
    ct = np.ascontiguousarray(_aos(n, off, pos, vel, S, npts))
    stt = np.ascontiguousarray(_aos(tn, toff, tpos, tvel, tS, npts))
    R2 = np.ascontiguousarray((2.0 * cl["r50_pc"] / 1000.0) ** 2)
    cn = np.ascontiguousarray(n.astype(np.int32))
    ntl = (npts + 7) // 8
    p1 = np.zeros((len(st), len(cl), ntl), np.uint8)
    p2 = np.zeros_like(p1)
    n_l2u = np.zeros(1, np.int64)
    got = harness.cc_scan(stt.ctypes.data, len(st), npts, ct.ctypes.data, cn.ctypes.data, R2.ctypes.data, len(cl),
                          npts, CHI2, p1.ctypes.data, p2.ctypes.data, n_l2u.ctypes.data)
    assert got == ntl
    q = _point_chi2(stt, ct, cn, R2)
    qt = np.full(p1.shape, np.inf)
    for j in range(ntl):
        qt[:, :, j] = q[:, :, 8 * j:8 * j + 8].min(-1)
    dropped = p2 == 0
    # every dropped tile: all its points above the cut (the direct chi2 agrees with the reference's to ~1e-12)
    worst = qt[dropped].min()
    assert worst >= CHI2 * (1 - 1e-9), f"a dropped tile holds chi2 {worst}"
    # level 2 only ever drops more
    assert not np.any((p2 == 1) & (p1 == 0))
    # the oracle's kept pairs: the tile of the minimum survives
    chi2, kb, _ = oracle.chi2_matrix(cl, st)
    ks, kc = np.nonzero(chi2 < CHI2)
    assert len(ks) > 50
    assert np.all(p2[ks, kc, kb[ks, kc] // 8] == 1)
    # not vacuous: each level drops most tiles (tiles that exist: j < ceil(n_c / 8))
    exists = np.arange(ntl)[None, None, :] < ((cn + 7) // 8)[None, :, None]
    f1 = p1[np.broadcast_to(exists, p1.shape)].mean()
    f2 = p2[np.broadcast_to(exists, p2.shape)].mean()
    assert f1 < 0.35 and f2 < 0.12, (f1, f2)
    # the interpolated covariances leave fewer tiles than one bound for the whole tile
    assert p2.sum() < 0.85 * n_l2u[0], (p2.sum(), n_l2u[0])
    # the scan's order: coarse tiles (two tiles) through both levels first, then the tiles of those left
    harness.cc_scan2.restype = C.c_int
    harness.cc_scan2.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                 C.c_double, C.c_void_p, C.c_void_p]
    p2c = np.zeros_like(p2)
    counts = np.zeros(4, np.int64)
    assert harness.cc_scan2(stt.ctypes.data, len(st), npts, ct.ctypes.data, cn.ctypes.data, R2.ctypes.data, len(cl),
                            npts, CHI2, p2c.ctypes.data, counts.ctypes.data) == ntl
    assert qt[p2c == 0].min() >= CHI2 * (1 - 1e-9)
    assert np.all(p2c <= p2)                       # only ever drops more
    assert np.all(p2c[ks, kc, kb[ks, kc] // 8] == 1)
    assert counts[0] < 0.6 * np.broadcast_to(exists, p1.shape).sum()   # coarse level-1 tests: about half
    assert counts[2] < 0.45 * counts[1]            # the coarse level 2 leaves a minority of the fine tests
    assert counts[3] == p2c.sum()
    # the record of a star slot past a batch's end passes level 1 against no cluster tile
    harness.cc_pad_passes.restype = C.c_long
    harness.cc_pad_passes.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double]
    assert harness.cc_pad_passes(ct.ctypes.data, cn.ctypes.data, R2.ctypes.data, len(cl), npts, CHI2) == 0


def test_level2_rotation_constants(harness):
    """The level-2 test rotates its dual vector by one point's frame rotation with literal cos / sin."""
    harness.cc_rot_check.restype = C.c_double
    assert harness.cc_rot_check() < 1e-15


def _sym(v):
    """3x3 from (xx xy xz yy yz zz)."""
    return np.array([[v[0], v[1], v[2]], [v[1], v[3], v[4]], [v[2], v[4], v[5]]], dtype=np.float64)


def _rot(k):
    th = -0.5 * (220.0 / 977.7922216807892) / 8.122 * k   # OMEGA_PT k (cull.h; km/s per kpc/Myr as physics.h)
    c, s = np.cos(th), np.sin(th)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


@pytest.mark.parametrize("coarse", [0, 1])
def test_interpolated_records_bound_every_point(harness, oracle, workload, coarse):
    """Each tile record's defining bounds, checked point by point in FP64 (numpy): every rotated covariance
    (plus R^2 I for a cluster) lies below (1 - w) F + w L + eta I, below lam I, and every rotated position
    within e of the stored FP32 axis a + b (k - tc); ispan = 1 / (kp - k0)."""
    harness.cc_recf.restype = C.c_int
    harness.cc_recf.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p]
    cl = workload.make_clusters(300, 7)
    st = workload.make_stars(6, 8)
    n, off, pos, vel, S = oracle.lin_clusters(cl)
    npts = int(np.ceil(float(cl["tmax_myr"].max()) / 0.5)) + 1
    tn, toff, tpos, tvel, tS = oracle.lin_targets(st, float(cl["tmax_myr"].max()))
    objs = [(t, int(k), 0.0) for t, k in zip(_aos(tn, toff, tpos, tvel, tS, npts), tn)]
    R2 = (2.0 * cl["r50_pc"] / 1000.0) ** 2
    objs += [(t, int(k), float(r)) for t, k, r in zip(_aos(n, off, pos, vel, S, npts)[:60], n[:60], R2[:60])]
    T = 16 if coarse else 8
    out = np.zeros(25, np.float32)
    kp = np.zeros(1, np.int32)
    worst_cov, worst_lam, worst_pos, tiles = np.inf, np.inf, np.inf, 0
    for trk, nk, r2 in objs:
        trk = np.ascontiguousarray(trk)
        for j in range((nk + T - 1) // T):
            k0 = harness.cc_recf(trk.ctypes.data, nk, j, coarse, r2, out.ctypes.data, kp.ctypes.data)
            kpp = int(kp[0])
            a, b = out[0:3].astype(np.float64), out[3:6].astype(np.float64)
            F, L = _sym(out[6:12].astype(np.float64)), _sym(out[12:18].astype(np.float64))
            e, eta, tp, tl, ispan, lam = (float(x) for x in (out[18], out[19], out[20], out[21], out[22], out[24]))
            tc = k0 + 0.5 * T
            assert tp == kpp - tc and tl == min(k0 + T, nk - 1) - tc
            assert ispan == (np.float32(1.0 / (kpp - k0)) if kpp > k0 else 0.0)
            for k in range(k0, kpp + 1):
                R = _rot(k)
                q = R @ _sym(trk[k, 6:12]) @ R.T + r2 * np.eye(3)
                w = (k - k0) / (kpp - k0) if kpp > k0 else 0.0
                scale = np.trace(q)
                worst_cov = min(worst_cov, np.linalg.eigvalsh((1 - w) * F + w * L + eta * np.eye(3) - q).min() / scale)
                worst_lam = min(worst_lam, (lam - np.linalg.eigvalsh(q).max()) / scale)
            for k in range(k0, min(k0 + T, nk - 1) + 1):
                p = _rot(k) @ trk[k, 0:3]
                worst_pos = min(worst_pos, e - np.linalg.norm(p - (a + b * (k - tc))))
            tiles += 1
    assert tiles > 100
    assert worst_cov >= -1e-12 and worst_lam >= -1e-12 and worst_pos >= 0.0, (worst_cov, worst_lam, worst_pos)


@pytest.mark.parametrize("chi2_max", [9.0, 100.0])
def test_two_level_scan_exact_at_other_cuts(harness, oracle, workload, chi2_max):
    """The scan's two-level order drops no tile holding a point under cuts other than 25 (other stars)."""
    harness.cc_scan2.restype = C.c_int
    harness.cc_scan2.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                 C.c_double, C.c_void_p, C.c_void_p]
    cl = workload.make_clusters(4087, 0)
    st = workload.make_stars(4, 11)
    n, off, pos, vel, S = oracle.lin_clusters(cl)
    npts = int(np.ceil(float(cl["tmax_myr"].max()) / 0.5)) + 1
    tn, toff, tpos, tvel, tS = oracle.lin_targets(st, float(cl["tmax_myr"].max()))
    ct = np.ascontiguousarray(_aos(n, off, pos, vel, S, npts))
    stt = np.ascontiguousarray(_aos(tn, toff, tpos, tvel, tS, npts))
    R2 = np.ascontiguousarray((2.0 * cl["r50_pc"] / 1000.0) ** 2)
    cn = np.ascontiguousarray(n.astype(np.int32))
    ntl = (npts + 7) // 8
    p2 = np.zeros((len(st), len(cl), ntl), np.uint8)
    counts = np.zeros(4, np.int64)
    assert harness.cc_scan2(stt.ctypes.data, len(st), npts, ct.ctypes.data, cn.ctypes.data, R2.ctypes.data, len(cl),
                            npts, chi2_max, p2.ctypes.data, counts.ctypes.data) == ntl
    q = _point_chi2(stt, ct, cn, R2)
    qt = np.full(p2.shape, np.inf)
    for j in range(ntl):
        qt[:, :, j] = q[:, :, 8 * j:8 * j + 8].min(-1)
    assert (qt < chi2_max).sum() > 100             # not vacuous
    assert qt[p2 == 0].min() >= chi2_max * (1 - 1e-9)
    assert p2.sum() < 3 * (qt < chi2_max).sum()    # and the bound is not vacuous either

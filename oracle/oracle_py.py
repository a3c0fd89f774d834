"""ORACLE -- test infrastructure only: ctypes binding of oracle/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module; the
product (phygas_amd/) never does.  Record arrays use phygas_amd.workload's CLUSTER_DTYPE /
TARGET_DTYPE (the byte layout of PhCluster / PhTarget)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from phygas_amd.workload import CLUSTER_DTYPE, TARGET_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))


def _cl(a) -> np.ndarray:
    """Cluster records in the C layout of PhCluster (numpy may pack a structured dtype on concatenate)."""
    return np.ascontiguousarray(np.atleast_1d(a), dtype=CLUSTER_DTYPE)


def _tg(a) -> np.ndarray:
    return np.ascontiguousarray(np.atleast_1d(a), dtype=TARGET_DTYPE)

ROW_DTYPE = np.dtype([("star", "<i4"), ("cluster", "<i4"), ("kb", "<i4"), ("flags", "<i4"), ("chi2", "<f8"),
                      ("sb", "<f8"), ("t", "<f8"), ("d_nom_pc", "<f8"), ("sig_pc", "<f8"), ("p_r50", "<f8"),
                      ("p_rtot", "<f8"), ("vrel", "<f8")])
RESULT_DTYPE = np.dtype({"names": ["n", "d05", "d50", "d95", "t05", "t50", "t95", "v05", "v50", "v95", "f_r50",
                                   "f_rtot"],
                         "formats": ["<i4"] + ["<f8"] * 11,
                         "offsets": [0] + [8 + 8 * i for i in range(11)], "itemsize": 96})

_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_L = C.c_long


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class Oracle:
    def __init__(self, path: str | None = None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.lib = C.CDLL(path)
        sig = {
            "orc_accel": (None, [_D, _D, _D, _P]),
            "orc_energy": (_D, [_D, _D, _D]),
            "orc_body_step": (None, [_P, _D]),
            "orc_astrometry_to_gc": (None, [_D] * 6 + [_P]),
            "orc_gc_to_astrometry": (None, [_P, _P]),
            "orc_quad_inv": (_D, [_P, _P]),
            "orc_seg_min": (_D, [_P, _P, _P]),
            "orc_params": (None, [_P]),
            "orc_lin_clusters": (_L, [_P, _I, _I, _I, _P, _P, _P, _P, _P, _L]),
            "orc_lin_targets": (_L, [_P, _I, _D, _I, _I, _P, _P, _P, _P, _P, _L]),
            "orc_lin_init_targets": (None, [_P, _I, _I, _P, _P]),
            "orc_screen_run": (_L, [_P, _I, _P, _I, _I, _I, _D, _I, _D, _I, _I, _P, _L]),
            "orc_chi2_matrix": (None, [_P, _I, _P, _I, _I, _D, _I, _P, _P, _P]),
            "orc_pair_run": (None, [_P, _P, _I, _P, _P, _I, _D, _I, _P]),
            "orc_pair_clones": (None, [_P, _P, _P, _I, _I, _P, _P, _P]),
            "ph_seed": (C.c_uint64, [C.c_char_p, C.c_char_p]),
            "ph_pair_draw": (_I, [_P, _P, _I, _D, C.c_uint64, _P, _P, _P]),
            "ph_pair_reduce": (None, [_P, _I, _P, _P, _P, _P]),
            "mc_rng_next_u64": (C.c_uint64, [_P]),
            "mc_gaussian": (_D, [_P]),
            "ph_init": (None, []),
            "orc_prefilter_mask": (None, [_P, _I, _P, _I, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L.ph_init()

    # ---- scalar helpers
    def accel(self, x, y, z):
        o = np.zeros(3)
        self.lib.orc_accel(x, y, z, _ptr(o))
        return o

    def energy(self, x, y, z):
        return self.lib.orc_energy(x, y, z)

    def body_step(self, b, dt):
        b = np.ascontiguousarray(b, dtype=np.float64).copy()
        self.lib.orc_body_step(_ptr(b), dt)
        return b

    def astrometry_to_gc(self, ra, dec, plx, pmra, pmdec, rv):
        o = np.zeros(6)
        self.lib.orc_astrometry_to_gc(ra, dec, plx, pmra, pmdec, rv, _ptr(o))
        return o

    def gc_to_astrometry(self, ps):
        ps = np.ascontiguousarray(ps, dtype=np.float64)
        o = np.zeros(6)
        self.lib.orc_gc_to_astrometry(_ptr(ps), _ptr(o))
        return o

    def quad_inv(self, M6, m3):
        M6 = np.ascontiguousarray(M6, dtype=np.float64)
        m3 = np.ascontiguousarray(m3, dtype=np.float64)
        return self.lib.orc_quad_inv(_ptr(M6), _ptr(m3))

    def seg_min(self, r0, r1):
        r0 = np.ascontiguousarray(r0, dtype=np.float64)
        r1 = np.ascontiguousarray(r1, dtype=np.float64)
        s = np.zeros(1)
        d = self.lib.orc_seg_min(_ptr(r0), _ptr(r1), _ptr(s))
        return d, float(s[0])

    def params(self):
        p = np.zeros(9)
        self.lib.orc_params(_ptr(p))
        return p

    def seed(self, a: str | bytes, b: str | bytes) -> int:
        a = a.encode() if isinstance(a, str) else a
        b = b.encode() if isinstance(b, str) else b
        return int(self.lib.ph_seed(a, b))

    def rng_u64(self, seed, n):
        st = np.array([seed], dtype=np.uint64)
        return np.array([self.lib.mc_rng_next_u64(_ptr(st)) for _ in range(n)], dtype=np.uint64)

    def gaussians(self, seed, n):
        st = np.array([seed], dtype=np.uint64)
        return np.array([self.lib.mc_gaussian(_ptr(st)) for _ in range(n)])

    # ---- tracks
    def _lin(self, fn, recs, *args, nthreads=8):
        cnt = len(recs)
        n = np.zeros(cnt, np.int32)
        off = np.zeros(cnt, np.int64)
        cap = 202 * max(cnt, 1)
        pos, vel, S = np.zeros((cap, 3)), np.zeros((cap, 3)), np.zeros((cap, 6))
        tot = fn(_ptr(recs), cnt, *args, nthreads, _ptr(n), _ptr(off), _ptr(pos), _ptr(vel), _ptr(S), cap)
        if tot < 0:
            raise RuntimeError("track buffer too small")
        return n, off, pos[:tot], vel[:tot], S[:tot]

    def lin_clusters(self, cl, n_samp=64, nthreads=8):
        return self._lin(self.lib.orc_lin_clusters, _cl(cl), n_samp, nthreads=nthreads)

    def lin_targets(self, tg, tmax, n_samp=64, nthreads=8):
        return self._lin(self.lib.orc_lin_targets, _tg(tg), C.c_double(tmax), n_samp, nthreads=nthreads)

    def lin_init_targets(self, tg, n_samp=64):
        tg = _tg(tg)
        ps0, sig = np.zeros((len(tg), 6)), np.zeros((len(tg), 36))
        self.lib.orc_lin_init_targets(_ptr(tg), len(tg), n_samp, _ptr(ps0), _ptr(sig))
        return ps0, sig

    # ---- screening
    def screen(self, cl, tg, chi2_max=25.0, n_samp=64, r_kernel=2.0, use_cut=True, nthreads=8, t0=0, t1=None):
        cl, tg = _cl(cl), _tg(tg)
        t1 = len(tg) if t1 is None else t1
        cap = 1 << 16
        while True:
            rows = np.zeros(cap, ROW_DTYPE)
            n = self.lib.orc_screen_run(_ptr(cl), len(cl), _ptr(tg), len(tg), t0, t1, chi2_max, n_samp, r_kernel,
                                        int(use_cut), nthreads, _ptr(rows), cap)
            if n >= 0:
                return rows[:n]
            cap = -n

    def chi2_matrix(self, cl, tg, n_samp=64, r_kernel=2.0, nthreads=8):
        cl, tg = _cl(cl), _tg(tg)
        chi2 = np.zeros((len(tg), len(cl)))
        kb = np.zeros((len(tg), len(cl)), np.int32)
        sb = np.zeros((len(tg), len(cl)))
        self.lib.orc_chi2_matrix(_ptr(cl), len(cl), _ptr(tg), len(tg), n_samp, r_kernel, nthreads, _ptr(chi2),
                                 _ptr(kb), _ptr(sb))
        return chi2, kb, sb

    # ---- Monte Carlo pairs
    def pair_draw(self, t, c, n_clones, dt_myr, seed):
        t, c = _tg(t)[:1], _cl(c)[:1]
        ps = np.zeros((max(n_clones, 1), 12))
        steps = np.zeros(1, np.int32)
        dt = np.zeros(1)
        m = self.lib.ph_pair_draw(_ptr(t), _ptr(c), n_clones, dt_myr, seed, _ptr(ps), _ptr(steps), _ptr(dt))
        return ps[:m], int(steps[0]), float(dt[0])

    def pair_clones(self, ps, steps, dts, nthreads=8):
        ps = np.ascontiguousarray(ps, dtype=np.float64)
        n = len(ps)
        steps = np.ascontiguousarray(np.broadcast_to(steps, (n,)), dtype=np.int32)
        dts = np.ascontiguousarray(np.broadcast_to(dts, (n,)), dtype=np.float64)
        b, t, v = np.zeros(n), np.zeros(n), np.zeros(n)
        self.lib.orc_pair_clones(_ptr(ps), _ptr(steps), _ptr(dts), n, nthreads, _ptr(b), _ptr(t), _ptr(v))
        return b, t, v

    def pair_reduce(self, c, best, tbest, vbest):
        c = _cl(c)[:1]
        best, tbest, vbest = (np.ascontiguousarray(a, dtype=np.float64) for a in (best, tbest, vbest))
        r = np.zeros(1, RESULT_DTYPE)
        self.lib.ph_pair_reduce(_ptr(c), len(best), _ptr(best), _ptr(tbest), _ptr(vbest), _ptr(r))
        return r[0]

    def prefilter_mask(self, tg, cl):
        tg, cl = _tg(tg), _cl(cl)
        m = np.zeros((len(tg), len(cl)), np.uint8)
        self.lib.orc_prefilter_mask(_ptr(tg), len(tg), _ptr(cl), len(cl), _ptr(m))
        return m.astype(bool)

    def pair_run(self, tg, cl, ti, ci, n_clones, dt_myr=0.05, nthreads=8):
        tg, cl = _tg(tg), _cl(cl)
        ti = np.ascontiguousarray(ti, dtype=np.int32)
        ci = np.ascontiguousarray(ci, dtype=np.int32)
        res = np.zeros(len(ti), RESULT_DTYPE)
        self.lib.orc_pair_run(_ptr(tg), _ptr(cl), len(ti), _ptr(ti), _ptr(ci), n_clones, dt_myr, nthreads, _ptr(res))
        return res

"""ctypes binding of the HIP library (libphygas_hip.so, C ABI of include/phygas_hip.h).

The host-side mirror of the reference's interfaces for this path:

* ``Engine`` -- the batched engine (phx_*): resident cluster tables, star batches screened on the GPU in one
  pass, Monte Carlo pairs with device quantiles.  Method names follow the reference operations they replace
  (``screen`` = the phygas_screen loop, src/phygas_screen.c:48-67; ``pair`` = ph_pair, src/phygas.c:240-249).
* ``pg_*`` -- the reference's GPU backend ABI (src/phygas_gpu.h:16-31), synchronous.

There is no CPU fallback: loading fails with an error when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .workload import CLUSTER_DTYPE, TARGET_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphygas_hip.so")

PLAIN, FAST, FAST64 = 0, 1, 2
DRAW_ON_DEVICE = 1

ROW_DTYPE = np.dtype([("star", "<i4"), ("cluster", "<i4"), ("kb", "<i4"), ("flags", "<i4"), ("chi2", "<f8"),
                      ("sb", "<f8"), ("t", "<f8"), ("d_nom_pc", "<f8"), ("sig_pc", "<f8"), ("p_r50", "<f8"),
                      ("p_rtot", "<f8"), ("vrel", "<f8")])
RESULT_DTYPE = np.dtype({"names": ["n", "d05", "d50", "d95", "t05", "t50", "t95", "v05", "v50", "v95", "f_r50",
                                   "f_rtot"],
                         "formats": ["<i4"] + ["<f8"] * 11,
                         "offsets": [0] + [8 + 8 * i for i in range(11)], "itemsize": 96})

_P, _D, _I, _L = C.c_void_p, C.c_double, C.c_int, C.c_long
_lib = None


class PhygasError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """The loaded HIP library; raises when it is missing (build it with ``python -m phygas_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PhygasError(f"{LIB_PATH} is missing: build it with `python -m phygas_amd.build`")
    L = C.CDLL(LIB_PATH)
    sig = {
        "pg_init": (_I, [_I, _I]), "pg_alloc": (_P, [C.c_size_t]), "pg_put": (_I, [_P, _P, C.c_size_t]),
        "pg_get": (_I, [_P, _P, C.c_size_t]), "pg_free": (None, [_P]),
        "pg_lin": (_I, [_P, _P, _P, _P, _P, _D, _D, _D, _I, _I, _I, _I, _I]),
        "pg_chi2": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _I, _P, _P, _P]),
        "pg_finish": (_I, [_P, _P, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P, _I, _I, _P]),
        "pg_pair": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I]),
        "phx_create": (_I, [_I, C.POINTER(_P)]), "phx_destroy": (None, [_P]),
        "phx_last_error": (C.c_char_p, []), "phx_sync": (_I, [_P]), "phx_stream": (_P, [_P]),
        "phx_load_clusters": (_I, [_P, _P, _I, _I, _D]), "phx_load_stars": (_I, [_P, _P, _I, _I]),
        "phx_clear_stars": (_I, [_P]), "phx_screen": (_I, [_P, _I, _D, _I, _D]),
        "phx_rows": (_I, [_P, _P, _L, C.POINTER(_L)]), "phx_kept": (_I, [_P, C.POINTER(_L)]),
        "phx_candidates": (_I, [_P, C.POINTER(_L)]),
        "phx_autotune": (_I, [_P, _I, _D, _I, _D, _P, _P]),
        "phx_pass_stats": (_I, [_P, _P, _I]),
        "phx_wu_run": (_I, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, _P, C.POINTER(_L)]),
        "phx_wu_resume": (_I, [C.c_char_p, C.c_char_p, _L, C.POINTER(_L), C.POINTER(_L)]),
        "phx_wu_defaults": (None, [_P]),
        "phx_stage_times": (_I, [_P, _P, _P, _I]), "phx_set_timing": (None, [_P, _I]),
        "phx_cluster_tracks": (_I, [_P, _P, _P, _P, _P, _I]), "phx_star_tracks": (_I, [_P, _I, _I, _P, _I]),
        "phx_chi2_matrix": (_I, [_P, _P, _P, _P]), "phx_npts": (_I, [_P]),
        "phx_pair_clones": (_I, [_P, _I, _P, _P, _P, _I, _P, _P, _P]),
        "phx_pair": (_I, [_P, _I, _P, _P, _I, _P, _P, _I, _D, _P]),
        "phx_seed": (C.c_uint64, [C.c_char_p, C.c_char_p]),
        "phx_pair_draw": (_I, [_P, _P, _I, _D, C.c_uint64, _P, _P, _P]),
        "phx_check_finish_screen": (_I, [_P, C.c_uint64, _L, _P]),
        "phx_check_crmath": (_I, [_I, _I, _P, _P, _L, _P]),
        "phx_pair_ex": (_I, [_P, _I, _I, _P, _P, _I, _P, _P, _I, _D, _P]),
        "phx_pair_loaded": (_I, [_P, _I, _I, _P, _P, _I, _D, _P]),
        "phx_pair_noalloc": (_I, [_I, _P, _P, _I, _P, _P, _I, _D, _P, _P]),
        "phx_pair_noalloc_limit": (C.c_size_t, []),
        "phx_pair_graph_stats": (_I, [_P, C.POINTER(_L), C.POINTER(_L)]),
        "phx_set_pair_bucket": (_I, [_P, _I]),
        "phx_set_pair_lanes": (_I, [_P, _I]),
        "phx_pair_lanes_for": (_I, [_P, _I, _L]),
        "phx_autotune_pairs": (_I, [_P, _I, _I, _P, _P, _I, _D, C.POINTER(_I), C.POINTER(_D)]),
        "phx_set_graphs": (None, [_P, _I]), "phx_graph_stats": (_I, [_P, C.POINTER(_L), C.POINTER(_L)]),
        "phx_set_split": (_I, [_P, _I]), "phx_set_track_offsets": (_I, [_P, _I]),
        "phx_set_grids": (_I, [_P, _P]),
        "phx_autotune_ex": (_I, [_P, _I, _D, _I, _D, _P]),
        "phx_set_tuning": (_I, [_P, _P]), "phx_get_tuning": (_I, [_P, _P]),
        "phx_prefilter": (_I, [_P, _P, _I, _P, _I, _P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().phx_last_error()
        raise PhygasError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def seed(a: str | bytes, b: str | bytes) -> int:
    """ph_seed (reference: src/phygas.c:452-458)."""
    a = a.encode() if isinstance(a, str) else a
    b = b.encode() if isinstance(b, str) else b
    return int(lib().phx_seed(a, b))


def pair_draw(t, c, n_clones: int, dt_myr: float, seed_: int):
    """ph_pair_draw on the host (reference: src/phygas.c:182-198): (ps[m][12], steps, dt)."""
    t = np.ascontiguousarray(np.atleast_1d(t)[:1], dtype=TARGET_DTYPE)
    c = np.ascontiguousarray(np.atleast_1d(c)[:1], dtype=CLUSTER_DTYPE)
    ps = np.zeros((max(n_clones, 1), 12))
    steps = np.zeros(1, np.int32)
    dt = np.zeros(1)
    m = lib().phx_pair_draw(_p(t), _p(c), n_clones, dt_myr, seed_, _p(ps), _p(steps), _p(dt))
    return ps[:m], int(steps[0]), float(dt[0])


def pair_noalloc(targets, clusters, ti, ci, n_clones: int, dt_myr: float = 0.05, form: int = PLAIN,
                 stream: int | None = None) -> np.ndarray:
    """The allocation-free Monte Carlo batch (phx_pair_noalloc): device draws on the current device, buffers
    from the library's static scratch array."""
    targets = np.ascontiguousarray(np.atleast_1d(targets), dtype=TARGET_DTYPE)
    clusters = np.ascontiguousarray(np.atleast_1d(clusters), dtype=CLUSTER_DTYPE)
    ti = np.ascontiguousarray(ti, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    res = np.zeros(len(ti), RESULT_DTYPE)
    _check(lib().phx_pair_noalloc(form, _p(targets), _p(clusters), len(ti), _p(ti), _p(ci), n_clones, dt_myr, _p(res),
                                  stream), "phx_pair_noalloc")
    return res


CRMATH_FUNCS = {"pow": 0, "pow15": 1, "hypot": 2, "sin": 3, "cos": 4}


def crmath(fn: str, x, y=None, device: bool = False) -> np.ndarray:
    """The library's correctly rounded pow / x^1.5 / hypot / sin / cos (the device side of the plain path's
    force and coordinates), evaluated on the device or by the host compilation of the same source."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float64)
    out = np.zeros_like(x)
    _check(lib().phx_check_crmath(CRMATH_FUNCS[fn], int(device), _p(x), _p(y), x.size, _p(out)), "phx_check_crmath")
    return out


class Tuning(C.Structure):
    """PhxTuning of include/phygas_hip.h."""
    _fields_ = [("scan_blocks", C.c_int), ("items_blocks", C.c_int), ("reduce_blocks", C.c_int),
                ("finish_blocks", C.c_int), ("stars_per_task", C.c_int), ("coarse_tiles", C.c_int),
                ("sub_passes", C.c_int), ("last_share_pct", C.c_int), ("track_offsets", C.c_int),
                ("pass_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class WuOptions(C.Structure):
    """PhxWuOptions of include/phygas_hip.h."""
    _fields_ = [("form", C.c_int), ("batch_targets", C.c_int), ("max_batches", C.c_int), ("device", C.c_int),
                ("draw_on_device", C.c_int), ("pair_form", C.c_int)]


WU_DONE, WU_CKPT_UNREADABLE, WU_STOPPED = 0, 4, 6


def wu_defaults() -> WuOptions:
    """The library's default work-unit options (phx_wu_defaults)."""
    o = WuOptions()
    lib().phx_wu_defaults(C.byref(o))
    return o


def wu_run(in_path: str, clusters_path: str, out_path: str, ckpt_path: str, form: int | None = None,
           batch_targets: int | None = None, max_batches: int = 0, device: int = 0, draw_on_device: bool = False,
           pair_form: int | None = None) -> tuple[int, int]:
    """Run a work unit (phx_wu_run): returns (status, targets done); status WU_DONE, WU_STOPPED (after
    max_batches) or WU_CKPT_UNREADABLE (the checkpoint exists but cannot be read).  Any other code is an error
    and raises.  Options left at None take the library's defaults (phx_wu_defaults)."""
    L = lib()
    o = wu_defaults()
    if form is not None:
        o.form = form
    if pair_form is not None:
        o.pair_form = pair_form
    if batch_targets is not None:
        o.batch_targets = batch_targets
    o.max_batches, o.device, o.draw_on_device = max_batches, device, int(draw_on_device)
    done = _L(0)
    rc = L.phx_wu_run(in_path.encode(), clusters_path.encode(), out_path.encode(), ckpt_path.encode(),
                      C.addressof(o), C.byref(done))
    if rc not in (WU_DONE, WU_STOPPED, WU_CKPT_UNREADABLE):
        _check(rc, "phx_wu_run")
    return rc, done.value


def wu_resume(out_path: str, ckpt_path: str, nt: int) -> tuple[int, int, int]:
    """The resume step alone (phx_wu_resume): (status, targets done, lines kept)."""
    L = lib()
    t, n = _L(0), _L(0)
    rc = L.phx_wu_resume(out_path.encode(), ckpt_path.encode(), nt, C.byref(t), C.byref(n))
    if rc not in (0, 4):
        _check(rc, "phx_wu_resume")
    return rc, t.value, n.value


class Engine:
    """A device context with resident cluster tables (phx_* ABI)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = _P()
        _check(self._L.phx_create(device, C.byref(h)), "phx_create")
        self.h = h
        self.nc = 0
        self.ns = 0
        self.clusters = None

    def close(self):
        if getattr(self, "h", None):
            self._L.phx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- tables and batches
    def load_clusters(self, cl: np.ndarray, n_samp: int = 64, r_kernel: float = 2.0):
        cl = np.ascontiguousarray(cl, dtype=CLUSTER_DTYPE)
        _check(self._L.phx_load_clusters(self.h, _p(cl), len(cl), n_samp, r_kernel), "phx_load_clusters")
        self.nc = len(cl)
        self.clusters = cl

    def load_stars(self, st: np.ndarray, group: int = 0):
        st = np.ascontiguousarray(st, dtype=TARGET_DTYPE)
        _check(self._L.phx_load_stars(self.h, _p(st), len(st), group), "phx_load_stars")
        self.ns += len(st)

    def clear_stars(self):
        _check(self._L.phx_clear_stars(self.h), "phx_clear_stars")
        self.ns = 0

    @property
    def npts(self) -> int:
        return self._L.phx_npts(self.h)

    # ---- screening
    def screen_async(self, form: int = PLAIN, chi2_max: float = 25.0, n_samp: int = 64, p_min: float = 0.0):
        _check(self._L.phx_screen(self.h, form, chi2_max, n_samp, p_min), "phx_screen")

    def rows(self, cap: int | None = None) -> np.ndarray:
        """Rows of the last pass in (group, star, cluster) order; cap: only the first cap of them."""
        n = _L(0)
        _check(self._L.phx_rows(self.h, None, 0, C.byref(n)), "phx_rows")
        m = n.value if cap is None else min(cap, n.value)
        out = np.zeros(m, ROW_DTYPE)
        if m:
            _check(self._L.phx_rows(self.h, _p(out), m, C.byref(n)), "phx_rows")
        return out

    def screen(self, form: int = PLAIN, chi2_max: float = 25.0, n_samp: int = 64, p_min: float = 0.0) -> np.ndarray:
        self.screen_async(form, chi2_max, n_samp, p_min)
        return self.rows()

    def kept(self) -> int:
        n = _L(0)
        _check(self._L.phx_kept(self.h, C.byref(n)), "phx_kept")
        return n.value

    def candidates(self) -> int:
        """Pairs the last pass evaluated point by point (all pairs for PLAIN; the tile scan's survivors for the
        fast forms)."""
        n = _L(0)
        _check(self._L.phx_candidates(self.h, C.byref(n)), "phx_candidates")
        return n.value

    def pass_stats(self) -> dict:
        """Work counts of the last pass (see phx_pass_stats)."""
        st = np.zeros(7, np.int64)
        _check(self._L.phx_pass_stats(self.h, st.ctypes.data, 7), "phx_pass_stats")
        return dict(zip(("l1_tests", "l2_tests", "candidates", "items", "kept", "rows", "points"), (int(x) for x in st)))

    def autotune(self, form: int = FAST, chi2_max: float = 25.0, n_samp: int = 64, p_min: float = 0.0) -> dict:
        """Pick the fast pass's launch grids by timing whole passes on the loaded batch; returns the choices and
        the pass time of each."""
        grids = np.zeros(4, np.int32)
        ms = np.zeros(4)
        _check(self._L.phx_autotune(self.h, form, chi2_max, n_samp, p_min, grids.ctypes.data, ms.ctypes.data),
               "phx_autotune")
        return {f"{k}_blocks": int(g) for k, g in zip(("scan", "items", "reduce", "finish"), grids)} | \
            {f"{k}_ms": float(t) for k, t in zip(("scan", "items", "reduce", "finish"), ms)}

    def autotune_ex(self, form: int = FAST, chi2_max: float = 25.0, n_samp: int = 64, p_min: float = 0.0) -> dict:
        """Every tunable setting of the fast pass (sub-passes and the last one's share, the scan's first-level
        time tile, stars per task, the four launch grids) chosen by timing whole passes on the loaded batch; the
        choices stay set."""
        t = Tuning()
        _check(self._L.phx_autotune_ex(self.h, form, chi2_max, n_samp, p_min, C.byref(t)), "phx_autotune_ex")
        return t.as_dict()

    def tuning(self) -> dict:
        t = Tuning()
        _check(self._L.phx_get_tuning(self.h, C.byref(t)), "phx_get_tuning")
        return t.as_dict()

    def set_tuning(self, scan: int = 0, items: int = 0, reduce: int = 0, finish: int = 0, stars_per_task: int = 16,
                   coarse_tiles: int = 1, sub_passes: int = 0, last_share_pct: int = 0, track_offsets: int = 0):
        t = Tuning(scan, items, reduce, finish, stars_per_task, coarse_tiles, sub_passes, last_share_pct,
                   track_offsets, 0.0)
        _check(self._L.phx_set_tuning(self.h, C.byref(t)), "phx_set_tuning")

    def set_grids(self, scan: int = 0, items: int = 0, reduce: int = 0, finish: int = 0):
        """Launch grids of the fast pass's scan, items, reduce and finish (0: the library's default)."""
        g = np.array([scan, items, reduce, finish], np.int32)
        _check(self._L.phx_set_grids(self.h, g.ctypes.data), "phx_set_grids")

    def sync(self):
        _check(self._L.phx_sync(self.h), "phx_sync")

    def set_timing(self, on: bool = True):
        self._L.phx_set_timing(self.h, int(on))

    def set_graphs(self, on: bool = True):
        self._L.phx_set_graphs(self.h, int(on))

    def set_split(self, parts: int):
        """Sub-passes (1..4) of a fast pass over a large batch, each on its own stream."""
        _check(self._L.phx_set_split(self.h, int(parts)), "phx_set_split")

    def set_track_offsets(self, offsets: int):
        """Offsets per lane of the fast tracks' offset wavefronts (1, 2, 3; 0 = the default, 2)."""
        _check(self._L.phx_set_track_offsets(self.h, int(offsets)), "phx_set_track_offsets")

    def graph_stats(self) -> tuple[int, int]:
        c, r = _L(0), _L(0)
        self._L.phx_graph_stats(self.h, C.byref(c), C.byref(r))
        return c.value, r.value

    def stage_times(self) -> dict[str, float]:
        ms = np.zeros(16)
        names = (C.c_char_p * 16)()
        n = self._L.phx_stage_times(self.h, _p(ms), C.cast(names, _P), 16)
        if n < 0:
            raise PhygasError("phx_stage_times failed")
        return {names[i].decode(): float(ms[i]) for i in range(n)}

    def stream(self) -> int:
        return self._L.phx_stream(self.h)

    # ---- parity access
    def cluster_tracks(self):
        npts = self.npts
        n = np.zeros(self.nc, np.int32)
        pos = np.zeros((self.nc, npts, 3))
        vel = np.zeros((self.nc, npts, 3))
        S = np.zeros((self.nc, npts, 6))
        _check(self._L.phx_cluster_tracks(self.h, _p(n), _p(pos), _p(vel), _p(S), npts), "phx_cluster_tracks")
        return n, pos, vel, S

    def star_tracks(self, first: int = 0, count: int | None = None):
        count = self.ns - first if count is None else count
        trk = np.zeros((count, self.npts, 12))
        _check(self._L.phx_star_tracks(self.h, first, count, _p(trk), self.npts), "phx_star_tracks")
        return trk

    def chi2_matrix(self):
        chi2 = np.zeros((self.ns, self.nc))
        kb = np.zeros((self.ns, self.nc), np.int32)
        sb = np.zeros((self.ns, self.nc))
        _check(self._L.phx_chi2_matrix(self.h, _p(chi2), _p(kb), _p(sb)), "phx_chi2_matrix")
        return chi2, kb, sb

    def check_finish_screen(self, seed: int = 12345, n: int = 1 << 22) -> float:
        w = np.zeros(1)
        _check(self._L.phx_check_finish_screen(self.h, seed, n, _p(w)), "phx_check_finish_screen")
        return float(w[0])

    # ---- Monte Carlo pairs
    def prefilter(self, targets, clusters) -> np.ndarray:
        """ph_prefilter of every (target, cluster) as phygas_cli runs it (reference: src/phygas_cli.c:103-117)."""
        targets = np.ascontiguousarray(np.atleast_1d(targets), dtype=TARGET_DTYPE)
        clusters = np.ascontiguousarray(np.atleast_1d(clusters), dtype=CLUSTER_DTYPE)
        mask = np.zeros((len(targets), len(clusters)), np.uint8)
        _check(self._L.phx_prefilter(self.h, _p(targets), len(targets), _p(clusters), len(clusters), _p(mask)),
               "phx_prefilter")
        return mask.astype(bool)

    def pair_clones(self, ps, steps, dts, form: int = PLAIN):
        ps = np.ascontiguousarray(ps, dtype=np.float64)
        n = len(ps)
        steps = np.ascontiguousarray(np.broadcast_to(steps, (n,)), dtype=np.int32)
        dts = np.ascontiguousarray(np.broadcast_to(dts, (n,)), dtype=np.float64)
        b, t, v = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(self._L.phx_pair_clones(self.h, form, _p(ps), _p(steps), _p(dts), n, _p(b), _p(t), _p(v)),
               "phx_pair_clones")
        return b, t, v

    def pair(self, targets, clusters, ti, ci, n_clones: int, dt_myr: float = 0.05, form: int = PLAIN,
             device_draws: bool = False) -> np.ndarray:
        """ph_pair for every (targets[ti[i]], clusters[ci[i]]) (reference: src/phygas.c:240-249)."""
        targets = np.ascontiguousarray(np.atleast_1d(targets), dtype=TARGET_DTYPE)
        clusters = np.ascontiguousarray(np.atleast_1d(clusters), dtype=CLUSTER_DTYPE)
        ti = np.ascontiguousarray(ti, dtype=np.int32)
        ci = np.ascontiguousarray(ci, dtype=np.int32)
        res = np.zeros(len(ti), RESULT_DTYPE)
        _check(self._L.phx_pair_ex(self.h, form, DRAW_ON_DEVICE if device_draws else 0, _p(targets), _p(clusters),
                                   len(ti), _p(ti), _p(ci), n_clones, dt_myr, _p(res)), "phx_pair_ex")
        return res

    def pair_loaded(self, star, cluster, n_clones: int, dt_myr: float = 0.05, form: int = PLAIN) -> np.ndarray:
        """Monte Carlo pairs of the loaded stars and resident clusters with device draws (phx_pair_loaded)."""
        star = np.ascontiguousarray(star, dtype=np.int32)
        cluster = np.ascontiguousarray(cluster, dtype=np.int32)
        res = np.zeros(len(star), RESULT_DTYPE)
        _check(self._L.phx_pair_loaded(self.h, form, len(star), _p(star), _p(cluster), n_clones, dt_myr, _p(res)),
               "phx_pair_loaded")
        return res

    def set_pair_bucket(self, chunks: int):
        _check(self._L.phx_set_pair_bucket(self.h, chunks), "phx_set_pair_bucket")

    def set_pair_lanes(self, lanes: int):
        """Lanes per clone of the clone kernel: 1, 2, or 0 (chosen from the batch)."""
        _check(self._L.phx_set_pair_lanes(self.h, lanes), "phx_set_pair_lanes")

    def pair_lanes(self, n_clones: int, form: int = PLAIN) -> int:
        """Lanes per clone a batch of n_clones clone slots of this form runs with under the current setting."""
        return int(self._L.phx_pair_lanes_for(self.h, form, n_clones))

    def autotune_pairs(self, star, cluster, n_clones: int, dt_myr: float = 0.05, form: int = FAST) -> dict:
        """Lanes per clone and bucket width of the clone kernel chosen by timing this batch (phx_autotune_pairs)."""
        star = np.ascontiguousarray(star, dtype=np.int32)
        cluster = np.ascontiguousarray(cluster, dtype=np.int32)
        b, ms = _I(0), _D(0)
        _check(self._L.phx_autotune_pairs(self.h, form, len(star), _p(star), _p(cluster), n_clones, dt_myr,
                                          C.byref(b), C.byref(ms)), "phx_autotune_pairs")
        return {"bucket": b.value, "lanes": self.pair_lanes(len(star) * n_clones, form), "ms": ms.value}

    def pair_graph_stats(self) -> tuple[int, int]:
        c, r = _L(0), _L(0)
        _check(self._L.phx_pair_graph_stats(self.h, C.byref(c), C.byref(r)), "phx_pair_graph_stats")
        return c.value, r.value

"""Serving loop binding (phx_server_* C ABI) and a seeded Poisson request mix.

``Server`` keeps the cluster tables resident on the GPU in every screening slot and batches requests
continuously: screening chunks (star batches) and pulsar Monte Carlo jobs, admitted by HBM footprint.
``request_mix`` builds the mix of the serving work unit: 70 % screening chunks of 256..20,000 stars
(log-uniform), 30 % pulsar jobs of 100..600 clusters x 1,000 clones, Poisson arrivals."""
from __future__ import annotations

import ctypes as C
import math
import threading
import time

import numpy as np

from . import hip
from . import workload as wl
from .workload import CLUSTER_DTYPE, TARGET_DTYPE

_P, _D, _I, _L = C.c_void_p, C.c_double, C.c_int, C.c_long


class ServeConfig(C.Structure):
    _fields_ = [("form", C.c_int), ("screen_slots", C.c_int), ("pair_slots", C.c_int), ("hbm_budget", C.c_size_t),
                ("max_batch_stars", C.c_long), ("max_batch_clones", C.c_long), ("chi2_max", C.c_double),
                ("p_min", C.c_double), ("r_kernel", C.c_double), ("n_samp", C.c_int)]


def _lib():
    L = hip.lib()
    if not hasattr(L, "_serve_bound"):
        sig = {
            "phx_serve_defaults": (None, [C.POINTER(ServeConfig)]),
            "phx_server_create": (_I, [_I, _P, _I, C.POINTER(ServeConfig), C.POINTER(_P)]),
            "phx_server_destroy": (None, [_P]),
            "phx_server_submit_screen": (_L, [_P, _P, _I]),
            "phx_server_submit_pair": (_L, [_P, _P, _P, _I, _I, _D]),
            "phx_server_wait": (_I, [_P, _L, _D]),
            "phx_server_screen_rows": (_I, [_P, _L, _P, _L, C.POINTER(_L)]),
            "phx_server_pair_results": (_I, [_P, _L, _P, _I]),
            "phx_server_info": (_I, [_P, _L, _P, C.POINTER(_I)]),
            "phx_server_release": (_I, [_P, _L]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._serve_bound = True
    return L


class Server:
    def __init__(self, clusters: np.ndarray, form: int = hip.FAST, device: int = 0, **kw):
        self._L = _lib()
        cfg = ServeConfig()
        self._L.phx_serve_defaults(C.byref(cfg))
        cfg.form = form
        for k, v in kw.items():
            setattr(cfg, k, v)
        self.cfg = cfg
        self.clusters = np.ascontiguousarray(clusters, dtype=CLUSTER_DTYPE)
        h = _P()
        hip._check(self._L.phx_server_create(device, self.clusters.ctypes.data, len(self.clusters), C.byref(cfg),
                                             C.byref(h)), "phx_server_create")
        self.h = h
        self._sizes = {}

    def close(self):
        if getattr(self, "h", None):
            self._L.phx_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit_screen(self, stars: np.ndarray) -> int:
        stars = np.ascontiguousarray(stars, dtype=TARGET_DTYPE)
        rid = self._L.phx_server_submit_screen(self.h, stars.ctypes.data, len(stars))
        if rid < 0:
            hip._check(int(rid), "phx_server_submit_screen")
        return rid

    def submit_pair(self, target: np.ndarray, cluster_idx, n_clones: int = 1000, dt_myr: float = 0.05) -> int:
        t = np.ascontiguousarray(np.atleast_1d(target)[:1], dtype=TARGET_DTYPE)
        ci = np.ascontiguousarray(cluster_idx, dtype=np.int32)
        rid = self._L.phx_server_submit_pair(self.h, t.ctypes.data, ci.ctypes.data, len(ci), n_clones, dt_myr)
        if rid < 0:
            hip._check(int(rid), "phx_server_submit_pair")
        self._sizes[rid] = len(ci)
        return rid

    def wait(self, rid: int, timeout: float = -1.0) -> bool:
        rc = self._L.phx_server_wait(self.h, rid, timeout)
        if rc < 0:
            hip._check(rc, "phx_server_wait")
        return rc == 0

    def rows(self, rid: int) -> np.ndarray:
        n = _L(0)
        hip._check(self._L.phx_server_screen_rows(self.h, rid, None, 0, C.byref(n)), "phx_server_screen_rows")
        out = np.zeros(n.value, hip.ROW_DTYPE)
        if n.value:
            hip._check(self._L.phx_server_screen_rows(self.h, rid, out.ctypes.data, n.value, C.byref(n)),
                       "phx_server_screen_rows")
        return out

    def pair_results(self, rid: int) -> np.ndarray:
        out = np.zeros(self._sizes[rid], hip.RESULT_DTYPE)
        hip._check(self._L.phx_server_pair_results(self.h, rid, out.ctypes.data, len(out)), "phx_server_pair_results")
        return out

    def info(self, rid: int):
        t = np.zeros(3)
        b = _I(0)
        st = self._L.phx_server_info(self.h, rid, t.ctypes.data, C.byref(b))
        return {"status": st, "arrive": t[0], "start": t[1], "done": t[2], "batch": b.value,
                "latency": t[2] - t[0], "queue": t[1] - t[0]}

    def release(self, rid: int):
        self._L.phx_server_release(self.h, rid)


def request_mix(n: int, rate_hz: float, seed: int = 5, n_clusters: int = 4087, max_stars: int = 20000):
    """The serving mix: (arrival offset s, kind, payload) with Poisson arrivals at rate_hz."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    out = []
    for i in range(n):
        if rng.uniform() < 0.7:
            ns = int(round(math.exp(rng.uniform(math.log(256), math.log(max_stars)))))
            stars = wl.make_stars(ns, seed * 100003 + i, prefix=f"Q{i:05d}S")
            out.append((float(t[i]), "screen", stars))
        else:
            k = int(rng.integers(100, 601))
            idx = np.sort(rng.choice(n_clusters, size=min(k, n_clusters), replace=False)).astype(np.int32)
            pul = wl.make_pulsars(1, seed * 100003 + i)
            pul["id"] = f"Q{i:05d}P".encode()
            out.append((float(t[i]), "pair", (pul, idx)))
    return out


def run_mix(server: Server, mix, n_clones: int = 1000, dt_myr: float = 0.05) -> dict:
    """Submit the mix at its arrival times (from a feeder thread), wait for every request, and report
    latency percentiles and throughput."""
    ids = [None] * len(mix)
    t0 = time.monotonic()

    def feed():
        for i, (ta, kind, payload) in enumerate(mix):
            d = t0 + ta - time.monotonic()
            if d > 0:
                time.sleep(d)
            if kind == "screen":
                ids[i] = server.submit_screen(payload)
            else:
                ids[i] = server.submit_pair(payload[0], payload[1], n_clones, dt_myr)

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    for rid in ids:
        server.wait(rid)
    t1 = time.monotonic()
    lat, lat_s, lat_p, lat_big, pairs, clones = [], [], [], [], 0, 0
    first = last = None
    for (ta, kind, payload), rid in zip(mix, ids):
        inf = server.info(rid)
        lat.append(inf["latency"])
        first = inf["arrive"] if first is None else min(first, inf["arrive"])
        last = inf["done"] if last is None else max(last, inf["done"])
        if kind == "screen":
            lat_s.append(inf["latency"])
            if len(payload) >= 10000:
                lat_big.append(inf["latency"])
            pairs += len(payload) * len(server.clusters)
        else:
            lat_p.append(inf["latency"])
            clones += len(payload[1]) * n_clones
        server.release(rid)
    span = last - first

    def pct(v, q):
        return float(np.percentile(v, q)) if v else None

    return {"requests": len(mix), "screen_requests": len(lat_s), "pair_requests": len(lat_p),
            "p50_s": pct(lat, 50), "p95_s": pct(lat, 95), "p99_s": pct(lat, 99),
            "screen_p50_s": pct(lat_s, 50), "screen_p99_s": pct(lat_s, 99),
            "pair_p50_s": pct(lat_p, 50), "pair_p99_s": pct(lat_p, 99),
            "large_screen_requests": len(lat_big), "large_screen_p50_s": pct(lat_big, 50),
            "large_screen_p99_s": pct(lat_big, 99),
            "span_s": span, "pairs_per_s": pairs / span if span > 0 else None,
            "clones_per_s": clones / span if span > 0 else None, "wall_s": t1 - t0}

"""Readers of the reference's text input formats into PhCluster / PhTarget record arrays.

Semantics of ph_read_clusters / ph_read_targets (reference: src/phygas.c:460-506): whitespace-separated
fields, 19 per cluster line and 14 per target line, names read as %39s, RV and its error read as text so
``nan`` works (has_rv = neither is NaN); a line that does not parse is skipped without a message."""
from __future__ import annotations

import math

import numpy as np

from .workload import CLUSTER_DTYPE, TARGET_DTYPE


def _tokens(line: str, widths: dict[int, int]) -> list[str]:
    """Split like sscanf: a %Ns conversion takes at most N characters, the rest starts the next field."""
    out: list[str] = []
    for tok in line.split():
        while True:
            w = widths.get(len(out))
            if w is not None and len(tok) > w:
                out.append(tok[:w])
                tok = tok[w:]
                continue
            out.append(tok)
            break
    return out


def _num(s: str) -> float:
    return math.nan if s == "nan" else float(s)


def read_clusters(path) -> np.ndarray:
    recs = []
    with open(path, "rb") as f:
        for raw in f:
            t = _tokens(raw.decode("latin-1"), {1: 39, 2: 3, 11: 31, 12: 31})
            if len(t) < 19:
                continue
            try:
                int(t[0])
                v = [float(x) for x in t[3:11]] + [_num(t[11]), _num(t[12])] + [float(x) for x in t[13:19]]
            except ValueError:
                continue
            r = np.zeros((), CLUSTER_DTYPE)
            r["name"], r["kind"] = t[1].encode("latin-1"), t[2][:1].encode("latin-1")
            for k, f_ in enumerate(["ra", "dec", "dist_pc", "e_dist_pc", "pmra", "e_pmra", "pmdec", "e_pmdec", "rv",
                                    "e_rv", "r50_pc", "rtot_pc", "age16", "age50", "age84", "tmax_myr"]):
                r[f_] = v[k]
            r["has_rv"] = int(not math.isnan(v[8]) and not math.isnan(v[9]))
            recs.append(r)
    return np.array(recs, CLUSTER_DTYPE) if recs else np.zeros(0, CLUSTER_DTYPE)


def read_targets(path) -> np.ndarray:
    recs = []
    with open(path, "rb") as f:
        for raw in f:
            t = _tokens(raw.decode("latin-1"), {0: 39, 1: 3, 12: 31, 13: 31})
            if len(t) < 14:
                continue
            try:
                v = [float(x) for x in t[2:12]] + [_num(t[12]), _num(t[13])]
            except ValueError:
                continue
            r = np.zeros((), TARGET_DTYPE)
            r["id"], r["kind"] = t[0].encode("latin-1"), t[1][:1].encode("latin-1")
            for k, f_ in enumerate(["ra", "dec", "plx", "e_plx", "dist_pc", "e_lndist", "pmra", "e_pmra", "pmdec",
                                    "e_pmdec", "rv", "e_rv"]):
                r[f_] = v[k]
            r["has_rv"] = int(not math.isnan(v[10]) and not math.isnan(v[11]))
            recs.append(r)
    return np.array(recs, TARGET_DTYPE) if recs else np.zeros(0, TARGET_DTYPE)

"""Output rows of the trace-back path, in the reference's four formats.

* screen CLI (reference: src/phygas_screen.c:60-61), 12 columns:
    target cluster kind chi2 t_myr d_nom_pc sig_pc p_r50 p_rtot vrel_kms age16 age84
* screen work unit (reference: src/main_phygas_boinc.cpp:165 GPU, :442 CPU), 10 columns (no ages)
* pair CLI (reference: src/phygas_cli.c:121-123), 17 columns:
    target cluster kind n d05 d50 d95 t05 t50 t95 v05 v50 v95 f_r50 f_rtot age16 age84
* pair work unit (reference: src/main_phygas_boinc.cpp:256, :453), 15 columns (no ages)

Row order: targets in input order, then clusters in file order.  Python's %-formatting rounds the
exact binary value like glibc's printf, so the text is the reference's byte for byte.
"""
from __future__ import annotations

import numpy as np


def _s(b) -> str:
    return b.decode() if isinstance(b, (bytes, np.bytes_)) else str(b)


def screen_lines(rows, targets, clusters, with_ages: bool = True) -> list[str]:
    out = []
    for r in rows:
        t, c = targets[r["star"]], clusters[r["cluster"]]
        s = "%s %s %s %.3f %.3f %.2f %.2f %.4f %.4f %.2f" % (
            _s(t["id"]), _s(c["name"]), _s(c["kind"]), r["chi2"], r["t"], r["d_nom_pc"], r["sig_pc"], r["p_r50"],
            r["p_rtot"], r["vrel"])
        if with_ages:
            s += " %.3f %.3f" % (c["age16"], c["age84"])
        out.append(s)
    return out


def pair_lines(results, target_idx, cluster_idx, targets, clusters, with_ages: bool = True) -> list[str]:
    out = []
    for r, j, i in zip(results, target_idx, cluster_idx):
        t, c = targets[j], clusters[i]
        s = "%s %s %s %d %.3f %.3f %.3f %.4f %.4f %.4f %.2f %.2f %.2f %.4f %.4f" % (
            _s(t["id"]), _s(c["name"]), _s(c["kind"]), r["n"], r["d05"], r["d50"], r["d95"], r["t05"], r["t50"],
            r["t95"], r["v05"], r["v50"], r["v95"], r["f_r50"], r["f_rtot"])
        if with_ages:
            s += " %.3f %.3f" % (c["age16"], c["age84"])
        out.append(s)
    return out

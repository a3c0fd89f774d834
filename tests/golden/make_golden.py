"""Write the golden fixtures of tests/golden/ from the compiled reference.

Run in a container that has the reference sources (the build reads them where they lie):

    make -C oracle ref && python tests/golden/make_golden.py

It writes
  inputs/     small input files in the reference's text formats (seeded synthetic records)
  unit_*.txt  the golden harness (oracle/golden_harness.c built against the reference) per section
  rows_*.txt  stdout of the reference's own phygas_screen / phygas_cli on the small inputs
Only numbers and the inputs that made them are kept: no reference source enters the repository.
"""
from __future__ import annotations

import importlib.util
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
SECTIONS = ["accel", "rng", "seed", "coords", "sample", "lininit", "lin", "chi2", "pair", "track", "read"]


def _workload():
    spec = importlib.util.spec_from_file_location("workload", os.path.join(ROOT, "phygas_amd", "workload.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def write_inputs(d: str) -> None:
    w = _workload()
    os.makedirs(d, exist_ok=True)
    cl = w.make_clusters(4087, 0)
    order = np.argsort(cl["tmax_myr"], kind="stable")
    pick = list(order[[0, 1, 2, 100, 1000, 2000, 3000, 4000, 4086]]) + list(np.where(cl["has_rv"] == 0)[0][:3])
    tg = np.concatenate([w.make_stars(6, 0)[:4], w.make_pulsars(2, 1)])
    tg[3]["rv"], tg[3]["e_rv"], tg[3]["has_rv"] = np.nan, np.nan, 0      # a star without an RV
    tg[2]["plx"], tg[2]["e_plx"] = 0.15, 0.3                              # draws that fail and retry
    w.write_clusters(os.path.join(d, "unit_clusters.txt"), cl[pick])
    w.write_targets(os.path.join(d, "unit_targets.txt"), tg)
    w.write_clusters(os.path.join(d, "clusters400.txt"), cl[:400])
    w.write_targets(os.path.join(d, "stars20.txt"), w.make_stars(20, 0))
    w.write_targets(os.path.join(d, "pulsar1.txt"), w.make_pulsars(1, 1))


def run(cmd: list[str], out: str) -> None:
    with open(out, "w") as f:
        subprocess.run(cmd, stdout=f, stderr=subprocess.DEVNULL, check=True)


def main() -> int:
    if not os.path.exists(os.path.join(REF, "golden_ref")):
        print("build the reference first: make -C oracle ref", file=sys.stderr)
        return 1
    inp = os.path.join(HERE, "inputs")
    write_inputs(inp)
    uc, ut = os.path.join(inp, "unit_clusters.txt"), os.path.join(inp, "unit_targets.txt")
    for s in SECTIONS:
        run([os.path.join(REF, "golden_ref"), s, uc, ut], os.path.join(HERE, f"unit_{s}.txt"))
    c400, s20, p1 = (os.path.join(inp, f) for f in ("clusters400.txt", "stars20.txt", "pulsar1.txt"))
    scr = os.path.join(REF, "phygas_screen")
    run([scr, c400, s20], os.path.join(HERE, "rows_screen.txt"))
    run([scr, c400, s20, "25", "64", "2", "control"], os.path.join(HERE, "rows_screen_control.txt"))
    run([scr, c400, s20, "9", "64", "0"], os.path.join(HERE, "rows_screen_rtot.txt"))
    run([os.path.join(REF, "phygas_cli"), c400, p1, "40", "0.05"], os.path.join(HERE, "rows_pair.txt"))
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Writes tests/golden/scan_counts.json: the tile tests of the scan's two-level order on 40 stars (seed 2) x
4,087 clusters (seed 0), chi2 cut 25, counted on the CPU (tests/native/cull_harness.cpp, cc_scan2, on the
oracle's FP64 tracks).  tests/test_gpu_fast.py::test_scan_test_counts_match_the_cpu_order compares the GPU
scan's own counts (phx_pass_stats) with them.

    python tests/golden/make_scan_counts.py
"""
import ctypes as C
import json
import os
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

from oracle.oracle_py import Oracle  # noqa: E402
from phygas_amd import workload  # noqa: E402
from test_cull import _aos  # noqa: E402

N_STARS, CHI2 = 40, 25.0


def main() -> None:
    out_so = os.path.join(tempfile.mkdtemp(), "libcull.so")
    cmd = [os.environ.get("HIPCC", "/opt/rocm/bin/hipcc"), "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC",
           "-shared", "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "phygas_amd", "csrc"),
           os.path.join(ROOT, "tests", "native", "cull_harness.cpp"), "-o", out_so]
    subprocess.run(cmd, check=True, capture_output=True)
    lib = C.CDLL(out_so)
    lib.cc_scan2.restype = C.c_int
    lib.cc_scan2.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                             C.c_double, C.c_void_p, C.c_void_p]
    o = Oracle()
    cl = workload.make_clusters(4087, 0)
    st = workload.make_stars(N_STARS, 2)
    n, off, pos, vel, S = o.lin_clusters(cl)
    npts = int(np.ceil(float(cl["tmax_myr"].max()) / 0.5)) + 1
    tn, toff, tpos, tvel, tS = o.lin_targets(st, float(cl["tmax_myr"].max()))
    ct = np.ascontiguousarray(_aos(n, off, pos, vel, S, npts))
    stt = np.ascontiguousarray(_aos(tn, toff, tpos, tvel, tS, npts))
    R2 = np.ascontiguousarray((2.0 * cl["r50_pc"] / 1000.0) ** 2)
    cn = np.ascontiguousarray(n.astype(np.int32))
    ntl = (npts + 7) // 8
    p2 = np.zeros((len(st), len(cl), ntl), np.uint8)
    counts = np.zeros(4, np.int64)
    lib.cc_scan2(stt.ctypes.data, len(st), npts, ct.ctypes.data, cn.ctypes.data, R2.ctypes.data, len(cl), npts, CHI2,
                 p2.ctypes.data, counts.ctypes.data)
    res = {"stars": N_STARS, "star_seed": 2, "clusters": len(cl), "cluster_seed": 0, "chi2_max": CHI2,
           "coarse_l1_tests": int(counts[0]), "coarse_l1_passes": int(counts[1]), "fine_tests": int(counts[2]),
           "fine_passes": int(counts[3]), "pairs_with_a_tile_left": int((p2.reshape(len(st), len(cl), -1).max(-1) > 0).sum())}
    with open(os.path.join(ROOT, "tests", "golden", "scan_counts.json"), "w") as f:
        json.dump(res, f, indent=1)
    print(res)


if __name__ == "__main__":
    main()

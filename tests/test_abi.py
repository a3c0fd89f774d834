"""CPU checks of the C ABI: the library loads and exports every function include/*.h declares; host-side
helpers (seeds, clone draws) are bit-identical to the oracle.  No device calls."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "phygas_amd", "libphygas_hip.so")


def _declared():
    """Every function include/*.h declares (static inline helpers of a header are not exported)."""
    names = set()
    for h in ("phygas_hip.h", "phygas_core.h"):
        with open(os.path.join(ROOT, "include", h)) as f:
            txt = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
        for m in re.finditer(r"^(?!static)[A-Za-z_][\w \*]*?\b((?:pg|phx|ph|mc)_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(LIB):
        from phygas_amd import build
        build.build()
    return LIB


def test_exports_every_declared_symbol(built):
    out = subprocess.run(["nm", "-D", "--defined-only", built], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    declared = _declared()
    assert len(declared) >= 60
    missing = sorted(declared - exported)
    assert not missing, missing


def test_library_loads_and_seeds_match_oracle(built, oracle):
    from phygas_amd import hip
    hip.lib()
    for a, b in [("lin", "C00001"), ("S0000001", "C01234"), ("", ""), ("J0534+2200", "Alessi_12")]:
        assert hip.seed(a, b) == oracle.seed(a, b)


def test_host_clone_draws_identical_to_oracle(built, oracle, workload):
    from phygas_amd import hip
    tg = np.concatenate([workload.make_pulsars(2, 1), workload.make_stars(2, 5)])
    tg[3]["plx"], tg[3]["e_plx"] = 0.15, 0.3
    cl = workload.make_clusters(5, 1)
    for t in tg:
        for c in cl:
            s = hip.seed(t["id"], c["name"])
            a = hip.pair_draw(t, c, 25, 0.05, s)
            b = oracle.pair_draw(t, c, 25, 0.05, s)
            assert a[0].tobytes() == b[0].tobytes() and a[1:] == b[1:]


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from phygas_amd import hip
    monkeypatch.setattr(hip, "_lib", None)
    monkeypatch.setattr(hip, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(hip.PhygasError):
        hip.lib()

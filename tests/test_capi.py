"""The reference's core C API as the product exports it (include/phygas_core.h, libphygas_hip.so; host code, no
device calls): the golden harness that pins the oracle, built a third time against the product's header and
library, must print the reference's golden vectors byte for byte (every section but coords, which is not part
of that API); and programs written against the reference's API (its CLIs' calls) link against the library."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "phygas_amd")
SECTIONS = ["accel", "rng", "seed", "sample", "lininit", "lin", "chi2", "pair", "track", "read"]
CORE_SYMBOLS = ["ph_init", "ph_potential", "ph_sample_target", "ph_sample_cluster", "ph_track_target",
                "ph_track_cluster", "ph_track_free", "ph_prefilter", "ph_pair", "ph_pair_draw", "ph_pair_clone",
                "ph_pair_reduce", "ph_lin_target", "ph_lin_target_init", "ph_lin_cluster", "ph_lin_free",
                "ph_screen_r", "ph_screen", "ph_screen_chi2", "ph_screen_chi2_cut", "ph_screen_finish", "ph_seed",
                "ph_read_clusters", "ph_read_targets", "mc_gaussian"]


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libphygas_hip.so")):
        from phygas_amd import build
        build.build()
    exe = str(tmp_path_factory.mktemp("core") / "golden_core")
    cmd = ["gcc", "-O2", "-std=c99", "-ffp-contract=off", "-DUSE_CORE", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "oracle", "golden_harness.c"), "-o", exe, "-L", LIBDIR, "-lphygas_hip",
           "-Wl,-rpath," + LIBDIR, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr[-3000:])
    return exe


def test_exports_the_core_api():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIBDIR, "libphygas_hip.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    assert not sorted(set(CORE_SYMBOLS) - exported)


@pytest.mark.parametrize("section", SECTIONS)
def test_core_api_prints_the_golden_vectors(harness, golden_dir, section):
    inp = os.path.join(golden_dir, "inputs")
    out = subprocess.run([harness, section, os.path.join(inp, "unit_clusters.txt"),
                          os.path.join(inp, "unit_targets.txt")], capture_output=True, check=True).stdout
    with open(os.path.join(golden_dir, f"unit_{section}.txt"), "rb") as f:
        assert out == f.read()


def test_screen_r_is_the_library_global():
    """ph_screen_r is a data symbol the CLIs assign (reference: src/phygas_screen.c:26): writes through it change
    the kernel radius the library's scan uses."""
    L = C.CDLL(os.path.join(LIBDIR, "libphygas_hip.so"))
    r = C.c_double.in_dll(L, "ph_screen_r")
    assert r.value == 2.0
    r.value = 0.0
    assert C.c_double.in_dll(L, "ph_screen_r").value == 0.0
    r.value = 2.0


REF_BIN = os.path.join(ROOT, "oracle", "_ref")


@pytest.mark.parametrize("prog,args,golden", [
    ("phygas_screen_on_product", ["clusters400.txt", "stars20.txt"], "rows_screen.txt"),
    ("phygas_screen_on_product", ["clusters400.txt", "stars20.txt", "25", "64", "2", "control"],
     "rows_screen_control.txt"),
    ("phygas_screen_on_product", ["clusters400.txt", "stars20.txt", "9", "64", "0"], "rows_screen_rtot.txt"),
    ("phygas_cli_on_product", ["clusters400.txt", "pulsar1.txt", "40", "0.05"], "rows_pair.txt"),
])
def test_reference_clis_linked_against_the_library(golden_dir, prog, args, golden):
    """The reference's phygas_screen.c / phygas_cli.c compiled from their own sources and linked against
    libphygas_hip.so in place of phygas.o (oracle/Makefile, `make ref`, where the reference sources exist):
    their stdout is the reference's rows byte for byte."""
    exe = os.path.join(REF_BIN, prog)
    if not os.path.exists(exe):
        pytest.skip(f"{prog} not built (make -C oracle ref needs the reference sources)")
    inp = os.path.join(golden_dir, "inputs")
    out = subprocess.run([exe] + [os.path.join(inp, a) if a.endswith(".txt") else a for a in args],
                         capture_output=True, check=True).stdout
    with open(os.path.join(golden_dir, golden), "rb") as f:
        assert out == f.read()

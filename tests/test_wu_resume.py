"""Resume step of the work-unit runner (phx_wu_resume, host code only): the output keeps exactly the lines the
checkpoint counts; anything the checkpoint cannot vouch for means a fresh start; an unreadable checkpoint
leaves everything alone.  (Reference behaviour being corrected: its app resumes from a bare target count and
keeps rows written after it, src/main_phygas_boinc.cpp:414-421.)"""
import os

import pytest

from phygas_amd import hip


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def _read(path):
    with open(path) as f:
        return f.read()


ROWS = "".join(f"S{i} C{i} o 1.000 2.000 3.00 4.00 0.5000 0.6000 7.00\n" for i in range(7))


def test_fresh_start_truncates(tmp_path):
    out, ck = str(tmp_path / "out"), str(tmp_path / "ckpt")
    _write(out, ROWS)
    assert hip.wu_resume(out, ck, 10) == (0, 0, 0)
    assert _read(out) == ""


def test_keeps_exactly_the_counted_lines(tmp_path):
    out, ck = str(tmp_path / "out"), str(tmp_path / "ckpt")
    _write(out, ROWS + "S7 C7 o 1.0")                 # rows after the checkpoint and a torn tail
    _write(ck, "phygas_checkpoint 2 3 5\n")
    assert hip.wu_resume(out, ck, 10) == (0, 3, 5)
    assert _read(out) == "".join(ROWS.splitlines(keepends=True)[:5])


@pytest.mark.parametrize("ckpt", ["phygas_checkpoint 2 3 9\n",      # counts more lines than out holds
                                  "phygas_checkpoint 2 -5 2\n",      # out of range
                                  "phygas_checkpoint 2 20 2\n",      # beyond the targets
                                  "phygas_checkpoint 1 3 2\n",       # another version
                                  "3\n",                             # the reference app's bare count
                                  "garbage\n", ""])
def test_untrusted_checkpoint_means_fresh_start(tmp_path, ckpt):
    out, ck = str(tmp_path / "out"), str(tmp_path / "ckpt")
    _write(out, ROWS)
    _write(ck, ckpt)
    assert hip.wu_resume(out, ck, 10) == (0, 0, 0)
    assert _read(out) == ""


def test_unreadable_checkpoint_leaves_everything(tmp_path):
    out, ck = str(tmp_path / "out"), str(tmp_path / "ckpt")
    _write(out, ROWS)
    os.mkdir(ck)                                       # exists, cannot be read as a file
    rc, t, n = hip.wu_resume(out, ck, 10)
    assert rc == 4 and (t, n) == (0, 0)
    assert _read(out) == ROWS


def test_status_codes_are_distinct(tmp_path):
    """Stopped, unreadable checkpoint and error are three different answers; an error raises."""
    assert len({hip.WU_DONE, hip.WU_STOPPED, hip.WU_CKPT_UNREADABLE, 1, 2}) == 5
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "phygas_hip.h")).read()
    assert "PHX_WU_CKPT_UNREADABLE = 4, PHX_WU_STOPPED = 6" in hdr
    # a missing cluster file fails before any device work: the code is an error, and the wrapper raises
    inp = str(tmp_path / "in")
    _write(inp, "PHYGAS 1 screen 25 64 2 0\n")
    with pytest.raises(hip.PhygasError):
        hip.wu_run(inp, str(tmp_path / "no_such_clusters"), str(tmp_path / "out"), str(tmp_path / "ck"))

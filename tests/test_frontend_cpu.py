"""The run.in front end rejects malformed or unsupported input with
DMGX_EINVAL and a message naming the file and line, before any device work
(run_parser.cpp:158-315 grammar)."""
import ctypes

import pytest


def _run(dmgx, tmp_path, text):
    (tmp_path / "run.in").write_text(text)
    L = dmgx.lib()
    st = L.dmgx_run_directory(str(tmp_path).encode(), None, None)
    return st, L.dmgx_last_error().decode()


@pytest.mark.parametrize("text,needle", [
    ("time_step 1\nrun 10\n", "does not contain a potential command"),
    ("potential nep.txt\nminimize sd 1e-6 100\n", "unsupported command 'minimize'"),
    ("potential nep.txt\nensemble nvt_ber 300 300\n", "should have 3 parameters"),
    ("potential nep.txt\nensemble nvt_ber 300 300 0.5\n", "temperature coupling should be >= 1"),
    ("potential nep.txt\nensemble npt_ber 300 300 100 0 0 1000\n", "bulk modulus should be > 0"),
    ("potential nep.txt\ntime_step 1 0\n", "maximum distance per step should be > 0"),
    ("potential nep.txt\ntime_step 1 0.1 3\n", "time_step should have 1 or 2 parameters"),
    ("potential nep.txt\ncorrect_velocity 5\n", "velocity correction interval should be >= 10"),
    ("potential nep.txt\ncorrect_velocity 10 -1\n", "grouping method should be >= 0"),
    ("potential nep.txt\ncorrect_velocity\n", "correct_velocity should have 1 or 2 parameters"),
    ("potential nep.txt\ndump_xyz 1 a.xyz group 0\n", "option 'group' requires a grouping method and group ID"),
    ("potential nep.txt\ndump_xyz 1 a.xyz group 0 -1\n", "grouping method and group ID should be >= 0"),
    ("potential nep.txt\ndump_xyz 1 a.xyz group 0 1 group 0 1\n", "option 'group' is specified more than once"),
    ("potential nep.txt\nensemble nvt_bdp 300 300 100\ncorrect_velocity 10\nrun 10\n",
     "ensemble nvt_bdp does not support correct_velocity"),
    ("potential nep.txt\ntime_step 1 0.05\nensemble nvt_nhc 300 300 100\nrun 10\n",
     "ensemble nvt_nhc does not support adaptive time_step"),
    ("potential nep.txt\ntime_step 1 0.05\nensemble npt_scr 300 300 100 0 100 1000\nrun 10\n",
     "does not support adaptive time_step"),
    ("potential nep.txt\ndump_thermo 0\n", "thermo dump interval should be > 0"),
    ("potential nep.txt\ndump_xyz 10 out.xyz precision quad\n", "invalid dump_xyz precision"),
    ("potential nep.txt\ndump_xyz 10 out.xyz bogus\n", "unrecognized dump_xyz argument"),
    ("potential nep.txt\nrun ten\n", "number of steps should be an integer"),
    ("potential a.txt\npotential b.txt\n", "multiple potential commands"),
])
def test_run_in_errors(dmgx, tmp_path, text, needle):
    st, msg = _run(dmgx, tmp_path, text)
    assert st == -1, (st, msg)
    assert needle in msg, msg
    assert "run.in" in msg


def test_missing_directory(dmgx):
    L = dmgx.lib()
    assert L.dmgx_run_directory(b"/nonexistent/dir", None, None) == -1
    assert "cannot open" in L.dmgx_last_error().decode()


def test_comment_and_blank_lines_are_skipped(dmgx, tmp_path):
    # parses fine up to the missing potential file (a runtime error of the
    # potential parser, reported as EINVAL with the path)
    st, msg = _run(dmgx, tmp_path, "# comment\n\npotential nep.txt # trailing\ntime_step 0.5\nrun 0\n")
    assert st == -1 and "nep.txt" in msg, msg

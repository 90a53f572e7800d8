"""Tuned-configuration files in the reference's format
(src/Dedispersion.cpp:20-81): the packaged files are what the sweep records
give, and the Python and C++ readers return the same entries."""
import os
import subprocess

import pytest

from dedispersion_amd import Conf, learned, tuner
from tests.test_library import host_check  # noqa: F401  (fixture)

MODES = ("single_step", "step_one", "step_two")


def test_packaged_files_follow_from_the_sweep_records():
    lines = tuner.tuned_lines_from_records(learned.load_records(tuner.PLAN_RECORDS_PATH))
    assert sorted(lines) == sorted(MODES)
    for mode in MODES:
        with open(tuner.tuned_file_path(mode)) as f:
            body = [l.rstrip("\n") for l in f if l[:1].isalpha()]
        assert body == lines[mode], mode


def test_reader_rules(tmp_path):
    p = tmp_path / "t.conf"
    p.write_text("# comment\nMI355X 64 0 1 4 64 4 1 16 2 1\n\n7 ignored\nMI355X 64 0 0 2 64 8 1 8 1 1\n"
                 "gfx950 32 0 1 8 64 2 1 4 8 1\n")
    t = tuner.read_tuned_file(str(p))
    assert t["MI355X"][64] == Conf(64, 4, 16, 2, 4, 1)  # the first line for (device, DMs) stays
    assert t["gfx950"][32] == Conf(64, 2, 4, 8, 8, 1)
    assert tuner.select_tuned_conf(t, "gfx950", 32).unroll == 8
    with pytest.raises(KeyError):
        tuner.select_tuned_conf(t, "MI355X", 128)
    with pytest.raises(OSError):
        tuner.read_tuned_file(str(p) + ".missing")


@pytest.mark.parametrize("mode", MODES)
def test_cpp_reader_agrees(host_check, mode):  # noqa: F811
    path = tuner.tuned_file_path(mode)
    out = subprocess.run([host_check, "readconf", path], check=True, capture_output=True, text=True).stdout
    cpp = sorted(out.splitlines())
    py = sorted(f"{dev} {dms} {c.print()}" for dev, m in tuner.read_tuned_file(path).items() for dms, c in m.items())
    assert cpp == py and len(py) > 0

"""The C-ABI library loads and exports what include/dedispersion_amd.h
declares; the host-side C++ API (no GPU needed) matches the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from dedispersion_amd import _lib, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dedispersion_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dd_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    declared = declared_functions()
    assert len(declared) >= 15
    assert set(declared) == set(_lib.PROTOTYPES), set(declared) ^ set(_lib.PROTOTYPES)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.dd_abi_version() == 3
    assert lib.dd_status_string(5) == b"invalid kernel configuration"


@pytest.fixture(scope="module")
def host_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cpp") / "host_api_check")
    lib_dir = os.path.join(ROOT, "dedispersion_amd")
    subprocess.run(
        ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
         os.path.join(ROOT, "tests", "cpp", "host_api_check.cpp"), "-L", lib_dir, "-ldedisp_amd",
         "-Wl,-rpath," + lib_dir, "-o", exe],
        check=True)
    return exe


@pytest.mark.parametrize("make,padding", [(workloads.plain_128, 128), (workloads.arts_two_step, 128),
                                          (workloads.arts_two_step, 0)])
def test_cpp_get_shifts_matches_oracle(host_check, make, padding):
    o = make().obs
    args = [host_check, "shifts", str(o.nr_channels), str(o.nr_subbands), repr(o.min_freq),
            repr(o.channel_bandwidth), str(o.nr_samples_per_batch), repr(o.sampling_time), str(o.downsampling),
            str(padding)]
    lines = subprocess.run(args, check=True, capture_output=True, text=True).stdout.split()
    n1, n2 = int(lines[0]), int(lines[1])
    vals = np.array([int(x, 16) for x in lines[2:]], dtype=np.uint32)
    orc = oracle.Oracle(o, padding)
    np.testing.assert_array_equal(vals[:n1], orc.shifts.view(np.uint32))
    np.testing.assert_array_equal(vals[n1:n1 + n2], orc.shifts2.view(np.uint32))


def test_cpp_tuned_conf_round_trip(host_check, tmp_path):
    out = subprocess.run([host_check, "conf", str(tmp_path / "tuned.conf")], check=True, capture_output=True,
                         text=True).stdout.splitlines()
    assert out == [
        "MI355X 64 0 0 2 64 8 1 8 1 1",
        "MI355X 2048 0 1 4 64 4 1 16 2 1",
        "gfx950 32 0 0 2 64 8 1 8 1 1",
        "FileError",
    ]


def test_cpp_zapped_file_and_beam_mapping(host_check, tmp_path):
    f = tmp_path / "zapped.txt"
    f.write_text("# zapped channels\n3 7\n7 15 # duplicate 7 counts once\n")
    out = subprocess.run([host_check, "zapped", str(f)], check=True, capture_output=True, text=True).stdout.splitlines()
    count, flags = out[0].split(":")
    assert int(count) == 3
    flags = [int(x) for x in flags.split()]
    assert len(flags) == 16 and [i for i, v in enumerate(flags) if v] == [3, 7, 15]
    n, mapping = out[1].split(":")
    mapping = np.array([int(x) for x in mapping.split()]).reshape(3, 16)
    assert int(n) == 48
    np.testing.assert_array_equal(mapping[:, 0], [0, 1, 0])

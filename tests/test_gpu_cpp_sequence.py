"""The reference's C++ call sequence on the GPU (tests/cpp/reference_sequence.cpp):
getShifts -> readTunedDedispersionConf -> get...HIP -> launch on hipMalloc'd
buffers, checked inside the program against the CPU functions of
include/Dedispersion.hpp and here against the oracle (src/DedispersionTest.cpp:
116-125, 129-146, 339-416).  The program is built by the build step
(tests/cpp/Makefile) and links the engine's library."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from dedispersion_amd import Observation
from tests.conftest import host_threads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "build", "reference_sequence")
PADDING = 128


def observation():
    """The observation reference_sequence.cpp builds (Test:80-89 setters)."""
    return Observation(nr_channels=256, min_freq=1300.0, channel_bandwidth=0.5, sampling_time=0.00008192,
                       nr_samples_per_batch=1500, nr_dms=40, first_dm=3.3, dm_step=0.7, nr_beams=2,
                       nr_synthesized_beams=3, nr_subbands=8, nr_dms_subbanding=12, first_dm_subbanding=12.7,
                       dm_step_subbanding=4.3)


def test_reference_call_sequence(torch_cuda, tmp_path):
    assert os.path.exists(EXE), "tests/cpp/build/reference_sequence is missing: run the build step"
    single = tmp_path / "single.conf"
    single.write_text("# device nrDMs split local unroll t0 t1 t2 i0 i1 i2\nMI355X 40 0 1 4 64 4 1 8 4 1\n")
    one = tmp_path / "one.conf"
    one.write_text("MI355X 12 0 1 2 64 2 1 16 2 1\n")
    two = tmp_path / "two.conf"
    two.write_text("MI355X 40 0 1 3 64 2 1 8 4 1\n")
    zapped_file = tmp_path / "zapped.txt"
    zapped_file.write_text("# channels\n0 17 128\n255\n")
    r = subprocess.run([EXE, str(tmp_path), str(single), str(one), str(two), str(zapped_file)], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout)
    lines = {l.split()[0]: l.split()[1:] for l in r.stdout.splitlines() if l.strip()}
    assert r.returncode == 0, r.stdout + r.stderr
    for case in ("single_step", "single_step_sbeams_1_2", "single_step_4bit", "single_step_float", "step_one",
                 "step_two", "invalid_conf_throws"):
        assert case in lines and int(lines[case][0]) == 0 and int(lines[case][1]) > 0, (case, lines.get(case))

    obs = observation()
    orc = oracle.Oracle(obs, PADDING, threads=host_threads())
    o = orc.o
    assert [int(x) for x in lines["sizes"]] == [o.nr_samples_per_dispersed_batch, o.nr_samples_per_batch_sub,
                                                o.nr_samples_per_dispersed_batch_sub]
    zapped = np.zeros(oracle.pad(256, PADDING // 4), dtype=np.uint32)
    zapped[[0, 17, 128, 255]] = 1
    mapping = orc.default_map(256)
    mapping[2, :100] = 0

    def raw(name, dtype):
        return np.fromfile(str(tmp_path / name), dtype=dtype)

    S = obs.nr_samples_per_batch
    want = orc.dedispersion(raw("single_in.u8", np.uint8), zapped, mapping)
    got = raw("single_out.f32", np.float32).reshape(want.shape)
    np.testing.assert_array_equal(got[..., :S], want[..., :S])
    orc_f = oracle.Oracle(obs, PADDING, threads=host_threads(), input_bits=32)
    want = orc_f.dedispersion(raw("float_in.f32", np.float32), zapped, mapping)
    got = raw("float_out.f32", np.float32).reshape(want.shape)
    np.testing.assert_array_equal(got[..., :S], want[..., :S])
    want_mid = orc.step_one(raw("two_in.u8", np.uint8), zapped)
    mid = raw("two_mid.f32", np.float32).reshape(want_mid.shape)
    np.testing.assert_array_equal(mid[..., : o.nr_samples_per_batch_sub], want_mid[..., : o.nr_samples_per_batch_sub])
    want = orc.step_two(want_mid.ravel())
    got = raw("two_out.f32", np.float32).reshape(want.shape)
    np.testing.assert_array_equal(got[..., :S], want[..., :S])

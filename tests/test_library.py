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
    assert lib.dd_abi_version() == 6
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


def integration_cpp_block() -> str:
    """The C++ example of INTEGRATION.md that follows 'This engine'."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    start = text.index("```cpp", text.index("This engine (`include/Dedispersion.hpp`"))
    return text[start + len("```cpp"):text.index("```", start + 6)]


def test_integration_cpp_example_compiles(tmp_path):
    """The documented call sequence compiles against include/ as written,
    wrapped in a function that declares the caller's variables."""
    block = integration_cpp_block()
    includes = [l for l in block.splitlines() if l.startswith("#include")]
    body = [l for l in block.splitlines() if not l.startswith("#include")]
    assert len(includes) == 3 and "readTunedDedispersionConf" in block
    src = tmp_path / "example.cpp"
    src.write_text("\n".join(includes) + """
#include <string>
#include <vector>
void example(AstroData::Observation &observation, unsigned int padding, const std::string &device,
             std::vector<unsigned int> &zappedChannels, std::vector<unsigned int> &beamMapping,
             const inputDataType *d_input, outputDataType *d_output, std::vector<inputDataType> &hostInput,
             std::vector<outputDataType> &hostOutput, unsigned int firstSynthesizedBeam, void *stream) {
""" + "\n".join(body) + "\n}\n")
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src)], check=True)


TEMPLATE_CASES = {
    # C NSB S D FIRST STEP DC FIRST1 STEP1 BEAMS SBEAMS DS S1 PADDING
    "two-beams": [48, 4, 300, 10, 2.5, 1.1, 5, 3.3, 4.7, 2, 3, 1, 0, 32],
    "unpadded": [40, 8, 257, 7, 0.0, 0.9, 3, 0.0, 6.0, 1, 1, 1, 0, 0],
    "downsampled": [32, 4, 200, 6, 1.2, 1.0, 3, 1.5, 5.0, 1, 2, 2, 260, 128],
}


@pytest.mark.parametrize("case", sorted(TEMPLATE_CASES))
def test_cpp_cpu_functions_match_oracle(host_check, tmp_path, case):
    """The CPU functions of include/Dedispersion.hpp (the reference's
    dedispersion / subbandDedispersionStepOne / subbandDedispersionStepTwo,
    DH:59-61) against the oracle: 8-bit, 2-bit packed and float rows for the
    single step, 8-bit and float rows for step one, general floats for step
    two, with zapped channels, a mixed beam mapping, nonzero first DMs,
    padded and unpadded rows and downsampling 2."""
    from dedispersion_amd import Observation

    C, nsb, S, D, first, step, dc, first1, step1, beams, sbeams, ds, s1, padding = TEMPLATE_CASES[case]
    args = [host_check, "templates", str(tmp_path)] + [repr(v) if isinstance(v, float) else str(v)
                                                       for v in TEMPLATE_CASES[case]]
    sizes = subprocess.run(args, check=True, capture_output=True, text=True).stdout.split()
    obs = Observation(nr_channels=C, min_freq=1300.0, channel_bandwidth=0.5, sampling_time=0.00008192,
                      nr_samples_per_batch=S, nr_dms=D, first_dm=first, dm_step=step, nr_beams=beams,
                      nr_synthesized_beams=sbeams, nr_subbands=nsb, downsampling=ds, nr_dms_subbanding=dc,
                      first_dm_subbanding=first1, dm_step_subbanding=step1, nr_samples_per_batch_subbanding=s1)
    orc = {bits: oracle.Oracle(obs, padding, threads=2, input_bits=bits) for bits in (8, 2, 32)}
    o = orc[8].o
    assert [int(x) for x in sizes] == [o.nr_samples_per_dispersed_batch, o.nr_samples_per_batch_sub,
                                       o.nr_samples_per_dispersed_batch_sub]
    zapped = np.zeros(oracle.pad(C, padding // 4), dtype=np.uint32)
    zapped[[1, C // 2]] = 1
    mapping = orc[8].default_map(C)
    if sbeams > 1 and beams > 1:
        mapping[1, 2] = 0

    def raw(name, dtype):
        return np.fromfile(str(tmp_path / name), dtype=dtype)

    for name, bits in (("single_u8", 8), ("single_2bit", 2), ("single_f32", 32)):
        dtype = np.float32 if bits == 32 else np.uint8
        want = orc[bits].dedispersion(raw(name + "_in", dtype), zapped, mapping)
        np.testing.assert_array_equal(raw(name + "_out", np.float32).reshape(want.shape), want, err_msg=name)
    for name, bits in (("one_u8", 8), ("one_f32", 32)):
        dtype = np.float32 if bits == 32 else np.uint8
        want = orc[bits].step_one(raw(name + "_in", dtype), zapped)
        np.testing.assert_array_equal(raw(name + "_out", np.float32).reshape(want.shape), want, err_msg=name)
    want = orc[8].step_two(raw("two_in", np.float32))
    np.testing.assert_array_equal(raw("two_out", np.float32).reshape(want.shape), want, err_msg="step two")


def test_peak_search_arguments_are_checked_on_the_host():
    """dd_detect_peaks, dd_detect_peaks_batch and dd_read_probe refuse bad
    arguments before any device call (so these run without a GPU)."""
    from dedispersion_amd import DDError, detect_peaks, detect_peaks_batch

    fake = 1 << 20  # never dereferenced: every call below fails its checks first
    with pytest.raises(DDError, match="nr_widths"):
        detect_peaks(fake, 4, 64, 64, 8, fake)
    with pytest.raises(DDError, match="nr_samples"):
        detect_peaks(fake, 4, 65, 64, 3, fake)
    with pytest.raises(DDError, match="null argument"):
        detect_peaks(0, 4, 64, 64, 3, fake)
    with pytest.raises(DDError, match="too large"):
        detect_peaks(fake, 1, 0xFFFFFFF0, 0xFFFFFFF0, 3, fake)
    detect_peaks(fake, 0, 64, 64, 3, fake)  # no rows: nothing to do
    detect_peaks_batch([], 7)  # no sets: nothing to do
    with pytest.raises(DDError, match="nr_widths"):
        detect_peaks_batch([(fake, 1, 64, 64, fake)], 0)
    with pytest.raises(DDError, match="job 2: need 0 < nr_samples <= pitch"):
        detect_peaks_batch([(fake, 1, 64, 64, fake), (0, 0, 0, 0, 0), (fake, 1, 64, 63, fake)], 7)
    lib = _lib.load()
    sink = fake
    assert lib.dd_read_probe(fake + 4, 1024, sink, None) == 1   # misaligned buffer
    assert lib.dd_read_probe(fake, 1000, sink, None) == 1       # not whole 16-byte pieces
    assert lib.dd_read_probe(fake, 1024, 0, None) == 1          # no sink

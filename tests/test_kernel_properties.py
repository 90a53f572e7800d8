"""Properties of the compiled device code, checked on the CPU over the gfx950
assembly of every HIP source of the library (built here with the library's
own flags):

* no register-indexed access (s_set_gpr_idx_*, v_movrel*): a rolled loop
  that indexes a register array at run time compiles to them, and an
  indexed write whose index runs past the array clobbers live registers;
* no scalar-cache writes (scalar memory stores and atomics, the scalar
  cache's write-back and discard): device code writes through vector stores
  only;
* no scratch (private segment) in any kernel except the flush forms of the
  packed 16-sample x 8-DM row-sum kernels (single step over more than 257
  live channels over packed rows), which spill at most 64 bytes;
* the row-sum kernels held to four waves per SIMD fit 128 VGPRs.

This file names scalar-cache write instructions to look for them; no GPU
run loads it.
"""
import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dedispersion_amd")
HIPCC = "/opt/rocm/bin/hipcc"

FORBIDDEN = {
    "register-indexed access": re.compile(r"^\s*(s_set_gpr_idx_\w+|v_movrel\w*)\b", re.M),
    "scalar-cache write": re.compile(r"^\s*(s_(buffer_|scratch_)?(store|atomic)\w*|s_dcache_(wb|discard)\w*)\b", re.M),
}
# template arguments RUN = 16, DPW = 8, FLUSH = true, PACKED = true
SPILL_ALLOWED = re.compile(r"ILi16ELi8ELb1ELb1EE")
SPILL_LIMIT = 64


def hip_sources():
    text = open(os.path.join(PKG, "Makefile")).read()
    block = text[text.index("HIP_SRCS"):text.index("CPP_SRCS")]
    return re.findall(r"csrc/\w+\.hip", block)


def device_asm(src: str, out_dir: str) -> str:
    out = os.path.join(out_dir, os.path.basename(src) + ".s")
    cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I../include", "-Icsrc", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S", "-o", out, src]
    subprocess.run(cmd, cwd=PKG, check=True, capture_output=True)
    with open(out) as f:
        return f.read()


def kernels(asm: str):
    """(name, scratch bytes, VGPRs) of every kernel in the code-object metadata."""
    meta = asm[asm.index("amdhsa.kernels:"):]
    out = []
    for m in re.finditer(r"\.name:\s+(\S+)\n(.*?)(?=\n  - |\Z)", meta, re.S):
        body = m.group(0)
        scratch = re.search(r"\.private_segment_fixed_size:\s+(\d+)", body)
        vgprs = re.search(r"\.vgpr_count:\s+(\d+)", body)
        out.append((m.group(1), int(scratch.group(1)), int(vgprs.group(1))))
    return out


@pytest.fixture(scope="module")
def assembly(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("asm"))
    srcs = hip_sources()
    assert len(srcs) >= 8
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        return dict(zip(srcs, pool.map(lambda s: device_asm(s, d), srcs)))


def test_no_register_indexing_and_no_scalar_cache_writes(assembly):
    for src, asm in assembly.items():
        for what, pattern in FORBIDDEN.items():
            hits = pattern.findall(asm)
            assert not hits, f"{src}: {what}: {hits[:3]}"


def test_no_scratch_outside_the_packed_flush_kernels(assembly):
    spilling = []
    total = 0
    for src, asm in assembly.items():
        # a source may hold device data only (the allocation-free run's array)
        ks = kernels(asm) if "amdhsa.kernels:" in asm else []
        total += len(ks)
        for name, scratch, _ in ks:
            if scratch == 0:
                continue
            assert SPILL_ALLOWED.search(name) and "rowsum_u8" in name, f"{src}: {name} uses {scratch} bytes of scratch"
            assert scratch <= SPILL_LIMIT, f"{src}: {name} spills {scratch} bytes"
            spilling.append(name)
    assert total >= 100
    # the allowed spills are what they are documented to be: at most the
    # three forms (per beam, persistent, tile list) of the one shape
    assert len(spilling) <= 3, spilling


def test_four_wave_row_sum_kernels_fit_128_vgprs(assembly):
    """Byte row sums whose sums fit 64 VGPRs (RUN x DPW <= 128, no flush
    sums; <= 64 over packed rows) ask for four waves per SIMD."""
    asm = assembly["csrc/kernels_u8_staged.hip"]
    checked = 0
    for name, _, vgprs in kernels(asm):
        m = re.search(r"ILi(\d+)ELi(\d+)ELb([01])ELb([01])EE", name)
        run, dpw, flush, packed = int(m.group(1)), int(m.group(2)), m.group(3) == "1", m.group(4) == "1"
        if not flush and run * dpw <= (64 if packed else 128):
            assert vgprs <= 128, (name, vgprs)
            checked += 1
    assert checked >= 30

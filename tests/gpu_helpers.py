"""Shared helpers of the GPU parity tests: device buffers come from torch,
every computation goes through the C ABI, the oracle checks the result."""
from __future__ import annotations

import numpy as np

import oracle
from dedispersion_amd import DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO, Conf, Plan, synth_gaussian_u8, synth_uniform_u8
from tests.conftest import host_threads


def stream_of(torch):
    return torch.cuda.current_stream().cuda_stream


def make_input(torch, plan: Plan, workload, rows_pitch: int, rows_len: int, seed: int):
    """Device uint8 input [beams][C][pitch]: noise over the dispersed length,
    zeros in the padding, then the workload's pulses in beam 0."""
    o = plan.obs
    n = o.nr_beams * o.nr_channels * rows_pitch
    d = torch.zeros(n, dtype=torch.uint8, device="cuda")
    s = stream_of(torch)
    if workload.noise == "uniform":
        synth_uniform_u8(d.data_ptr(), o.nr_beams * o.nr_channels, rows_len, rows_pitch, seed, s)
    else:
        synth_gaussian_u8(d.data_ptr(), o.nr_beams * o.nr_channels, rows_len, rows_pitch, 128.0, 16.0, seed, s)
    for p in workload.pulses:
        plan.inject_pulse(d.data_ptr(), rows_pitch, rows_len, p.dm_value, p.t0, p.width, p.amplitude, s)
    return d


def zapped_flags(plan: Plan, zapped) -> np.ndarray:
    flags = np.zeros(oracle.pad(plan.obs.nr_channels, plan.padding // 4), dtype=np.uint32)
    flags[list(zapped)] = 1
    return flags


def single_step(torch, workload, padding=128, conf: Conf | None = None, mapping=None, seed=None):
    """Runs the single-step path on the GPU; returns (plan, host input, host
    output, oracle, flags)."""
    plan = Plan(workload.obs, padding)
    flags = zapped_flags(plan, workload.zapped)
    plan.set_zapped_channels(flags)
    if mapping is not None:
        plan.set_beam_mapping(single=mapping)
    if conf is not None:
        plan.set_conf(DD_SINGLE_STEP, conf)
    z = plan.sizes
    d_in = make_input(torch, plan, workload, z.in_pitch, z.nr_dispersed_samples,
                      workload.seed if seed is None else seed)
    d_out = torch.full((z.output_elements,), -1.0, dtype=torch.float32, device="cuda")
    plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), stream=stream_of(torch))
    torch.cuda.synchronize()
    orc = oracle.Oracle(workload.obs, padding, threads=host_threads())
    return plan, d_in.cpu().numpy(), d_out.cpu().numpy(), orc, flags


def compare_rows(got: np.ndarray, want: np.ndarray, valid: int, what: str):
    """Bit-exact comparison of the valid samples of every row."""
    g = got[..., :valid]
    w = want[..., :valid]
    bad = np.argwhere(g != w)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first at {bad[:5].tolist()}: got {g[tuple(bad[0])]} want {w[tuple(bad[0])]}"


_HIP = None


def device_to_numpy(ptr: int, count: int, dtype) -> np.ndarray:
    """Copies `count` elements at a raw device address to a new host array
    (hipMemcpy through the HIP runtime the process already loaded)."""
    import ctypes

    global _HIP
    if _HIP is None:
        _HIP = ctypes.CDLL("libamdhip64.so")
        _HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _HIP.hipMemcpy.restype = ctypes.c_int
    out = np.empty(count, dtype=dtype)
    rc = _HIP.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2)  # hipMemcpyDeviceToHost
    assert rc == 0, f"hipMemcpy failed: {rc}"
    return out

"""MI355X-native incoherent dedispersion engine (Python plumbing).

The product is the native library ``libdedisp_amd.so`` (HIP kernels for gfx950
behind the C ABI of ``include/dedispersion_amd.h``, with the C++ API of
``include/Dedispersion.hpp`` on top).  This package binds that C ABI with
ctypes for tests, benchmarks and scripting; device memory and streams come
from the caller (for example PyTorch tensors: pass ``tensor.data_ptr()`` and
``torch.cuda.current_stream().cuda_stream``).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field, fields
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO, DDError, check, load

__all__ = [
    "Observation",
    "Conf",
    "Plan",
    "Batch",
    "DDError",
    "DD_SINGLE_STEP",
    "DD_STEP_ONE",
    "DD_STEP_TWO",
    "load",
    "synth_gaussian_u8",
    "synth_uniform_u8",
    "detect_peaks",
    "detect_peaks_batch",
    "PEAK_DTYPE",
]


@dataclass
class Observation:
    """Mirror of dd_observation (AstroData::Observation fields on the path)."""

    nr_channels: int
    min_freq: float
    channel_bandwidth: float
    sampling_time: float
    nr_samples_per_batch: int
    nr_dms: int
    first_dm: float = 0.0
    dm_step: float = 0.0
    nr_beams: int = 1
    nr_synthesized_beams: int = 1
    nr_subbands: int = 1
    downsampling: int = 1
    nr_dms_subbanding: int = 0
    first_dm_subbanding: float = 0.0
    dm_step_subbanding: float = 0.0
    nr_samples_per_dispersed_batch: int = 0
    nr_samples_per_batch_subbanding: int = 0
    nr_samples_per_dispersed_batch_subbanding: int = 0

    def to_c(self) -> _lib.DDObservation:
        c = _lib.DDObservation()
        for f in fields(self):
            setattr(c, f.name, getattr(self, f.name))
        return c


@dataclass
class Conf:
    """Mirror of dd_conf / Dedispersion::DedispersionConf."""

    threads_d0: int = 64
    threads_d1: int = 4
    items_d0: int = 16
    items_d1: int = 4
    unroll: int = 4
    local: int = 1
    split_batches: int = 0

    def to_c(self) -> _lib.DDConf:
        c = _lib.DDConf()
        for f in fields(self):
            setattr(c, f.name, int(getattr(self, f.name)))
        return c

    @staticmethod
    def from_c(c: _lib.DDConf) -> "Conf":
        return Conf(*c.as_tuple())

    def print(self) -> str:
        """DedispersionConf::print(): "split local unroll t0 t1 t2 i0 i1 i2"."""
        return (
            f"{self.split_batches} {self.local} {self.unroll} "
            f"{self.threads_d0} {self.threads_d1} 1 {self.items_d0} {self.items_d1} 1"
        )


def _u32_ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _f32_ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Plan:
    """A dd_plan: observation, padding, device-resident tables, launches."""

    def __init__(self, obs: Observation, padding: int = 128, input_bits: int = 8):
        """input_bits: the reference's inputBits (1, 2, 4 = packed input rows,
        8, or 32 = float input rows for dedisperse_f32 / step_one_f32)."""
        self.lib = load()
        self.obs = obs
        self.padding = padding
        self.input_bits = input_bits
        self._c_obs = obs.to_c()
        handle = ctypes.c_void_p()
        check(self.lib.dd_plan_create(ctypes.byref(handle), ctypes.byref(self._c_obs), padding))
        self.handle = handle
        if input_bits != 8:
            check(self.lib.dd_plan_set_input_bits(self.handle, input_bits))
        s = _lib.DDSizes()
        check(self.lib.dd_plan_get_sizes(self.handle, ctypes.byref(s)))
        self.sizes = s

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.dd_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- padded sizes (elements) ----
    def pad(self, n: int, elem_bytes: int) -> int:
        k = self.padding // elem_bytes
        return n if k == 0 else ((n + k - 1) // k) * k

    # ---- configuration and tables ----
    def set_conf(self, mode: int, conf: Conf) -> None:
        c = conf.to_c()
        check(self.lib.dd_plan_set_conf(self.handle, mode, ctypes.byref(c)))

    def apply_tuned(self, mode: int, tuned: dict, device: str = "MI355X") -> Conf:
        """Sets the configuration of `mode` from a tuned-configuration map
        (tuner.read_tuned_file, the reference's tunedDedispersionConf) at
        this plan's DM count for the mode: D for the single step and step
        two, Dc for step one.  KeyError when the map has no entry."""
        o = self.obs
        nr_dms = o.nr_dms_subbanding if mode == DD_STEP_ONE else o.nr_dms
        conf = tuned[device][nr_dms]
        self.set_conf(mode, conf)
        return conf

    def set_step_one_shifts(self, absolute: bool) -> None:
        """Step one's delay per channel: relative to its subband's top channel
        (the reference's CPU function, the default) or its own (the
        reference's OpenCL kernel)."""
        check(self.lib.dd_plan_set_step_one_shifts(self.handle, 1 if absolute else 0))
        s = _lib.DDSizes()
        check(self.lib.dd_plan_get_sizes(self.handle, ctypes.byref(s)))
        self.sizes = s

    def get_conf(self, mode: int) -> Conf:
        c = _lib.DDConf()
        check(self.lib.dd_plan_get_conf(self.handle, mode, ctypes.byref(c)))
        return Conf.from_c(c)

    def set_zapped_channels(self, zapped: Sequence[int]) -> None:
        flags = np.zeros(self.pad(self.obs.nr_channels, 4), dtype=np.uint32)
        z = np.asarray(zapped, dtype=np.uint32)
        flags[: len(z)] = z
        check(self.lib.dd_plan_set_zapped_channels(self.handle, _u32_ptr(flags)))

    def set_beam_mapping(self, single: Optional[np.ndarray] = None, step_two: Optional[np.ndarray] = None) -> None:
        a = None if single is None else np.ascontiguousarray(single, dtype=np.uint32)
        b = None if step_two is None else np.ascontiguousarray(step_two, dtype=np.uint32)
        check(
            self.lib.dd_plan_set_beam_mapping(
                self.handle, None if a is None else _u32_ptr(a), None if b is None else _u32_ptr(b)
            )
        )

    def set_shifts(self, shifts: Optional[np.ndarray] = None, shifts_step_two: Optional[np.ndarray] = None) -> None:
        a = None if shifts is None else np.ascontiguousarray(shifts, dtype=np.float32)
        b = None if shifts_step_two is None else np.ascontiguousarray(shifts_step_two, dtype=np.float32)
        check(
            self.lib.dd_plan_set_shifts(
                self.handle, None if a is None else _f32_ptr(a), None if b is None else _f32_ptr(b)
            )
        )
        s = _lib.DDSizes()
        check(self.lib.dd_plan_get_sizes(self.handle, ctypes.byref(s)))
        self.sizes = s

    def get_shifts(self):
        a = np.zeros(self.pad(self.obs.nr_channels, 4), dtype=np.float32)
        b = np.zeros(self.pad(self.obs.nr_subbands, 4), dtype=np.float32)
        check(self.lib.dd_plan_get_shifts(self.handle, _f32_ptr(a), _f32_ptr(b)))
        return a, b

    def get_integer_shifts(self, mode: int) -> np.ndarray:
        o = self.obs
        if mode == DD_SINGLE_STEP:
            shape = (o.nr_channels, o.nr_dms)
        elif mode == DD_STEP_ONE:
            shape = (o.nr_channels, o.nr_dms_subbanding)
        else:
            shape = (o.nr_subbands, o.nr_dms)
        t = np.zeros(shape, dtype=np.uint32)
        check(self.lib.dd_plan_get_integer_shifts(self.handle, mode, _u32_ptr(t)))
        return t

    # ---- launches (device pointers as ints) ----
    def dedisperse(self, d_input: int, d_output: int, first_sbeam: int = 0, nr_sbeams: int = 0, stream: int = 0):
        check(self.lib.dd_dedisperse(self.handle, d_input, d_output, first_sbeam, nr_sbeams, stream or None))

    def step_one(self, d_input: int, d_intermediate: int, stream: int = 0):
        check(self.lib.dd_subband_step_one(self.handle, d_input, d_intermediate, stream or None))

    def dedisperse_f32(self, d_input: int, d_output: int, first_sbeam: int = 0, nr_sbeams: int = 0, stream: int = 0):
        """The single step over float input rows (a plan with input_bits 32)."""
        check(self.lib.dd_dedisperse_f32(self.handle, d_input, d_output, first_sbeam, nr_sbeams, stream or None))

    def step_one_f32(self, d_input: int, d_intermediate: int, stream: int = 0):
        """Step one over float input rows (a plan with input_bits 32)."""
        check(self.lib.dd_subband_step_one_f32(self.handle, d_input, d_intermediate, stream or None))

    def step_two(self, d_intermediate: int, d_output: int, first_sbeam: int = 0, nr_sbeams: int = 0, stream: int = 0):
        check(self.lib.dd_subband_step_two(self.handle, d_intermediate, d_output, first_sbeam, nr_sbeams, stream or None))

    def inject_pulse(self, d_beam_input: int, pitch: int, row_len: int, dm_value: float, t0: int, width: int,
                     amplitude: int, stream: int = 0):
        check(
            self.lib.dd_synth_inject_pulse(
                self.handle, d_beam_input, pitch, row_len, dm_value, t0, width, amplitude, stream or None
            )
        )


class Batch:
    """A dd_batch: jobs of several plans run by one persistent launch per
    kernel kind (optionally replayed from a HIP graph).  The plans and the
    device buffers of the jobs must outlive the batch."""

    def __init__(self):
        self.lib = load()
        handle = ctypes.c_void_p()
        check(self.lib.dd_batch_create(ctypes.byref(handle)))
        self.handle = handle
        self._keep = []  # plans and buffers referenced by jobs

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.dd_batch_destroy(self.handle)
            self.handle = None
        self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_single_step(self, plan: Plan, d_input: int, d_output: int, first_sbeam: int = 0, nr_sbeams: int = 0,
                        keep=()):
        check(self.lib.dd_batch_add_single_step(self.handle, plan.handle, d_input, d_output, first_sbeam, nr_sbeams))
        self._keep.append((plan, keep))

    def add_two_step(self, plan: Plan, d_input: int, d_intermediate: int, d_output: int, first_sbeam: int = 0,
                     nr_sbeams: int = 0, keep=()):
        check(self.lib.dd_batch_add_two_step(self.handle, plan.handle, d_input, d_intermediate, d_output,
                                             first_sbeam, nr_sbeams))
        self._keep.append((plan, keep))

    def set_conf(self, mode: int, conf: Conf) -> None:
        c = conf.to_c()
        check(self.lib.dd_batch_set_conf(self.handle, mode, ctypes.byref(c)))

    def set_step_one_shifts(self, absolute: bool) -> None:
        """Step one's delay per channel: relative to its subband's top channel
        (the reference's CPU function, the default) or its own (the
        reference's OpenCL kernel)."""
        check(self.lib.dd_plan_set_step_one_shifts(self.handle, 1 if absolute else 0))
        s = _lib.DDSizes()
        check(self.lib.dd_plan_get_sizes(self.handle, ctypes.byref(s)))
        self.sizes = s

    def get_conf(self, mode: int) -> Conf:
        c = _lib.DDConf()
        check(self.lib.dd_batch_get_conf(self.handle, mode, ctypes.byref(c)))
        return Conf.from_c(c)

    def set_pipeline(self, chunks: int) -> None:
        check(self.lib.dd_batch_set_pipeline(self.handle, chunks))

    def set_lds_reserve(self, mode: int, nbytes: int) -> None:
        check(self.lib.dd_batch_set_lds_reserve(self.handle, mode, nbytes))

    def set_compact_intermediate(self, enable: bool) -> None:
        check(self.lib.dd_batch_set_compact_intermediate(self.handle, 1 if enable else 0))

    def set_paired_step_two(self, enable: bool) -> None:
        check(self.lib.dd_batch_set_paired_step_two(self.handle, 1 if enable else 0))

    def run(self, use_graph: bool = False, stream: int = 0) -> None:
        check(self.lib.dd_batch_run(self.handle, 1 if use_graph else 0, stream or None))

    def run_mode(self, mode: int, stream: int = 0) -> None:
        check(self.lib.dd_batch_run_mode(self.handle, mode, stream or None))

    def run_per_job(self, stream: int = 0) -> None:
        check(self.lib.dd_batch_run_per_job(self.handle, stream or None))

    def job_placement(self, job: int) -> dict:
        """Where job `job` runs: {"beside": beside the batch kernels, with the
        float intermediate; "compact": its intermediate holds uint16 sums}."""
        beside, compact = ctypes.c_uint32(), ctypes.c_uint32()
        check(self.lib.dd_batch_get_job_placement(self.handle, job, ctypes.byref(beside), ctypes.byref(compact)))
        return {"beside": bool(beside.value), "compact": bool(compact.value)}

    def info(self) -> dict:
        i = _lib.DDBatchInfo()
        check(self.lib.dd_batch_get_info(self.handle, ctypes.byref(i)))
        return {"jobs": i.jobs, "tiles": list(i.tiles), "workgroups": list(i.workgroups),
                "lds_bytes": list(i.lds_bytes), "units": i.units, "compact_intermediate": i.compact_intermediate,
                "pipeline_chunks": i.pipeline_chunks, "solo_jobs": i.solo_jobs,
                "paired_step_two": i.paired_step_two}


def synth_gaussian_u8(d_data: int, rows: int, cols: int, pitch: int, mean: float, std: float, seed: int,
                      stream: int = 0) -> None:
    check(load().dd_synth_gaussian_u8(d_data, rows, cols, pitch, mean, std, seed, stream or None))


def synth_uniform_u8(d_data: int, rows: int, cols: int, pitch: int, seed: int, stream: int = 0) -> None:
    check(load().dd_synth_uniform_u8(d_data, rows, cols, pitch, seed, stream or None))


# One dd_peak as a numpy record (copy a device array of them to the host with
# this dtype).
PEAK_DTYPE = np.dtype([("snr", "<f4"), ("mean", "<f4"), ("sigma", "<f4"), ("sample", "<u4"), ("width", "<u4")])


def detect_peaks(d_rows: int, nr_rows: int, nr_samples: int, pitch: int, nr_widths: int, d_peaks: int,
                 stream: int = 0) -> None:
    """Boxcar peak search (dd_detect_peaks) over nr_rows device rows of
    nr_samples non-negative integer-valued floats (below 2^24), pitch floats
    apart; d_peaks receives nr_rows dd_peak records (PEAK_DTYPE)."""
    check(load().dd_detect_peaks(d_rows, nr_rows, nr_samples, pitch, nr_widths, d_peaks, stream or None))


def detect_peaks_batch(jobs, nr_widths: int, stream: int = 0) -> None:
    """The same search over many sets of rows in one launch per 64 sets
    (dd_detect_peaks_batch).  jobs: (d_rows, nr_rows, nr_samples, pitch,
    d_peaks) per set."""
    from ._lib import DDDetectJob

    arr = (DDDetectJob * max(1, len(jobs)))()
    for a, (d_rows, nr_rows, nr_samples, pitch, d_peaks) in zip(arr, jobs):
        a.d_rows, a.nr_rows, a.nr_samples, a.pitch, a.d_peaks = d_rows, nr_rows, nr_samples, pitch, d_peaks
    check(load().dd_detect_peaks_batch(arr, len(jobs), nr_widths, stream or None))

"""ctypes binding of the CPU oracle (dd_oracle.c).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py to check results; the library never uses it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_float, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libdd_oracle.so")


class OrcObs(ctypes.Structure):
    _fields_ = [
        ("nr_beams", c_uint32),
        ("nr_synthesized_beams", c_uint32),
        ("nr_channels", c_uint32),
        ("nr_subbands", c_uint32),
        ("min_freq", c_float),
        ("channel_bandwidth", c_float),
        ("sampling_time", c_float),
        ("downsampling", c_uint32),
        ("nr_samples_per_batch", c_uint32),
        ("nr_samples_per_dispersed_batch", c_uint32),
        ("nr_samples_per_batch_sub", c_uint32),
        ("nr_samples_per_dispersed_batch_sub", c_uint32),
        ("nr_dms", c_uint32),
        ("first_dm", c_float),
        ("dm_step", c_float),
        ("nr_dms_sub", c_uint32),
        ("first_dm_sub", c_float),
        ("dm_step_sub", c_float),
        ("input_bits", c_uint32),
        ("step_one_absolute", c_uint32),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = POINTER
    lib.orc_pad.restype = c_uint32
    lib.orc_pad.argtypes = [c_uint32, c_uint32]
    for name in ("orc_max_freq", "orc_subband_bandwidth", "orc_subband_min_freq", "orc_subband_max_freq"):
        getattr(lib, name).restype = c_float
        getattr(lib, name).argtypes = [P(OrcObs)]
    lib.orc_channels_per_subband.restype = c_uint32
    lib.orc_channels_per_subband.argtypes = [P(OrcObs)]
    lib.orc_get_shifts.restype = None
    lib.orc_get_shifts.argtypes = [P(OrcObs), c_uint32, P(c_float)]
    lib.orc_get_shifts_step_two.restype = None
    lib.orc_get_shifts_step_two.argtypes = [P(OrcObs), c_uint32, P(c_float)]
    lib.orc_dispersed_samples.restype = c_uint32
    lib.orc_dispersed_samples.argtypes = [c_uint32, c_float, c_float, c_float, c_uint32]
    lib.orc_input_pitch.restype = c_uint64
    lib.orc_input_pitch.argtypes = [c_uint32, c_uint32, c_uint32]
    lib.orc_shift.restype = c_uint32
    lib.orc_shift.argtypes = [c_float, c_float, c_uint32, c_float]
    lib.orc_dedispersion_mt.restype = None
    lib.orc_dedispersion_mt.argtypes = [P(OrcObs), P(c_uint32), P(c_uint32), c_void_p, P(c_float), P(c_float),
                                        c_uint32, c_uint64, c_uint64, c_int]
    lib.orc_step_one_mt.restype = None
    lib.orc_step_one_mt.argtypes = [P(OrcObs), P(c_uint32), c_void_p, P(c_float), P(c_float), c_uint32,
                                    c_uint64, c_uint64, c_int]
    lib.orc_step_two_mt.restype = None
    lib.orc_step_two_mt.argtypes = [P(OrcObs), P(c_uint32), c_void_p, P(c_float), P(c_float), c_uint32,
                                    c_uint64, c_uint64, c_int]
    lib.orc_detect_peaks.restype = None
    lib.orc_detect_peaks.argtypes = [P(c_float), c_uint64, c_uint32, c_uint64, c_uint32, P(c_float), P(c_float),
                                     P(c_float), P(c_uint32), P(c_uint32)]
    _lib = lib
    return lib


def _u32(a):
    return a.ctypes.data_as(POINTER(c_uint32))


def _f32(a):
    return a.ctypes.data_as(POINTER(c_float))


def pad(v: int, k: int) -> int:
    return v if k == 0 else ((v + k - 1) // k) * k


class Oracle:
    """The reference's CPU functions over one observation (numpy in/out)."""

    def __init__(self, obs, padding: int, threads: int = 0, input_bits: int = 8, step_one_absolute: bool = False):
        """obs: a dedispersion_amd.Observation-like object with the same field
        names; dispersed sizes of 0 are derived with the reference's rules.
        input_bits: the reference's inputBits (1, 2, 4 = packed rows, 8, or 32
        = float input).  step_one_absolute: step one applies each channel's
        own delay (the reference's OpenCL kernel) instead of its delay
        relative to the subband's top channel (the CPU function)."""
        assert input_bits in (1, 2, 4, 8, 32)
        self.lib = load()
        self.padding = padding
        self.input_bits = input_bits
        self.threads = threads or max(1, os.cpu_count() or 1)
        o = OrcObs()
        o.nr_beams = obs.nr_beams
        o.nr_synthesized_beams = obs.nr_synthesized_beams
        o.nr_channels = obs.nr_channels
        o.nr_subbands = obs.nr_subbands
        o.min_freq = obs.min_freq
        o.channel_bandwidth = obs.channel_bandwidth
        o.sampling_time = obs.sampling_time
        o.downsampling = obs.downsampling
        o.nr_samples_per_batch = obs.nr_samples_per_batch
        o.nr_dms = obs.nr_dms
        o.first_dm = obs.first_dm
        o.dm_step = obs.dm_step
        o.nr_dms_sub = obs.nr_dms_subbanding
        o.first_dm_sub = obs.first_dm_subbanding
        o.dm_step_sub = obs.dm_step_subbanding
        o.input_bits = input_bits
        o.step_one_absolute = 1 if step_one_absolute else 0
        self.o = o
        self.shifts = self.get_shifts()
        self.shifts2 = self.get_shifts_step_two()
        L = self.lib
        o.nr_samples_per_dispersed_batch = obs.nr_samples_per_dispersed_batch or L.orc_dispersed_samples(
            o.nr_samples_per_batch, float(self.shifts[0]), o.first_dm, o.dm_step, o.nr_dms)
        if o.nr_dms_sub:
            o.nr_samples_per_batch_sub = obs.nr_samples_per_batch_subbanding or L.orc_dispersed_samples(
                o.nr_samples_per_batch, float(self.shifts2[0]), o.first_dm, o.dm_step, o.nr_dms)
            o.nr_samples_per_dispersed_batch_sub = (
                obs.nr_samples_per_dispersed_batch_subbanding
                or L.orc_dispersed_samples(o.nr_samples_per_batch_sub, float(self.shifts[0]), o.first_dm_sub,
                                           o.dm_step_sub, o.nr_dms_sub))

    # sizes (elements)
    @property
    def input_dtype(self):
        return np.float32 if self.input_bits == 32 else np.uint8

    @property
    def in_pitch(self):
        """Input elements (bytes, or floats) per channel row of the single-step input."""
        return self.lib.orc_input_pitch(self.o.nr_samples_per_dispersed_batch, self.input_bits, self.padding)

    @property
    def out_pitch(self):
        return pad(self.o.nr_samples_per_batch // self.o.downsampling, self.padding // 4)

    @property
    def sub_in_pitch(self):
        """Input elements (bytes, or floats) per channel row of the step-one input."""
        return self.lib.orc_input_pitch(self.o.nr_samples_per_dispersed_batch_sub, self.input_bits, self.padding)

    @property
    def sub_out_pitch(self):
        """Floats per row of step one's output: pad(S1) (include/Dedispersion.hpp:148)."""
        return pad(self.o.nr_samples_per_batch_sub, self.padding // 4)

    @property
    def sub_two_in_pitch(self):
        """Floats per row of step two's input: pad(S1 / downsampling) (include/Dedispersion.hpp:169),
        which equals sub_out_pitch only when downsampling is 1."""
        return pad(self.o.nr_samples_per_batch_sub // self.o.downsampling, self.padding // 4)

    def get_shifts(self) -> np.ndarray:
        out = np.zeros(pad(self.o.nr_channels, self.padding // 4), dtype=np.float32)
        self.lib.orc_get_shifts(ctypes.byref(self.o), self.padding, _f32(out))
        return out

    def get_shifts_step_two(self) -> np.ndarray:
        out = np.zeros(pad(self.o.nr_subbands, self.padding // 4), dtype=np.float32)
        self.lib.orc_get_shifts_step_two(ctypes.byref(self.o), self.padding, _f32(out))
        return out

    def shift(self, first: float, step: float, dm: int, table: float) -> int:
        return self.lib.orc_shift(first, step, dm, table)

    def default_map(self, rows: int) -> np.ndarray:
        p = pad(rows, self.padding // 4)
        m = np.zeros((self.o.nr_synthesized_beams, p), dtype=np.uint32)
        m[:] = (np.arange(self.o.nr_synthesized_beams) % self.o.nr_beams)[:, None]
        return m

    def dedispersion(self, inp: np.ndarray, zapped=None, beam_mapping=None, shifts=None, rows=None):
        o = self.o
        C = o.nr_channels
        z = np.zeros(pad(C, self.padding // 4), dtype=np.uint32) if zapped is None else np.ascontiguousarray(zapped, dtype=np.uint32)
        m = self.default_map(C) if beam_mapping is None else np.ascontiguousarray(beam_mapping, dtype=np.uint32)
        s = self.shifts if shifts is None else np.ascontiguousarray(shifts, dtype=np.float32)
        assert np.asarray(inp).dtype == self.input_dtype, "input type of the observation's input_bits"
        inp = np.ascontiguousarray(inp)
        assert inp.size >= o.nr_beams * C * self.in_pitch
        out = np.zeros(o.nr_synthesized_beams * o.nr_dms * self.out_pitch, dtype=np.float32)
        r0, r1 = (0, o.nr_synthesized_beams * o.nr_dms) if rows is None else rows
        self.lib.orc_dedispersion_mt(ctypes.byref(o), _u32(z), _u32(m), inp.ctypes.data, _f32(out), _f32(s),
                                     self.padding, r0, r1, self.threads)
        return out.reshape(o.nr_synthesized_beams, o.nr_dms, self.out_pitch)

    def step_one(self, inp: np.ndarray, zapped=None, shifts=None, rows=None):
        o = self.o
        C = o.nr_channels
        z = np.zeros(pad(C, self.padding // 4), dtype=np.uint32) if zapped is None else np.ascontiguousarray(zapped, dtype=np.uint32)
        s = self.shifts if shifts is None else np.ascontiguousarray(shifts, dtype=np.float32)
        assert np.asarray(inp).dtype == self.input_dtype, "input type of the observation's input_bits"
        inp = np.ascontiguousarray(inp)
        assert inp.size >= o.nr_beams * C * self.sub_in_pitch
        out = np.zeros(o.nr_beams * o.nr_dms_sub * o.nr_subbands * self.sub_out_pitch, dtype=np.float32)
        r0, r1 = (0, o.nr_beams * o.nr_dms_sub) if rows is None else rows
        self.lib.orc_step_one_mt(ctypes.byref(o), _u32(z), inp.ctypes.data, _f32(out), _f32(s), self.padding,
                                 r0, r1, self.threads)
        return out.reshape(o.nr_beams, o.nr_dms_sub, o.nr_subbands, self.sub_out_pitch)

    def step_two(self, inp: np.ndarray, beam_mapping=None, shifts=None, rows=None):
        o = self.o
        m = self.default_map(o.nr_subbands) if beam_mapping is None else np.ascontiguousarray(beam_mapping, dtype=np.uint32)
        s = self.shifts2 if shifts is None else np.ascontiguousarray(shifts, dtype=np.float32)
        inp = np.ascontiguousarray(inp, dtype=np.float32)
        assert inp.size >= o.nr_beams * o.nr_dms_sub * o.nr_subbands * self.sub_two_in_pitch
        out = np.zeros(o.nr_synthesized_beams * o.nr_dms_sub * o.nr_dms * self.out_pitch, dtype=np.float32)
        r0, r1 = (0, o.nr_synthesized_beams * o.nr_dms_sub * o.nr_dms) if rows is None else rows
        self.lib.orc_step_two_mt(ctypes.byref(o), _u32(m), inp.ctypes.data, _f32(out), _f32(s), self.padding,
                                 r0, r1, self.threads)
        return out.reshape(o.nr_synthesized_beams, o.nr_dms_sub * o.nr_dms, self.out_pitch)


def detect_peaks(rows: np.ndarray, nr_samples: int, nr_widths: int) -> dict:
    """orc_detect_peaks over a [rows][pitch] float array: per row snr, mean,
    sigma, sample and width of the best boxcar."""
    lib = load()
    rows = np.ascontiguousarray(rows, dtype=np.float32)
    n, pitch = rows.shape
    out = {k: np.zeros(n, dtype=np.float32) for k in ("snr", "mean", "sigma")}
    out.update({k: np.zeros(n, dtype=np.uint32) for k in ("sample", "width")})
    lib.orc_detect_peaks(_f32(rows), n, nr_samples, pitch, nr_widths, _f32(out["snr"]), _f32(out["mean"]),
                         _f32(out["sigma"]), _u32(out["sample"]), _u32(out["width"]))
    return out

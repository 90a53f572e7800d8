"""ctypes binding of the C ABI declared in include/dedispersion_amd.h.

The shared library libdedisp_amd.so is built in-tree (``make -C
dedispersion_amd`` or ``__graft_entry__.build()``).  Loading fails loudly when
it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_uint32, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdedisp_amd.so")

DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO = 0, 1, 2

STATUS_NAMES = {
    0: "DD_OK",
    1: "DD_ERR_INVALID_ARGUMENT",
    2: "DD_ERR_UNSUPPORTED",
    3: "DD_ERR_HIP",
    4: "DD_ERR_OUT_OF_MEMORY",
    5: "DD_ERR_CONFIGURATION",
    6: "DD_ERR_NOT_READY",
    7: "DD_ERR_IO",
}


class DDObservation(ctypes.Structure):
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
        ("nr_dms", c_uint32),
        ("first_dm", c_float),
        ("dm_step", c_float),
        ("nr_dms_subbanding", c_uint32),
        ("first_dm_subbanding", c_float),
        ("dm_step_subbanding", c_float),
        ("nr_samples_per_dispersed_batch", c_uint32),
        ("nr_samples_per_batch_subbanding", c_uint32),
        ("nr_samples_per_dispersed_batch_subbanding", c_uint32),
    ]


class DDConf(ctypes.Structure):
    _fields_ = [
        ("threads_d0", c_uint32),
        ("threads_d1", c_uint32),
        ("items_d0", c_uint32),
        ("items_d1", c_uint32),
        ("unroll", c_uint32),
        ("local", c_uint32),
        ("split_batches", c_uint32),
    ]

    def as_tuple(self):
        return tuple(getattr(self, f) for f, _ in self._fields_)


class DDSizes(ctypes.Structure):
    _fields_ = [
        ("nr_samples", c_uint32),
        ("nr_dispersed_samples", c_uint32),
        ("nr_samples_subbanding", c_uint32),
        ("nr_dispersed_samples_subbanding", c_uint32),
        ("in_pitch", c_uint64),
        ("out_pitch", c_uint64),
        ("sub_in_pitch", c_uint64),
        ("sub_out_pitch", c_uint64),
        ("input_elements", c_uint64),
        ("output_elements", c_uint64),
        ("sub_input_elements", c_uint64),
        ("intermediate_elements", c_uint64),
        ("sub_output_elements", c_uint64),
        ("max_shift", c_uint32),
        ("max_shift_step_one", c_uint32),
        ("max_shift_step_two", c_uint32),
        ("input_bits", c_uint32),
        ("nr_samples_step_one", c_uint32),
        ("step_two_in_pitch", c_uint64),
        ("step_two_input_elements", c_uint64),
    ]


class DDPeak(ctypes.Structure):
    _fields_ = [
        ("snr", c_float),
        ("mean", c_float),
        ("sigma", c_float),
        ("sample", c_uint32),
        ("width", c_uint32),
    ]


class DDDetectJob(ctypes.Structure):
    _fields_ = [
        ("d_rows", c_void_p),
        ("nr_rows", c_uint64),
        ("pitch", c_uint64),
        ("nr_samples", c_uint32),
        ("reserved", c_uint32),
        ("d_peaks", c_void_p),
    ]


class DDBatchInfo(ctypes.Structure):
    _fields_ = [
        ("jobs", c_uint32),
        ("tiles", c_uint32 * 3),
        ("workgroups", c_uint32 * 3),
        ("lds_bytes", c_uint32 * 3),
        ("units", c_uint64),
        ("compact_intermediate", c_uint32),
        ("pipeline_chunks", c_uint32),
        ("solo_jobs", c_uint32),
        ("paired_step_two", c_uint32),
    ]


class DDStreamInfo(ctypes.Structure):
    _fields_ = [
        ("slots", c_uint32),
        ("hip_streams", c_uint32),
        ("chunk_samples", c_uint64),
        ("history_samples", c_uint64),
        ("slot_bytes", c_uint64),
        ("history_bytes", c_uint64),
        ("queued", c_uint32),
        ("in_flight", c_uint32),
        ("submitted", c_uint64),
        ("completed", c_uint64),
    ]


class DDStreamEvent(ctypes.Structure):
    _fields_ = [
        ("tag", c_uint64),
        ("beam", c_uint32),
        ("slot", c_uint32),
        ("chunk", c_uint64),
        ("first_sample", ctypes.c_int64),
        ("output", c_void_p),
        ("intermediate", c_void_p),
        ("device_seconds", ctypes.c_double),
    ]


# name -> (restype, argtypes); every exported symbol of include/dedispersion_amd.h
PROTOTYPES = {
    "dd_last_error": (c_char_p, []),
    "dd_status_string": (c_char_p, [c_int]),
    "dd_abi_version": (c_uint32, []),
    "dd_plan_create": (c_int, [POINTER(c_void_p), POINTER(DDObservation), c_uint32]),
    "dd_plan_destroy": (c_int, [c_void_p]),
    "dd_plan_get_sizes": (c_int, [c_void_p, POINTER(DDSizes)]),
    "dd_plan_set_input_bits": (c_int, [c_void_p, c_uint32]),
    "dd_plan_set_conf": (c_int, [c_void_p, c_int, POINTER(DDConf)]),
    "dd_plan_get_conf": (c_int, [c_void_p, c_int, POINTER(DDConf)]),
    "dd_plan_set_zapped_channels": (c_int, [c_void_p, POINTER(c_uint32)]),
    "dd_plan_set_beam_mapping": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint32)]),
    "dd_plan_set_shifts": (c_int, [c_void_p, POINTER(c_float), POINTER(c_float)]),
    "dd_plan_get_shifts": (c_int, [c_void_p, POINTER(c_float), POINTER(c_float)]),
    "dd_plan_get_integer_shifts": (c_int, [c_void_p, c_int, POINTER(c_uint32)]),
    "dd_dedisperse": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_void_p]),
    "dd_subband_step_one": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "dd_dedisperse_f32": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_void_p]),
    "dd_plan_set_step_one_shifts": (c_int, [c_void_p, c_int]),
    "dd_subband_step_one_f32": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "dd_subband_step_two": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_void_p]),
    "dd_detect_peaks": (c_int, [c_void_p, c_uint64, c_uint32, c_uint64, c_uint32, c_void_p, c_void_p]),
    "dd_detect_peaks_batch": (c_int, [POINTER(DDDetectJob), c_uint32, c_uint32, c_void_p]),
    "dd_read_probe": (c_int, [c_void_p, c_uint64, c_void_p, c_void_p]),
    "dd_tune_time": (
        c_int,
        [c_void_p, c_int, POINTER(DDConf), c_void_p, c_void_p, c_uint32, c_uint32, c_void_p, POINTER(c_float),
         POINTER(c_float)],
    ),
    "dd_batch_create": (c_int, [POINTER(c_void_p)]),
    "dd_batch_destroy": (c_int, [c_void_p]),
    "dd_batch_add_single_step": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32]),
    "dd_batch_add_two_step": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32]),
    "dd_batch_set_conf": (c_int, [c_void_p, c_int, POINTER(DDConf)]),
    "dd_batch_get_conf": (c_int, [c_void_p, c_int, POINTER(DDConf)]),
    "dd_batch_set_compact_intermediate": (c_int, [c_void_p, c_int]),
    "dd_batch_set_pipeline": (c_int, [c_void_p, c_uint32]),
    "dd_batch_set_lds_reserve": (c_int, [c_void_p, c_int, c_uint32]),
    "dd_batch_run": (c_int, [c_void_p, c_int, c_void_p]),
    "dd_batch_run_mode": (c_int, [c_void_p, c_int, c_void_p]),
    "dd_batch_run_per_job": (c_int, [c_void_p, c_void_p]),
    "dd_batch_get_info": (c_int, [c_void_p, POINTER(DDBatchInfo)]),
    "dd_batch_run_static": (c_int, [c_void_p, c_void_p]),
    "dd_batch_set_paired_step_two": (c_int, [c_void_p, c_int]),
    "dd_batch_get_job_placement": (c_int, [c_void_p, c_uint32, POINTER(c_uint32), POINTER(c_uint32)]),
    "dd_stream_create": (c_int, [POINTER(c_void_p), c_void_p, c_uint32, c_uint64, c_uint32, c_uint32]),
    "dd_stream_destroy": (c_int, [c_void_p]),
    "dd_stream_get_info": (c_int, [c_void_p, POINTER(DDStreamInfo)]),
    "dd_stream_submit": (c_int, [c_void_p, c_uint32, c_void_p, c_uint64, c_uint64]),
    "dd_stream_progress": (c_int, [c_void_p, POINTER(DDStreamEvent), c_uint32, POINTER(c_uint32)]),
    "dd_stream_release": (c_int, [c_void_p, c_uint32]),
    "dd_stream_drain": (c_int, [c_void_p]),
    "dd_mlp_create": (c_int, [POINTER(c_void_p), c_uint32, c_uint32, c_uint32]),
    "dd_mlp_destroy": (c_int, [c_void_p]),
    "dd_mlp_param_count": (c_int, [c_void_p, POINTER(c_uint32)]),
    "dd_mlp_set_params": (c_int, [c_void_p, POINTER(c_float)]),
    "dd_mlp_get_params": (c_int, [c_void_p, POINTER(c_float)]),
    "dd_mlp_init": (c_int, [c_void_p, c_uint64]),
    "dd_mlp_forward": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_void_p]),
    "dd_mlp_loss_grad": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p]),
    "dd_mlp_adam_step": (c_int, [c_void_p, c_void_p, c_float, c_void_p]),
    "dd_mlp_train": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_float, POINTER(c_float), c_void_p]),
    "dd_synth_gaussian_u8": (c_int, [c_void_p, c_uint64, c_uint64, c_uint64, c_float, c_float, c_uint64, c_void_p]),
    "dd_synth_uniform_u8": (c_int, [c_void_p, c_uint64, c_uint64, c_uint64, c_uint64, c_void_p]),
    "dd_synth_inject_pulse": (
        c_int,
        [c_void_p, c_void_p, c_uint64, c_uint32, c_float, c_uint32, c_uint32, c_uint32, c_void_p],
    ),
}

_lib = None


class DDError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def load() -> ctypes.CDLL:
    """Loads libdedisp_amd.so once and declares every prototype."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `make -C dedispersion_amd` or __graft_entry__.build()"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != 0:
        raise DDError(status, load().dd_last_error().decode())

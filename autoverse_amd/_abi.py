"""ctypes mirror of include/flowlenia_hip.h and the loader of the HIP library.

The product library is ``autoverse_amd/libflowlenia_hip.so`` built in-tree by
``autoverse_amd/build.py``.  There is no fallback: if the library is missing the
loader raises, so nothing silently runs on the CPU.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_uint32, c_uint64, c_void_p

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libflowlenia_hip.so")

# hp_status (flowlenia_hip.h); the first five mirror ConfigError (config.rs:169-180)
HP_OK = 0
HP_ERR_INVALID_DIMENSIONS = 1
HP_ERR_INVALID_CHANNELS = 2
HP_ERR_INVALID_TIME_STEP = 3
HP_ERR_INVALID_KERNEL_RADIUS = 4
HP_ERR_INVALID_CHANNEL_INDEX = 5
HP_ERR_NO_DEVICE = 6
HP_ERR_HIP = 7
HP_ERR_SHAPE = 8
HP_ERR_NULL = 9
HP_ERR_NONFINITE = 10
HP_ERR_UNSUPPORTED_SIZE = 11
HP_ERR_IO = 12
HP_ERR_PARSE = 13

STATUS_NAMES = {
    0: "HP_OK", 1: "HP_ERR_INVALID_DIMENSIONS", 2: "HP_ERR_INVALID_CHANNELS",
    3: "HP_ERR_INVALID_TIME_STEP", 4: "HP_ERR_INVALID_KERNEL_RADIUS",
    5: "HP_ERR_INVALID_CHANNEL_INDEX", 6: "HP_ERR_NO_DEVICE", 7: "HP_ERR_HIP", 8: "HP_ERR_SHAPE",
    9: "HP_ERR_NULL", 10: "HP_ERR_NONFINITE", 11: "HP_ERR_UNSUPPORTED_SIZE", 12: "HP_ERR_IO",
    13: "HP_ERR_PARSE",
}

# hp_pattern_kind
PATTERN_KINDS = {
    "GaussianBlob": 0, "MultiBlob": 1, "Ring": 2, "GaussianSphere": 3, "MultiSphere": 4,
    "Shell": 5, "Torus3D": 6, "Noise": 7, "Custom": 8, "Custom3D": 9,
}

# hp_kernel_class
KERNEL_CLASSES = ["rows_fwd", "cols_fwd", "spectral", "cols_inv", "rows_inv", "stencil", "general"]
HP_K_NUM_CLASSES = len(KERNEL_CLASSES)


class HpRing(ctypes.Structure):
    _fields_ = [("amplitude", c_float), ("distance", c_float), ("width", c_float)]


class HpKernel(ctypes.Structure):
    _fields_ = [("radius", c_float), ("weight", c_float), ("mu", c_float), ("sigma", c_float),
                ("source_channel", c_uint32), ("target_channel", c_uint32),
                ("ring_offset", c_uint32), ("ring_count", c_uint32)]


class HpFlow(ctypes.Structure):
    _fields_ = [("beta_a", c_float), ("n", c_float), ("distribution_size", c_float)]


class HpConfig(ctypes.Structure):
    _fields_ = [("width", c_uint32), ("height", c_uint32), ("depth", c_uint32), ("channels", c_uint32),
                ("dt", c_float), ("kernel_radius", c_uint32), ("flow", HpFlow),
                ("num_kernels", c_uint32), ("kernels", POINTER(HpKernel)),
                ("num_rings", c_uint32), ("rings", POINTER(HpRing))]


class HpBlob(ctypes.Structure):
    _fields_ = [("center", c_float * 3), ("radius", c_float), ("amplitude", c_float), ("channel", c_uint32)]


class HpCustomValue(ctypes.Structure):
    _fields_ = [("x", c_uint32), ("y", c_uint32), ("z", c_uint32), ("channel", c_uint32), ("value", c_float)]


class HpPlan(ctypes.Structure):
    """hp_plan: how the step runs on the device (see include/flowlenia_hip.h)."""
    _fields_ = [("fft_path", c_int32), ("spectral_group", c_int32), ("radix_order", c_int32),
                ("column_split", c_int32), ("stencil_tile", c_int32)]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}

    @classmethod
    def from_dict(cls, d: dict) -> "HpPlan":
        p = cls(0, -1, 0, 1, 0)
        for k, v in d.items():
            setattr(p, k, int(v))
        return p


class HpSeed(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("center", c_float * 3), ("radius", c_float),
                ("inner_radius", c_float), ("outer_radius", c_float),
                ("major_radius", c_float), ("minor_radius", c_float), ("amplitude", c_float),
                ("channel", c_int32), ("noise_seed", c_uint64),
                ("num_blobs", c_uint32), ("blobs", POINTER(HpBlob)),
                ("num_values", c_uint32), ("values", POINTER(HpCustomValue))]


class HpEvalConfig(ctypes.Structure):
    _fields_ = [("steps", c_uint64), ("warmup_steps", c_uint64), ("sample_interval", c_uint64)]


class HpMetric(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("weight", c_float), ("period", c_uint64), ("tolerance", c_float),
                ("min_displacement", c_float), ("max_period", c_uint64), ("threshold", c_float)]


class HpBehavior(ctypes.Structure):
    _fields_ = [("final_mass", c_float), ("initial_mass", c_float), ("total_displacement", c_float),
                ("final_radius", c_float), ("active_cells", c_uint64), ("max_activation", c_float)]


class HpSample(ctypes.Structure):
    _fields_ = [("center_x", c_float), ("center_y", c_float), ("radius", c_float), ("mass", c_float),
                ("max_value", c_float), ("active_cells", c_uint32)]


class HpEmbedding(ctypes.Structure):
    _fields_ = [("enabled", c_int32), ("mixing_temperature", c_float), ("linear_mixing", c_int32)]


class HpCellParams(ctypes.Structure):
    _fields_ = [("mu", c_float), ("sigma", c_float), ("weight", c_float), ("beta_a", c_float), ("n", c_float)]


class HpRecorderConfig(ctypes.Structure):
    _fields_ = [("compression", c_int32), ("frame_skip", c_uint32), ("max_frames", c_uint64)]


class HpAnimationStats(ctypes.Structure):
    _fields_ = [("frame_count", c_uint64), ("total_bytes", c_uint64), ("average_frame_size", c_uint64),
                ("compression", c_int32)]


class HpAnimationInfo(ctypes.Structure):
    _fields_ = [("width", c_uint32), ("height", c_uint32), ("depth", c_uint32), ("channels", c_uint32),
                ("frame_count", c_uint64), ("dt", c_float), ("compression", c_int32), ("delta_encoding", c_int32)]


# hp_metric_kind, by the serde tag of FitnessMetric (src/schema/evolution.rs:319-375)
METRIC_KINDS = {
    "Persistence": 0, "Compactness": 1, "Locomotion": 2, "Periodicity": 3, "Complexity": 4,
    "MassConcentration": 5, "GliderScore": 6, "OscillatorScore": 7, "Stability": 8, "Custom": 9,
}

FloatPP = POINTER(POINTER(c_float))
FP = POINTER(c_float)

# (name, restype, argtypes) for every entry point declared in include/flowlenia_hip.h
EXPORTS = [
    ("hp_validate_config", c_int, [POINTER(HpConfig), POINTER(c_uint32), POINTER(c_uint32)]),
    ("hp_seed_generate", c_int, [POINTER(HpSeed), c_uint32, c_uint32, c_uint32, c_uint32, FloatPP]),
    ("hp_create", c_int, [POINTER(HpConfig), c_int, POINTER(c_void_p)]),
    ("hp_destroy", None, [c_void_p]),
    ("hp_last_error", c_char_p, [c_void_p]),
    ("hp_last_global_error", c_char_p, []),
    ("hp_get_config", c_int, [c_void_p, POINTER(HpConfig)]),
    ("hp_set_state", c_int, [c_void_p, FloatPP, c_float, c_uint64]),
    ("hp_get_state", c_int, [c_void_p, FloatPP, POINTER(c_float), POINTER(c_uint64)]),
    ("hp_step", c_int, [c_void_p, FloatPP, POINTER(c_float), POINTER(c_uint64)]),
    ("hp_run", c_int, [c_void_p, FloatPP, POINTER(c_float), POINTER(c_uint64), c_uint64]),
    ("hp_advance", c_int, [c_void_p, c_uint64]),
    ("hp_mass_f64", c_int, [c_void_p, POINTER(c_double), POINTER(c_double)]),
    ("hp_stats", c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint64)]),
    ("hp_set_conservation", c_int, [c_void_p, c_int]),
    ("hp_get_conservation", c_int, [c_void_p, POINTER(c_int)]),
    ("hp_profile_enable", c_int, [c_void_p, c_int]),
    ("hp_profile_read", c_int, [c_void_p, POINTER(c_double), POINTER(c_uint64)]),
    ("hp_profile_repeats", c_int, [c_void_p, c_uint32]),
    ("hp_kernel_bytes", c_int, [c_void_p, POINTER(c_double)]),
    ("hp_advect_mass_2d", c_int, [c_int, FP, FP, FP, FP, c_uint32, c_uint32, c_float, c_float]),
    ("hp_advect_mass_2d_noalloc", c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32,
                                          c_float, c_float]),
    ("hp_advect_mass_3d", c_int, [c_int, FP, FP, FP, FP, FP, c_uint32, c_uint32, c_uint32, c_float, c_float]),
    ("hp_sobel_gradient_2d", c_int, [c_int, FP, FP, FP, c_uint32, c_uint32]),
    ("hp_compute_flow_2d", c_int, [c_int, FP, FP, FP, FP, FP, c_uint32, HpFlow, FP, FP]),
    ("hp_growth_accumulate", c_int, [c_int, FP, FP, c_uint32, c_float, c_float, c_float]),
    ("hp_stencil_with_affinity", c_int, [c_void_p, FloatPP, FloatPP, FloatPP]),
    ("hp_affinity", c_int, [c_void_p, FloatPP, FloatPP]),
    ("hp_batch_create", c_int, [POINTER(HpConfig), c_uint32, c_int, POINTER(c_void_p)]),
    ("hp_batch_destroy", None, [c_void_p]),
    ("hp_batch_last_error", c_char_p, [c_void_p]),
    ("hp_batch_set_state", c_int, [c_void_p, c_uint32, FloatPP]),
    ("hp_batch_get_state", c_int, [c_void_p, c_uint32, FloatPP, POINTER(c_float), POINTER(c_uint64)]),
    ("hp_batch_advance", c_int, [c_void_p, c_uint64]),
    ("hp_batch_centroids", c_int, [c_void_p, POINTER(c_double), POINTER(c_double)]),
    ("hp_batch_mass_f64", c_int, [c_void_p, POINTER(c_double)]),
    ("hp_batch_stats", c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint64)]),
    ("hp_get_plan", c_int, [c_void_p, POINTER(HpPlan)]),
    ("hp_set_plan", c_int, [c_void_p, POINTER(HpPlan)]),
    ("hp_autotune", c_int, [c_void_p, c_uint32, POINTER(HpPlan), c_uint32, POINTER(HpPlan), POINTER(c_double),
                            POINTER(c_uint32)]),
    ("hp_batch_get_plan", c_int, [c_void_p, POINTER(HpPlan)]),
    ("hp_batch_set_plan", c_int, [c_void_p, POINTER(HpPlan)]),
    ("hp_batch_set_step_mode", c_int, [c_void_p, c_int32]),
    ("hp_batch_step_mode", c_int, [c_void_p, POINTER(c_int32), POINTER(c_uint64)]),
    ("hp_batch_autotune", c_int, [c_void_p, c_uint32, POINTER(HpPlan), c_uint32, POINTER(HpPlan),
                                  POINTER(c_double), POINTER(c_uint32)]),
    ("hp_eval_sample_count", c_int, [POINTER(HpEvalConfig), POINTER(c_uint32)]),
    ("hp_batch_evaluate", c_int, [c_void_p, POINTER(HpEvalConfig), POINTER(HpMetric), c_uint32, c_int, FP, FP,
                                  POINTER(HpBehavior), POINTER(HpSample)]),
    ("hp_config_from_json", c_int, [c_char_p, POINTER(c_void_p)]),
    ("hp_seed_from_json", c_int, [c_char_p, POINTER(c_void_p)]),
    ("hp_config_doc_config", POINTER(HpConfig), [c_void_p]),
    ("hp_config_doc_embedding", c_int, [c_void_p, POINTER(HpEmbedding)]),
    ("hp_config_doc_free", None, [c_void_p]),
    ("hp_seed_doc_seed", POINTER(HpSeed), [c_void_p]),
    ("hp_seed_doc_free", None, [c_void_p]),
    ("hp_load_config_and_seed", c_int, [c_char_p, POINTER(c_void_p), POINTER(c_void_p)]),
    ("hp_json_last_error", c_char_p, []),
    ("hp_recorder_open", c_int, [c_char_p, POINTER(HpConfig), POINTER(HpRecorderConfig), POINTER(c_void_p)]),
    ("hp_recorder_record_ctx", c_int, [c_void_p, c_void_p, POINTER(c_int)]),
    ("hp_recorder_record_member", c_int, [c_void_p, c_void_p, c_uint32, POINTER(c_int)]),
    ("hp_recorder_record_host", c_int, [c_void_p, FloatPP, POINTER(c_int)]),
    ("hp_recorder_finalize", c_int, [c_void_p, POINTER(HpAnimationStats)]),
    ("hp_animation_last_error", c_char_p, []),
    ("hp_player_open", c_int, [c_char_p, POINTER(c_void_p)]),
    ("hp_player_info", c_int, [c_void_p, POINTER(HpAnimationInfo)]),
    ("hp_player_read_frame", c_int, [c_void_p, c_uint64, FloatPP]),
    ("hp_player_load_ctx", c_int, [c_void_p, c_uint64, c_void_p, c_float, c_uint64]),
    ("hp_player_close", None, [c_void_p]),
    # parameter grids travel as one float pointer per channel (W*H cells of 5 floats)
    ("hp_emb_create", c_int, [POINTER(HpConfig), POINTER(HpEmbedding), c_int, POINTER(c_void_p)]),
    ("hp_emb_destroy", None, [c_void_p]),
    ("hp_emb_last_error", c_char_p, [c_void_p]),
    ("hp_emb_default_params", c_int, [POINTER(HpConfig), POINTER(HpCellParams)]),
    ("hp_emb_set_state", c_int, [c_void_p, FloatPP, FloatPP, c_float, c_uint64]),
    ("hp_emb_get_state", c_int, [c_void_p, FloatPP, FloatPP, POINTER(c_float), POINTER(c_uint64)]),
    ("hp_emb_step", c_int, [c_void_p, FloatPP, FloatPP, POINTER(c_float), POINTER(c_uint64)]),
    ("hp_emb_run", c_int, [c_void_p, FloatPP, FloatPP, POINTER(c_float), POINTER(c_uint64), c_uint64]),
    ("hp_emb_advance", c_int, [c_void_p, c_uint64]),
    ("hp_emb_mass_f64", c_int, [c_void_p, POINTER(c_double), POINTER(c_double)]),
    ("hp_emb_get_next_params", c_int, [c_void_p, FloatPP]),
    ("hp_emb_set_next_params", c_int, [c_void_p, FloatPP]),
    ("hp_emb_last_radius", c_int, [c_void_p, POINTER(c_int32)]),
    ("hp_emb_affinity", c_int, [c_void_p, FloatPP, FloatPP, FloatPP]),
    ("hp_emb_stencil", c_int, [c_void_p, FloatPP, FloatPP, FloatPP, FloatPP, FloatPP]),
    ("hp_advect_mass_and_params", c_int, [FP, FP, FP, FP, POINTER(HpEmbedding), c_float, c_float, c_uint32,
                                          c_uint32, FP, FP, c_int]),
    ("hp_batch_theta_count", c_int, [c_void_p, c_uint32, POINTER(c_uint32)]),
    ("hp_batch_tape_forward", c_int, [c_void_p, c_uint64, c_uint32, POINTER(c_uint32)]),
    ("hp_batch_tape_backward", c_int, [c_void_p, FloatPP, FloatPP, POINTER(c_double)]),
]

_LIB = None


class HipLibraryMissing(RuntimeError):
    pass


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Load the HIP library (no fallback).  Raises HipLibraryMissing if it was never built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise HipLibraryMissing(
            f"{p} not found: build it with `python -m autoverse_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(p)
    for name, res, args in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


class HipError(RuntimeError):
    def __init__(self, status: int, message: str = ""):
        self.status = status
        name = STATUS_NAMES.get(status, f"status {status}")
        super().__init__(f"{name}: {message}" if message else name)


def float_ptr_array(arrays) -> ctypes.Array:
    """Array of float* built from C-contiguous float32 numpy arrays (kept alive by caller)."""
    arr = (FP * len(arrays))()
    for i, a in enumerate(arrays):
        assert a.dtype.name == "float32" and a.flags["C_CONTIGUOUS"], "float32 C-contiguous required"
        arr[i] = a.ctypes.data_as(FP)
    return arr


def autotune_call(fn, ptr, steps: int):
    """Shared driver of hp_autotune / hp_batch_autotune: (status, best plan dict, [(plan dict,
    microseconds per step or None if the plan disagreed with the first)])."""
    cap = 64
    best = HpPlan()
    tried = (HpPlan * cap)()
    us = (c_double * cap)()
    n = c_uint32(0)
    st = fn(ptr, int(steps), ctypes.byref(best), cap, tried, us, ctypes.byref(n))
    table = [(tried[i].as_dict(), us[i] if us[i] >= 0 else None) for i in range(min(n.value, cap))]
    return st, best.as_dict(), table

"""ctypes binding of libcrdfft.so (include/crdfft.h, include/crdfft_runtime.h).

The library is built in-tree (crd_amd/libcrdfft.so, by crd_amd/csrc/Makefile). Loading fails loudly when it is
missing: there is no CPU fallback for any transform.
"""
import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcrdfft.so")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")

# status codes (crdfft_status)
OK, INVALID_SIZE, INVALID_ARGUMENT, OUT_OF_MEMORY, HIP_ERROR, SELF_CHECK_FAILED, UNSUPPORTED, NO_DEVICE = range(8)
CHECK_NOT_RUN, CHECK_PASSED, CHECK_DEMOTED = range(3)  # crdfft_check_status
FORWARD, INVERSE = 0, 1
C2C, R2C, C2R = 0, 1, 2
FP32, SPLIT_BF16, SPLIT_FP16 = 0, 1, 2
ALGO_COPY, ALGO_SINGLE, ALGO_FOUR_STEP, ALGO_REFERENCE, ALGO_BLUESTEIN, ALGO_FUSED, ALGO_COOP = 0, 1, 2, 3, 4, 5, 6
HINT_COOP_SOLO = 16
FLAG_NONE, FLAG_NO_SELF_CHECK, FLAG_NORMALIZE_INVERSE = 0, 1, 2
FLAG_SEPARATE_PASSES = 4
FLAG_CHIRP_Z = 8
HINTS_NONE = 0x100
HINT_TWO_PIPELINES = 0x20
DCT2, DCT3, DST2, DST3 = 0, 1, 2, 3
ND_TRANSPOSE_ALL = 1
WINDOW_BOXCAR, WINDOW_HANN = 0, 1
SPECTRAL_DETREND = 1


class CrdfftError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = status_string(status) if _lib is not None else str(status)
        if status == HIP_ERROR and _lib is not None:
            msg += f" (hipError {_lib.crdfft_last_hip_error()})"
        super().__init__(f"{what}: {msg}" if what else msg)


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint64),
        ("batch", ctypes.c_uint64),
        ("type", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("algorithm", ctypes.c_int32),
        ("complex_n", ctypes.c_uint64),
        ("n1", ctypes.c_uint64),
        ("n2", ctypes.c_uint64),
        ("e", ctypes.c_int32),
        ("e1", ctypes.c_int32),
        ("tile1", ctypes.c_int32),
        ("e2", ctypes.c_int32),
        ("tile2", ctypes.c_int32),
        ("chunk", ctypes.c_uint64),
        ("workspace_bytes", ctypes.c_uint64),
        ("table_bytes", ctypes.c_uint64),
        ("hints", ctypes.c_uint32),
        ("check", ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [
        ("n1", ctypes.c_uint64),
        ("e", ctypes.c_int32),
        ("e1", ctypes.c_int32),
        ("tile1", ctypes.c_int32),
        ("e2", ctypes.c_int32),
        ("tile2", ctypes.c_int32),
        ("chunk", ctypes.c_uint64),
        ("hints", ctypes.c_uint32),
        ("algorithm", ctypes.c_int32),
    ]


class PhaseTimes(ctypes.Structure):
    _fields_ = [
        ("ms", ctypes.c_double * 4),
        ("launches", ctypes.c_uint64 * 4),
        ("points", ctypes.c_uint64 * 4),
    ]


class DeviceProps(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 256),
        ("arch", ctypes.c_char * 64),
        ("compute_units", ctypes.c_int32),
        ("clock_khz", ctypes.c_int32),
        ("memory_clock_khz", ctypes.c_int32),
        ("memory_bus_width", ctypes.c_int32),
        ("global_mem_bytes", ctypes.c_uint64),
        ("l2_bytes", ctypes.c_uint64),
        ("max_lds_per_block", ctypes.c_int32),
    ]


MAX_STAGES = 16
FEATURES = 32


class RadixPlan(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("nstages", ctypes.c_uint32), ("radix", ctypes.c_uint32 * MAX_STAGES),
                ("group", ctypes.c_uint32)]

    def radices(self):
        return [int(self.radix[i]) for i in range(self.nstages)]

    def as_dict(self):
        return {"n": int(self.n), "radices": self.radices(), "group": int(self.group)}


class ManyInfo(ctypes.Structure):
    _fields_ = [("transforms", ctypes.c_uint64), ("lengths", ctypes.c_uint64), ("tiles", ctypes.c_uint64),
                ("points", ctypes.c_uint64), ("workgroups", ctypes.c_uint32), ("imbalance", ctypes.c_double)]


class TuneRecord(ctypes.Structure):
    _fields_ = [("plan", RadixPlan), ("candidate", ctypes.c_uint32), ("transforms", ctypes.c_uint32),
                ("us_per_transform", ctypes.c_double)]


class PlanRecord(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("batch", ctypes.c_uint64), ("type", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("algorithm", ctypes.c_int32), ("ndim", ctypes.c_uint32),
                ("dims", ctypes.c_uint64 * 3), ("stride", ctypes.c_uint64), ("tuning", Tuning),
                ("ms", ctypes.c_double)]


_lib = None

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int32
_u32 = ctypes.c_uint32
_SIGS = {
    "crdfft_plan_create": (_i32, [ctypes.POINTER(_vp), _u64, _u64, _u64, _u64, _i32, _i32, _u32]),
    "crdfft_plan_create_tuned": (_i32, [ctypes.POINTER(_vp), _u64, _u64, _u64, _u64, _i32, _i32, _u32,
                                        ctypes.POINTER(Tuning)]),
    "crdfft_plan_destroy": (_i32, [_vp]),
    "crdfft_graph_begin": (_i32, [_vp]),
    "crdfft_graph_end": (_i32, [_vp, ctypes.POINTER(_vp)]),
    "crdfft_graph_launch": (_i32, [_vp, _vp]),
    "crdfft_graph_destroy": (_i32, [_vp]),
    "crdfft_plan_coop_status": (_i32, [_vp, _vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "crdfft_bluestein_plan_create": (_i32, [ctypes.POINTER(_vp), _u64, _u64, _u64, _u64, _u32]),
    "crdfft_execute": (_i32, [_vp, _vp, _vp, _i32, _vp]),
    "crdfft_execute_batched": (_i32, [_vp, _vp, _u64, _i32, _vp]),
    "crdfft_execute_reference": (_i32, [_vp, _vp, _vp, _i32, _vp]),
    "crdfft_execute_noalloc": (_i32, [_vp, _vp, _vp, _i32, _vp]),
    "crdfft_noalloc_scratch_bytes": (_u64, []),
    "crdfft_self_check": (_i32, [_vp]),
    "crdfft_execute_profiled": (_i32, [_vp, _vp, _vp, _i32, _vp, ctypes.POINTER(PhaseTimes)]),
    "crdfft_plan_get_info": (_i32, [_vp, ctypes.POINTER(PlanInfo)]),
    "crdfft_nd_plan_create": (_i32, [ctypes.POINTER(_vp), ctypes.POINTER(_u64), _u32, _u64, _u32]),
    "crdfft_nd_execute": (_i32, [_vp, _vp, _i32, _vp]),
    "crdfft_nd_plan_destroy": (_i32, [_vp]),
    "crdfft_plan_describe": (_i32, [_u64, _u64, _i32, _i32, ctypes.POINTER(PlanInfo)]),
    "crdfft_plan_describe_tuned": (_i32, [_u64, _u64, _i32, _i32, ctypes.POINTER(Tuning), ctypes.POINTER(PlanInfo)]),
    "crdfft_status_string": (ctypes.c_char_p, [_i32]),
    "crdfft_last_hip_error": (_i32, []),
    "crdfft_version": (ctypes.c_char_p, []),
    "crdfft_device_count": (_i32, [ctypes.POINTER(_i32)]),
    "crdfft_device_props_get": (_i32, [_i32, ctypes.POINTER(DeviceProps)]),
    "crdfft_malloc": (_i32, [ctypes.POINTER(_vp), _u64]),
    "crdfft_free": (_i32, [_vp]),
    "crdfft_pointer_is_device": (_i32, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "crdfft_memcpy_h2d": (_i32, [_vp, _vp, _u64]),
    "crdfft_memcpy_d2h": (_i32, [_vp, _vp, _u64]),
    "crdfft_memcpy_d2d": (_i32, [_vp, _vp, _u64, _vp]),
    "crdfft_fill_bytes": (_i32, [_vp, _i32, _u64, _vp]),
    "crdfft_stream_copy": (_i32, [_vp, _vp, _u64, _vp]),
    "crdfft_stream_create": (_i32, [ctypes.POINTER(_vp)]),
    "crdfft_stream_destroy": (_i32, [_vp]),
    "crdfft_stream_synchronize": (_i32, [_vp]),
    "crdfft_device_synchronize": (_i32, []),
    "crdfft_event_create": (_i32, [ctypes.POINTER(_vp)]),
    "crdfft_event_destroy": (_i32, [_vp]),
    "crdfft_event_record": (_i32, [_vp, _vp]),
    "crdfft_event_synchronize": (_i32, [_vp]),
    "crdfft_event_elapsed_ms": (_i32, [ctypes.POINTER(ctypes.c_float), _vp, _vp]),
    # mixed-size persistent launch and its autotuner
    "crdfft_radix_candidates": (_i32, [_u64, ctypes.POINTER(RadixPlan), _u32, ctypes.POINTER(_u32)]),
    "crdfft_many_plan_create": (_i32, [ctypes.POINTER(_vp), ctypes.POINTER(_u64), ctypes.POINTER(_u64), _u64,
                                       ctypes.POINTER(RadixPlan), _u64, _u32]),
    "crdfft_many_plan_destroy": (_i32, [_vp]),
    "crdfft_many_plan_get_info": (_i32, [_vp, ctypes.POINTER(ManyInfo)]),
    "crdfft_many_execute": (_i32, [_vp, _vp, _vp, _i32, _vp]),
    "crdfft_many_graph_capture": (_i32, [_vp, _vp, _vp, _i32, _vp]),
    "crdfft_many_graph_launch": (_i32, [_vp, _vp]),
    "crdfft_dct_plan_create": (_i32, [ctypes.POINTER(_vp), _u64, _u64, _u64, _u64, _u32]),
    "crdfft_dct_plan_destroy": (_i32, [_vp]),
    "crdfft_dct_plan_get_info": (_i32, [_vp, ctypes.POINTER(PlanInfo)]),
    "crdfft_dct_execute": (_i32, [_vp, _i32, _vp, _vp, _vp]),
    "crdfft_spectral_plan_create": (_i32, [ctypes.POINTER(_vp), _u64, _u64, _i32, _u32]),
    "crdfft_spectral_plan_destroy": (_i32, [_vp]),
    "crdfft_spectral_frames": (_u64, [_vp, _u64]),
    "crdfft_stft": (_i32, [_vp, _vp, _u64, _vp, _vp]),
    "crdfft_spectrogram": (_i32, [_vp, _vp, _u64, ctypes.c_double, _vp, _vp]),
    "crdfft_welch": (_i32, [_vp, _vp, _u64, ctypes.c_double, _vp, _vp]),
    "crdfft_many_autotune": (_i32, [ctypes.POINTER(_u64), _u64, _u32, ctypes.POINTER(RadixPlan), _u64,
                                    ctypes.POINTER(_u64), ctypes.POINTER(TuneRecord), _u64, ctypes.POINTER(_u64)]),
    # single-length autotuner, records and features
    "crdfft_plan_candidates": (_i32, [_u64, _u64, _i32, ctypes.POINTER(Tuning), _u32, ctypes.POINTER(_u32)]),
    "crdfft_time_plan": (_i32, [_u64, _u64, _i32, ctypes.POINTER(Tuning), _u32, ctypes.POINTER(PlanRecord)]),
    "crdfft_time_nd": (_i32, [ctypes.POINTER(_u64), _u32, _u64, _u32, _u32, ctypes.POINTER(PlanRecord)]),
    "crdfft_autotune_plan": (_i32, [_u64, _u64, _i32, _u32, ctypes.POINTER(Tuning), ctypes.POINTER(PlanRecord),
                                    _u32, ctypes.POINTER(_u32)]),
    "crdfft_plan_features": (_i32, [ctypes.POINTER(PlanRecord), _u64, ctypes.POINTER(ctypes.c_float)]),
    # the cost model
    "crdfft_mlp_create": (_i32, [ctypes.POINTER(_vp), _u32, _u32, _u64]),
    "crdfft_mlp_destroy": (_i32, [_vp]),
    "crdfft_mlp_param_count": (_i32, [_vp, ctypes.POINTER(_u64)]),
    "crdfft_mlp_get_params": (_i32, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "crdfft_mlp_set_params": (_i32, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "crdfft_mlp_predict": (_i32, [_vp, ctypes.POINTER(ctypes.c_float), _u64, ctypes.POINTER(ctypes.c_float)]),
    "crdfft_mlp_loss_grad": (_i32, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _u64,
                                    ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "crdfft_mlp_train": (_i32, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _u64, _u32,
                                _u32, ctypes.c_float, _u64, ctypes.POINTER(ctypes.c_float)]),
}


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C crd_amd/csrc` (or __graft_entry__.build()); "
                "crd_amd has no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def status_string(status):
    return lib().crdfft_status_string(status).decode()


def check(status, what=""):
    if status != OK:
        raise CrdfftError(status, what)


def header_symbols():
    """Every CRDFFT_API function declared under include/."""
    names = []
    for fn in sorted(os.listdir(INCLUDE_DIR)):
        if fn.endswith(".h"):
            text = open(os.path.join(INCLUDE_DIR, fn)).read()
            names += re.findall(r"CRDFFT_API\s+[^;(]*?\b(crdfft_\w+)\s*\(", text)
    return names

"""ctypes binding of libmuscle_hip.so (declared in include/muscle_hip.h).

The library is the product: there is no CPU fallback.  If the shared
library is missing this module raises ImportError at load time; if no GPU is
present, creating a handle fails with the HIP error the runtime reports.
torch is imported first so that its HIP runtime (libamdhip64.so.7) is the
one the library binds to — device pointers from torch's allocator and the
library's kernels then live in one runtime.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (loads the HIP runtime before the library)

from .errors import ArgumentError, DimensionMismatch, MuscleHipError, UnsupportedError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libmuscle_hip.so"))

# status codes (muscle_hip_status_t)
OK, INVALID_ARGUMENT, DIM_MISMATCH, UNSUPPORTED, HIP_ERROR, OUT_OF_MEMORY, NOT_READY, REJECTED = range(8)
# dtypes (muscle_hip_dtype_t)
F32, F64, C32, C64 = range(4)
# precision modes (muscle_hip_precision_t)
PRECISION = {"native": 0, "fp32": 1, "bf16x3": 2, "fp16x2": 3, "ozaki_int8": 4}

DTYPE_CODE = {torch.float32: F32, torch.float64: F64, torch.complex64: C32, torch.complex128: C64}
CODE_DTYPE = {v: k for k, v in DTYPE_CODE.items()}


class TensorDesc(ctypes.Structure):
    _fields_ = [
        ("dtype", ctypes.c_int32),
        ("rank", ctypes.c_int32),
        ("data", ctypes.c_void_p),
        ("modes", ctypes.POINTER(ctypes.c_int32)),
        ("extents", ctypes.POINTER(ctypes.c_int64)),
        ("strides", ctypes.POINTER(ctypes.c_int64)),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("pairs", ctypes.c_int64),
        ("items", ctypes.c_int64),
        ("workgroups", ctypes.c_int32),
        ("graph", ctypes.c_int32),
        ("flops", ctypes.c_double),
        ("bytes", ctypes.c_double),
        ("table_bytes", ctypes.c_int64),
        ("shallow_items", ctypes.c_int64),
        ("split_items", ctypes.c_int64),
        ("skinny_items", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchPairInfo(ctypes.Structure):
    _fields_ = [
        ("route", ctypes.c_int32),
        ("tile", ctypes.c_int32),
        ("size_class", ctypes.c_int32),
        ("ksplit", ctypes.c_int32),
        ("m", ctypes.c_int64),
        ("n", ctypes.c_int64),
        ("k", ctypes.c_int64),
        ("batch", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# batched pair routes (muscle_hip_pair_route_t)
PAIR_ROUTES = {0: "empty", 1: "tiles", 2: "split_k", 3: "shallow", 4: "solo", 5: "skinny"}


class JobDesc(ctypes.Structure):
    _fields_ = [
        ("n_leaves", ctypes.c_int32),
        ("n_nodes", ctypes.c_int32),
        ("leaves", ctypes.POINTER(TensorDesc)),
        ("lhs", ctypes.POINTER(ctypes.c_int32)),
        ("rhs", ctypes.POINTER(ctypes.c_int32)),
        ("node_off", ctypes.POINTER(ctypes.c_int32)),
        ("node_modes", ctypes.POINTER(ctypes.c_int32)),
        ("n_labels", ctypes.c_int32),
        ("label_extents", ctypes.POINTER(ctypes.c_int64)),
    ]


class JobResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("result", ctypes.c_void_p),
        ("result_elems", ctypes.c_int64),
        ("submit_s", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
        ("ticks", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("submitted", "completed", "rejected", "ticks", "launched_pairs",
                                              "pool_bytes", "pool_peak_bytes", "active_jobs", "waiting_jobs")]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Info(ctypes.Structure):
    _fields_ = [
        ("kernel", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("slices", ctypes.c_int32),
        ("swapped", ctypes.c_int32),
        ("m", ctypes.c_int64),
        ("n", ctypes.c_int64),
        ("k", ctypes.c_int64),
        ("batch", ctypes.c_int64),
        ("flops", ctypes.c_double),
        ("bytes", ctypes.c_double),
        ("err_bound", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libmuscle_hip.so not found at {LIB_PATH}; build it with `make -C muscle.jl_amd/csrc` "
            "(the HIP backend has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    H = ctypes.c_void_p
    T = ctypes.POINTER(TensorDesc)
    I = ctypes.POINTER(Info)
    sig = {
        "muscle_hip_version": ([], ctypes.c_int),
        "muscle_hip_create": ([ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "muscle_hip_destroy": ([H], ctypes.c_int),
        "muscle_hip_set_stream": ([H, ctypes.c_void_p], ctypes.c_int),
        "muscle_hip_synchronize": ([H], ctypes.c_int),
        "muscle_hip_measure_bandwidth": ([H, ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "muscle_hip_set_option": ([H, ctypes.c_int, ctypes.c_int64], ctypes.c_int),
        "muscle_hip_status_string": ([ctypes.c_int], ctypes.c_char_p),
        "muscle_hip_last_error": ([], ctypes.c_char_p),
        "muscle_hip_plan_cache_size": ([H], ctypes.c_int64),
        "muscle_hip_plan_cache_clear": ([H], ctypes.c_int),
        "muscle_hip_binary_einsum": ([H, T, T, T, ctypes.c_int, ctypes.c_double, I], ctypes.c_int),
        "muscle_hip_binary_einsum_noalloc": ([H, T, T, T, I], ctypes.c_int),
        "muscle_hip_hadamard": ([H, T, T, T, I], ctypes.c_int),
        "muscle_hip_unary_einsum": ([H, T, T, I], ctypes.c_int),
        "muscle_hip_batch_create": ([H, ctypes.c_int64, T, T, T, ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "muscle_hip_batch_run": ([H, ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
        "muscle_hip_batch_get_info": ([ctypes.c_void_p, ctypes.POINTER(BatchInfo)], ctypes.c_int),
        "muscle_hip_batch_get_pair_info": ([ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(BatchPairInfo)],
                                           ctypes.c_int),
        "muscle_hip_batch_destroy": ([ctypes.c_void_p], ctypes.c_int),
        "muscle_hip_batch_autotune": ([H, ctypes.c_void_p, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "muscle_hip_tuning_save": ([H, ctypes.c_char_p], ctypes.c_int),
        "muscle_hip_tuning_load": ([H, ctypes.c_char_p], ctypes.c_int),
        "muscle_hip_tuning_class": ([ctypes.c_int64, ctypes.c_int64], ctypes.c_int),
        "muscle_hip_server_create": ([ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
                                     ctypes.c_int),
        "muscle_hip_server_submit": ([ctypes.c_void_p, ctypes.POINTER(JobDesc), ctypes.POINTER(ctypes.c_int64)],
                                     ctypes.c_int),
        "muscle_hip_server_wait": ([ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(JobResult)],
                                   ctypes.c_int),
        "muscle_hip_server_copy_result": ([ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p], ctypes.c_int),
        "muscle_hip_server_release": ([ctypes.c_void_p, ctypes.c_int64], ctypes.c_int),
        "muscle_hip_server_get_stats": ([ctypes.c_void_p, ctypes.POINTER(ServerStats)], ctypes.c_int),
        "muscle_hip_server_clock": ([ctypes.c_void_p], ctypes.c_double),
        "muscle_hip_server_destroy": ([ctypes.c_void_p], ctypes.c_int),
        "muscle_hip_svd_thin": ([ctypes.c_void_p] + [ctypes.POINTER(TensorDesc)] * 4, ctypes.c_int),
        "muscle_hip_qr_thin": ([ctypes.c_void_p] + [ctypes.POINTER(TensorDesc)] * 3, ctypes.c_int),
        "muscle_hip_describe": ([T, T, T, ctypes.c_int, ctypes.c_double, I, ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


lib = _load()


def check(status: int) -> None:
    """Raise the Julia-side exception a status code maps to."""
    if status == OK:
        return
    msg = (lib.muscle_hip_last_error() or b"").decode(errors="replace")
    name = lib.muscle_hip_status_string(status).decode()
    if status == INVALID_ARGUMENT:
        raise ArgumentError(msg)
    if status == DIM_MISMATCH:
        raise DimensionMismatch(msg)
    if status == UNSUPPORTED:
        raise UnsupportedError(msg)
    raise MuscleHipError(f"{name}: {msg}", status)


class _Handles(threading.local):
    def __init__(self):
        self.by_device = {}
        self.stream = {}  # device -> the stream last set on that handle
        self.option = {}  # (device, option) -> the value last set on that handle


_handles = _Handles()


def set_stream(device: int, h, stream: int) -> None:
    """Points the handle at `stream` (skipped when it already is)."""
    if _handles.stream.get(device) != stream:
        check(lib.muscle_hip_set_stream(h, ctypes.c_void_p(stream)))
        _handles.stream[device] = stream


def set_option(device: int, h, option: int, value: int) -> None:
    """muscle_hip_set_option (skipped when the handle already holds the value)."""
    if _handles.option.get((device, option)) != value:
        check(lib.muscle_hip_set_option(h, option, value))
        _handles.option[(device, option)] = value


def handle(device: int) -> ctypes.c_void_p:
    """The calling thread's library handle for a device (created on first use)."""
    h = _handles.by_device.get(device)
    if h is None:
        with torch.cuda.device(device):
            h = ctypes.c_void_p()
            check(lib.muscle_hip_create(ctypes.byref(h)))
        _handles.by_device[device] = h
    return h


class Desc:
    """Keeps a TensorDesc and the arrays it points into alive for one call."""

    __slots__ = ("desc", "_keep")

    def __init__(self, dtype: int, data_ptr: int, modes, extents, strides):
        r = len(modes)
        n = max(r, 1)
        m = (ctypes.c_int32 * n)(*modes)
        e = (ctypes.c_int64 * n)(*extents)
        s = (ctypes.c_int64 * n)(*strides)
        self._keep = (m, e, s)
        # ctypes turns each array into the pointer field's type on assignment
        self.desc = TensorDesc(dtype, r, data_ptr, m, e, s)

    @property
    def ref(self):
        return ctypes.byref(self.desc)


class DescSlot:
    """A reusable TensorDesc with room for 64 modes, for calls that read their
    descriptors only while they run (a synchronous entry point copies what it
    keeps).  Refilled in place: about 5x cheaper than building a Desc."""

    __slots__ = ("desc", "_m", "_e", "_s")

    def __init__(self):
        self._m = (ctypes.c_int32 * 64)()
        self._e = (ctypes.c_int64 * 64)()
        self._s = (ctypes.c_int64 * 64)()
        self.desc = TensorDesc(0, 0, None, self._m, self._e, self._s)

    def fill(self, dtype: int, data_ptr: int, modes, extents, strides) -> "DescSlot":
        r = len(modes)
        self._m[:r] = modes
        self._e[:r] = extents
        self._s[:r] = strides
        d = self.desc
        d.dtype = dtype
        d.rank = r
        d.data = data_ptr
        return self

    @property
    def ref(self):
        return ctypes.byref(self.desc)


def exported_symbols() -> list[str]:
    """Names of the C ABI entry points bound here."""
    return [n for n in dir(lib) if n.startswith("muscle_hip_")]

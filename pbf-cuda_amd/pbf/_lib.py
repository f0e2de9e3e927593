"""ctypes binding of libpbf.so (the C ABI declared in include/pbf.h).

The library is loaded from this directory only (built in-tree by
``make -C pbf-cuda_amd``).  There is no fallback: if it is missing, loading
raises.  torch is imported first so that libpbf.so binds to the HIP runtime
torch already loaded (both carry the soname libamdhip64.so.7); loading the
library before torch would put two HIP runtimes in one process.
"""
from __future__ import annotations

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpbf.so")

PBF_OK = 0
PBF_ERR_INVALID_ARGUMENT = 1
PBF_ERR_CAPACITY = 2
PBF_ERR_HIP = 3
PBF_ERR_NO_DEVICE = 4
PBF_ERR_STATE = 5

PBF_LAYOUT_FLOAT3 = 3
PBF_LAYOUT_FLOAT4 = 4

PHASES = ("predict", "sort", "grid", "lambda", "delta", "velocity", "xsph", "step")
PBF_NUM_PHASES = len(PHASES)


class PbfParams(ctypes.Structure):
    """struct pbf_params: the solver fields of GUIParams (fluids/GUIParams.h:6-17)."""

    _fields_ = [
        ("niter", ctypes.c_int),
        ("pho0", ctypes.c_float),
        ("g", ctypes.c_float),
        ("h", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("lambda_eps", ctypes.c_float),
        ("delta_q", ctypes.c_float),
        ("k_corr", ctypes.c_float),
        ("n_corr", ctypes.c_float),
        ("k_boundaryDensity", ctypes.c_float),
        ("c_XSPH", ctypes.c_float),
    ]


class PbfVec3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


class PbfSweep(ctypes.Structure):
    """struct pbf_sweep: the sweeping wall of fluids/FluidSystem.cpp:36-38, :104-110."""

    _fields_ = [("ulim0", PbfVec3), ("llim0", PbfVec3), ("a_ulim", PbfVec3), ("a_llim", PbfVec3),
                ("w", ctypes.c_float), ("start_frame", ctypes.c_int)]


class PbfShapes(ctypes.Structure):
    """struct pbf_shapes: launch shapes of the neighbour passes."""

    _fields_ = [("light_block", ctypes.c_int), ("chunk_blocks", ctypes.c_int)]


class PbfGraphInfo(ctypes.Structure):
    """struct pbf_graph_info: graph recordings, in-place updates, replays."""

    _fields_ = [("recorded", ctypes.c_longlong), ("instantiated", ctypes.c_longlong),
                ("updated", ctypes.c_longlong), ("replayed", ctypes.c_longlong),
                ("live", ctypes.c_int)]


class PBFError(RuntimeError):
    """A non-zero status from libpbf.so, with the library's message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{status_name(code)}: {message}")
        self.code = code
        self.message = message


# (name, restype, argtypes) of every symbol include/pbf.h declares.
_P = ctypes.c_void_p
_SIGNATURES = [
    ("pbf_create", ctypes.c_int,
     [ctypes.POINTER(PbfParams), PbfVec3, PbfVec3, ctypes.c_int, ctypes.c_longlong, ctypes.c_int,
      ctypes.POINTER(_P)]),
    ("pbf_destroy", None, [_P]),
    ("pbf_set_params", ctypes.c_int, [_P, ctypes.POINTER(PbfParams)]),
    ("pbf_get_params", ctypes.c_int, [_P, ctypes.POINTER(PbfParams)]),
    ("pbf_set_lim", ctypes.c_int, [_P, PbfVec3, PbfVec3]),
    ("pbf_get_lim", ctypes.c_int, [_P, ctypes.POINTER(PbfVec3), ctypes.POINTER(PbfVec3)]),
    ("pbf_step", ctypes.c_int, [_P, _P, _P, _P, _P, _P, ctypes.c_int, _P]),
    ("pbf_density_error", ctypes.c_int, [_P, _P, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_double)]),
    ("pbf_sweep_box", ctypes.c_int,
     [ctypes.POINTER(PbfSweep), ctypes.c_int, ctypes.POINTER(PbfVec3), ctypes.POINTER(PbfVec3)]),
    ("pbf_run_frames", ctypes.c_int,
     [_P, _P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(PbfSweep),
      _P, ctypes.POINTER(ctypes.c_int)]),
    ("pbf_set_timing", ctypes.c_int, [_P, ctypes.c_int]),
    ("pbf_get_timing", ctypes.c_int,
     [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)]),
    ("pbf_create_batched", ctypes.c_int,
     [ctypes.POINTER(PbfParams), ctypes.c_int, ctypes.POINTER(PbfVec3), ctypes.POINTER(PbfVec3),
      ctypes.POINTER(ctypes.c_int), ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(_P)]),
    ("pbf_get_scenes", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    ("pbf_step_batched", ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    ("pbf_set_graph", ctypes.c_int, [_P, ctypes.c_int]),
    ("pbf_get_graph_info", ctypes.c_int, [_P, ctypes.POINTER(PbfGraphInfo)]),
    ("pbf_set_shapes", ctypes.c_int, [_P, ctypes.POINTER(PbfShapes)]),
    ("pbf_get_shapes", ctypes.c_int, [_P, ctypes.POINTER(PbfShapes)]),
    ("pbf_set_heavy_threshold", ctypes.c_int, [_P, ctypes.c_int]),
    ("pbf_get_heavy_threshold", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int)]),
    ("pbf_autotune", ctypes.c_int,
     [_P, _P, _P, _P, ctypes.c_int, ctypes.c_int, _P, ctypes.POINTER(PbfShapes),
      ctypes.POINTER(ctypes.c_double)]),
    ("pbf_tape_create", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)]),
    ("pbf_tape_create_ex", ctypes.c_int,
     [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)]),
    ("pbf_tape_destroy", None, [_P]),
    ("pbf_tape_gradient_scale", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int)]),
    ("pbf_tape_forward", ctypes.c_int,
     [_P, _P, _P, _P, _P, ctypes.c_int, ctypes.POINTER(PbfSweep), _P]),
    ("pbf_tape_state", ctypes.c_int,
     [_P, ctypes.c_int, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    ("pbf_tape_read", ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, _P]),
    ("pbf_tape_backward", ctypes.c_int,
     [_P, _P, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_double), _P]),
    ("pbf_step_noalloc", ctypes.c_int,
     [ctypes.POINTER(PbfParams), PbfVec3, PbfVec3, _P, _P, _P, _P, _P, ctypes.c_int, _P]),
    ("pbf_last_error", ctypes.c_char_p, [_P]),
    ("pbf_status_string", ctypes.c_char_p, [ctypes.c_int]),
    ("pbf_abi_version", ctypes.c_int, []),
]

SYMBOLS = tuple(name for name, _, _ in _SIGNATURES)

_lib = None


def load() -> ctypes.CDLL:
    """Load libpbf.so once (after torch) and declare its prototypes."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (binds libpbf.so to torch's HIP runtime)

    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `make -C pbf-cuda_amd` (or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def status_name(code: int) -> str:
    names = {0: "PBF_OK", 1: "PBF_ERR_INVALID_ARGUMENT", 2: "PBF_ERR_CAPACITY",
             3: "PBF_ERR_HIP", 4: "PBF_ERR_NO_DEVICE", 5: "PBF_ERR_STATE"}
    return names.get(code, f"PBF_ERR_{code}")


def check(rc: int, handle=None) -> None:
    """Raise PBFError for a non-zero status, with pbf_last_error's message."""
    if rc != PBF_OK:
        msg = load().pbf_last_error(handle).decode(errors="replace")
        raise PBFError(rc, msg)

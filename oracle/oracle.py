"""ctypes wrapper of oracle/liboracle.so — TEST INFRASTRUCTURE, not product.

The C++ FP64 restatement of the reference step (oracle/pbf_oracle.cpp).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as
the checker; the product library never loads it.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle.so")


class OrcParams(ctypes.Structure):
    _fields_ = [("niter", ctypes.c_int)] + [
        (f, ctypes.c_float)
        for f in ("pho0", "g", "h", "dt", "lambda_eps", "delta_q", "k_corr", "n_corr",
                  "k_boundaryDensity", "c_XSPH")
    ]


class OrcBox(ctypes.Structure):
    _fields_ = [("ulim", ctypes.c_float * 3), ("llim", ctypes.c_float * 3)]


class OrcOptions(ctypes.Structure):
    _fields_ = [("stride", ctypes.c_int), ("brute_force", ctypes.c_int),
                ("threads", ctypes.c_int), ("query_cell_f64", ctypes.c_int)]


_D = ctypes.POINTER(ctypes.c_double)
_U = ctypes.POINTER(ctypes.c_uint32)
_F = ctypes.POINTER(ctypes.c_float)


class OrcDiag(ctypes.Structure):
    _fields_ = [("x_final", _D), ("v_final", _D), ("v_pre", _D), ("lambda_", _D), ("rho", _D),
                ("dp_raw_first", _D), ("perm", _U), ("keys", _U)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    lib.orc_grid_dims.argtypes = [ctypes.POINTER(OrcBox), ctypes.c_float, ctypes.c_int * 3]
    lib.orc_grid_dims.restype = ctypes.c_int
    lib.orc_cell_keys.argtypes = [_F, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrcBox),
                                  ctypes.c_float, ctypes.c_int * 3, _U]
    lib.orc_cell_keys.restype = None
    lib.orc_stable_order.argtypes = [_U, ctypes.c_int, _U]
    lib.orc_stable_order.restype = None
    lib.orc_cell_ranges.argtypes = [_U, ctypes.c_int, ctypes.c_longlong, _U, _U]
    lib.orc_cell_ranges.restype = None
    lib.orc_poly6.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.orc_poly6.restype = ctypes.c_double
    lib.orc_spiky_grad.argtypes = [ctypes.c_double, ctypes.c_double * 3, ctypes.c_double * 3]
    lib.orc_spiky_grad.restype = None
    lib.orc_boundary_density.argtypes = [ctypes.POINTER(OrcBox), ctypes.c_double,
                                         ctypes.c_double * 3]
    lib.orc_boundary_density.restype = ctypes.c_double
    lib.orc_step.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcBox), ctypes.c_int, _F,
                             _F, _F, _F, _U, ctypes.POINTER(OrcOptions), ctypes.POINTER(OrcDiag)]
    lib.orc_step.restype = ctypes.c_int
    lib.orc_density_error.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcBox),
                                      ctypes.c_int, _F, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double * 2, _D]
    lib.orc_density_error.restype = ctypes.c_int
    lib.orc_step_f32.argtypes = lib.orc_step.argtypes
    lib.orc_step_f32.restype = ctypes.c_int
    lib.orc_density_error_f32.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcBox),
                                          ctypes.c_int, _F, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double * 2, _D]
    lib.orc_density_error_f32.restype = ctypes.c_int
    lib.orc_rollout.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcBox),
                                ctypes.POINTER(OrcBox), ctypes.c_int, _D, _D, _U, ctypes.c_int,
                                ctypes.c_double, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _D,
                                _D]
    lib.orc_rollout.restype = ctypes.c_int
    lib.orc_max_threads.argtypes = []
    lib.orc_max_threads.restype = ctypes.c_int
    _lib = lib
    return lib


def params_c(p) -> OrcParams:
    """From any object with the GUIParams fluid fields."""
    return OrcParams(int(p.niter), float(p.pho0), float(p.g), float(p.h), float(p.dt),
                     float(p.lambda_eps), float(p.delta_q), float(p.k_corr), float(p.n_corr),
                     float(p.k_boundaryDensity), float(p.c_XSPH))


def box_c(ulim: Sequence[float], llim: Sequence[float]) -> OrcBox:
    b = OrcBox()
    for a in range(3):
        b.ulim[a] = float(ulim[a])
        b.llim[a] = float(llim[a])
    return b


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_F)


def _up(a: np.ndarray):
    return a.ctypes.data_as(_U)


def _dp(a: Optional[np.ndarray]):
    return a.ctypes.data_as(_D) if a is not None else None


def grid_dims(ulim, llim, h: float):
    d = (ctypes.c_int * 3)()
    if load().orc_grid_dims(ctypes.byref(box_c(ulim, llim)), float(h), d) != 0:
        raise ValueError("invalid box")
    return tuple(d)


def cell_keys(x: np.ndarray, ulim, llim, h: float) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    n, stride = x.shape
    dims = (ctypes.c_int * 3)(*grid_dims(ulim, llim, h))
    keys = np.empty(n, np.uint32)
    load().orc_cell_keys(_fp(x), stride, n, ctypes.byref(box_c(ulim, llim)), float(h), dims,
                         _up(keys))
    return keys


def stable_order(keys: np.ndarray) -> np.ndarray:
    keys = np.ascontiguousarray(keys, np.uint32)
    perm = np.empty(keys.shape[0], np.uint32)
    load().orc_stable_order(_up(keys), keys.shape[0], _up(perm))
    return perm


def cell_ranges(sorted_keys: np.ndarray, cells: int):
    sk = np.ascontiguousarray(sorted_keys, np.uint32)
    start = np.empty(cells, np.uint32)
    end = np.empty(cells, np.uint32)
    load().orc_cell_ranges(_up(sk), sk.shape[0], cells, _up(start), _up(end))
    return start, end


def poly6(h: float, r2: float) -> float:
    return load().orc_poly6(float(h), float(r2))


def spiky_grad(h: float, r: Sequence[float]) -> np.ndarray:
    g = (ctypes.c_double * 3)()
    load().orc_spiky_grad(float(h), (ctypes.c_double * 3)(*map(float, r)), g)
    return np.array(list(g))


def boundary_density(ulim, llim, h: float, p: Sequence[float]) -> float:
    return load().orc_boundary_density(ctypes.byref(box_c(ulim, llim)), float(h),
                                       (ctypes.c_double * 3)(*map(float, p)))


def step(params, ulim, llim, pos: np.ndarray, vel: np.ndarray, iid: np.ndarray, *,
         brute_force: bool = False, threads: int = 0, query_cell_f64: bool = False,
         fp32: bool = False) -> dict:
    """One oracle frame.  pos/vel: (n, 3|4) float32 in input order.  Returns
    the caller arrays after the step (all in sorted order) plus diagnostics.
    fp32=False: FP64 oracle; fp32=True: the FP32 mirror of the GPU's
    operation sequence (bit-exact reference for the GPU)."""
    pos = np.array(pos, np.float32, copy=True, order="C")
    vel = np.array(vel, np.float32, copy=True, order="C")
    iid = np.array(iid, np.uint32, copy=True)
    n, stride = pos.shape
    npos = np.zeros_like(pos)
    nvel = np.zeros_like(vel)
    out = {k: np.zeros((n, 3)) for k in ("x_final", "v_final", "v_pre", "dp_raw_first")}
    out["lambda"] = np.zeros(n)
    out["rho"] = np.zeros(n)
    out["perm"] = np.zeros(n, np.uint32)
    out["keys"] = np.zeros(n, np.uint32)
    diag = OrcDiag(_dp(out["x_final"]), _dp(out["v_final"]), _dp(out["v_pre"]),
                   _dp(out["lambda"]), _dp(out["rho"]), _dp(out["dp_raw_first"]),
                   _up(out["perm"]), _up(out["keys"]))
    opt = OrcOptions(stride, 1 if brute_force else 0, int(threads), 1 if query_cell_f64 else 0)
    fn = load().orc_step_f32 if fp32 else load().orc_step
    rc = fn(ctypes.byref(params_c(params)), ctypes.byref(box_c(ulim, llim)), n,
                         _fp(pos), _fp(npos), _fp(vel), _fp(nvel), _up(iid), ctypes.byref(opt),
                         ctypes.byref(diag))
    if rc != 0:
        raise ValueError("orc_step rejected its input")
    out.update(pos=pos, vel=vel, npos=npos, nvel=nvel, iid=iid)
    return out


def density_error(params, ulim, llim, x: np.ndarray, threads: int = 0,
                  per_particle: bool = False):
    """(mean, max) of |rho_i/pho0 - 1| with FP64 rho on a fresh grid; with
    per_particle, also the per-particle errors in input order."""
    x = np.ascontiguousarray(x, np.float32)
    out = (ctypes.c_double * 2)()
    err = np.zeros(x.shape[0]) if per_particle else None
    rc = load().orc_density_error(ctypes.byref(params_c(params)), ctypes.byref(box_c(ulim, llim)),
                                  x.shape[0], _fp(x), x.shape[1], int(threads), out, _dp(err))
    if rc != 0:
        raise ValueError("orc_density_error rejected its input")
    if per_particle:
        return float(out[0]), float(out[1]), err
    return float(out[0]), float(out[1])


def density_error_f32(params, ulim, llim, x: np.ndarray, threads: int = 0):
    """(mean, max, per-particle errors in input order) with FP32 rho."""
    x = np.ascontiguousarray(x, np.float32)
    out = (ctypes.c_double * 2)()
    err = np.zeros(x.shape[0])
    rc = load().orc_density_error_f32(ctypes.byref(params_c(params)),
                                      ctypes.byref(box_c(ulim, llim)), x.shape[0], _fp(x),
                                      x.shape[1], int(threads), out, _dp(err))
    if rc != 0:
        raise ValueError("orc_density_error_f32 rejected its input")
    return float(out[0]), float(out[1]), err


def max_threads() -> int:
    return load().orc_max_threads()


def rollout(params, ulim, llim, pos0: np.ndarray, vel0: np.ndarray, iid0: np.ndarray,
            nframes: int, target: Optional[np.ndarray] = None, *, pho0: float = 0.0,
            fp32: bool = False, threads: int = 0, grad: bool = True, boxes=None):
    """nframes frames from (pos0, vel0, iid0) (N x 3, input order); the state
    stays in FP64 (fp32=False) or follows the FP32 mirror (fp32=True).
    boxes: optional per-frame [(ulim, llim), ...] (a sweeping wall).
    Returns dict(x_final (N x 3 by identity), loss, g_pos0, g_vel0 (N x 3,
    input order), g_rho0): the loss (1/N) sum |x_final - target|^2 and its
    adjoint gradients (FP64), when target is given and grad is True."""
    lib = load()
    n = int(pos0.shape[0])
    P = np.ascontiguousarray(np.asarray(pos0, np.float64)[:, :3])
    V = np.ascontiguousarray(np.asarray(vel0, np.float64)[:, :3])
    I = np.ascontiguousarray(iid0, np.uint32)
    xf = np.zeros((n, 3))
    out = {"x_final": xf}
    tgt = None if target is None else np.ascontiguousarray(target, np.float64)
    loss = ctypes.c_double(0.0)
    gp = gv = None
    gr = ctypes.c_double(0.0)
    if tgt is not None and grad:
        gp, gv = np.zeros((n, 3)), np.zeros((n, 3))
    bx = None
    if boxes is not None:
        if len(boxes) != int(nframes):
            raise ValueError("boxes needs one (ulim, llim) per frame")
        bx = (OrcBox * len(boxes))(*[box_c(u, l) for u, l in boxes])
    rc = lib.orc_rollout(ctypes.byref(params_c(params)), ctypes.byref(box_c(ulim, llim)), bx, n,
                         _dp(P), _dp(V), _up(I), int(nframes), float(pho0), 1 if fp32 else 0,
                         int(threads), _dp(tgt), ctypes.byref(loss) if tgt is not None else None,
                         _dp(xf), _dp(gp), _dp(gv), ctypes.byref(gr) if gp is not None else None)
    if rc != 0:
        raise ValueError("orc_rollout rejected its input")
    if tgt is not None:
        out["loss"] = loss.value
    if gp is not None:
        out.update(g_pos0=gp, g_vel0=gv, g_rho0=gr.value)
    return out

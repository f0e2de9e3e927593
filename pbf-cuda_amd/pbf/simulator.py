"""Host-side mirror of the reference's Simulator (fluids/Simulator.h:7-61).

Same method names and argument meaning as the reference class, over the C ABI
of include/pbf.h.  Differences: buffers are device pointers (torch tensors or
integer addresses) instead of GL buffer names, parameters are passed
explicitly instead of read from the GUIParams singleton, and failures raise
PBFError instead of exiting the process.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

from . import _lib
from ._lib import PBFError, PbfParams, PbfShapes, PbfSweep, PbfVec3
from .params import FluidParams


def _vec3(v: Sequence[float]) -> PbfVec3:
    return PbfVec3(float(v[0]), float(v[1]), float(v[2]))


def _ptr(buf, layout: int, what: str, n: int, is_vec: bool) -> int:
    """Device address of a torch tensor (checked) or a raw integer address."""
    if isinstance(buf, int):
        return buf
    import torch

    if not isinstance(buf, torch.Tensor):
        raise TypeError(f"{what}: expected a torch tensor or a device address, got {type(buf)}")
    if not buf.is_cuda:
        raise ValueError(f"{what}: tensor must be on a HIP device")
    if not buf.is_contiguous():
        raise ValueError(f"{what}: tensor must be contiguous")
    if is_vec:
        if buf.dtype != torch.float32:
            raise ValueError(f"{what}: dtype must be float32")
        if buf.numel() < n * layout:
            raise ValueError(f"{what}: holds {buf.numel()} floats, need {n * layout}")
    else:
        if buf.dtype not in (torch.int32, torch.uint32):
            raise ValueError(f"{what}: dtype must be int32/uint32")
        if buf.numel() < n:
            raise ValueError(f"{what}: holds {buf.numel()} ids, need {n}")
    return buf.data_ptr()


class Sweep:
    """Sweeping wall (fluids/FluidSystem.cpp:36-38, :104-110): from frame
    start_frame on, ulim = ulim0 + a_ulim * sin(w * (frame - start_frame)),
    llim likewise; w in radians per frame."""

    def __init__(self, ulim0, llim0, a_ulim=(0.0, 0.0, 0.0), a_llim=(0.0, 0.0, 0.0),
                 w: float = 0.0, start_frame: int = 0):
        self.c = PbfSweep(_vec3(ulim0), _vec3(llim0), _vec3(a_ulim), _vec3(a_llim), float(w),
                          int(start_frame))

    def box(self, frame: int):
        """(ulim, llim) of `frame`, computed by the library (host only)."""
        u, l = PbfVec3(), PbfVec3()
        _lib.check(_lib.load().pbf_sweep_box(ctypes.byref(self.c), int(frame), ctypes.byref(u),
                                             ctypes.byref(l)), None)
        return (u.x, u.y, u.z), (l.x, l.y, l.z)


def _stream_handle(stream) -> Optional[int]:
    if stream is None:
        import torch

        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Simulator:
    """Simulator(params, ulim, llim) (fluids/Simulator.h:10).

    capacity: maximum particle count (the reference's MAX_PARTICLE_NUM,
    fluids/helper.h:10, lifted to a per-instance value).  max_cells: maximum
    grid cells over the solver's life (<= 0: four times the initial box's,
    as fluids/Simulator.h:13-14).  layout: 4 for float4 vector arrays, 3 for
    packed float3 (the reference's GL buffers).
    """

    def __init__(self, params: FluidParams, ulim: Sequence[float], llim: Sequence[float],
                 capacity: int, max_cells: int = 0, layout: int = _lib.PBF_LAYOUT_FLOAT4):
        self._lib = _lib.load()
        self._h = ctypes.c_void_p()
        self.layout = int(layout)
        self.capacity = int(capacity)
        p = params.to_c()
        rc = self._lib.pbf_create(ctypes.byref(p), _vec3(ulim), _vec3(llim), self.capacity,
                                  int(max_cells), self.layout, ctypes.byref(self._h))
        _lib.check(rc, None)

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        """~Simulator (fluids/Simulator.h:27-34)."""
        if self._h:
            self._lib.pbf_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _handle(self):
        if not self._h:
            raise PBFError(_lib.PBF_ERR_STATE, "simulator is closed")
        return self._h

    # -- reference API ----------------------------------------------------
    def step(self, pos, npos, vel, nvel, iid, nparticle: int, stream=None) -> None:
        """Simulator::step (fluids/Simulator.h:37, fluids/Simulator.cpp:10-99).

        Reads pos/vel/iid, writes npos/nvel; leaves every array in the new
        cell-sorted order (pos = sorted start positions, vel = velocities
        before XSPH).  Stream-ordered on `stream` (default: torch's current
        stream); does not synchronise."""
        n = int(nparticle)
        L = self.layout
        args = (_ptr(pos, L, "pos", n, True), _ptr(npos, L, "npos", n, True),
                _ptr(vel, L, "vel", n, True), _ptr(nvel, L, "nvel", n, True),
                _ptr(iid, L, "iid", n, False))
        h = self._handle()
        _lib.check(self._lib.pbf_step(h, *args, n, _stream_handle(stream)), h)

    def loadParams(self, params: FluidParams) -> None:
        """Simulator::loadParams (fluids/Simulator.cpp:101-115), with the
        parameters passed in instead of read from a singleton."""
        p = params.to_c()
        h = self._handle()
        _lib.check(self._lib.pbf_set_params(h, ctypes.byref(p)), h)

    def saveParams(self) -> FluidParams:
        """Simulator::saveParams (fluids/Simulator.cpp:117-130): returns them."""
        p = PbfParams()
        h = self._handle()
        _lib.check(self._lib.pbf_get_params(h, ctypes.byref(p)), h)
        return FluidParams.from_c(p)

    def setLim(self, ulim: Sequence[float], llim: Sequence[float]) -> None:
        """Simulator::setLim (fluids/Simulator.cpp:132-136)."""
        h = self._handle()
        _lib.check(self._lib.pbf_set_lim(h, _vec3(ulim), _vec3(llim)), h)

    def getLim(self):
        u, l = PbfVec3(), PbfVec3()
        h = self._handle()
        _lib.check(self._lib.pbf_get_lim(h, ctypes.byref(u), ctypes.byref(l)), h)
        return (u.x, u.y, u.z), (l.x, l.y, l.z)

    # -- additions --------------------------------------------------------
    def run_frames(self, pos_a, vel_a, pos_b, vel_b, iid, nparticle: int, first_frame: int,
                   nframes: int, sweep: Optional[Sweep] = None, stream=None) -> int:
        """FluidSystem::stepSimulate for nframes frames (tic-toc buffers, optional
        sweeping wall).  Returns 0 if the newest state is in (pos_a, vel_a),
        1 if in (pos_b, vel_b).  Stream-ordered; does not synchronise."""
        n = int(nparticle)
        L = self.layout
        args = (_ptr(pos_a, L, "pos_a", n, True), _ptr(vel_a, L, "vel_a", n, True),
                _ptr(pos_b, L, "pos_b", n, True), _ptr(vel_b, L, "vel_b", n, True),
                _ptr(iid, L, "iid", n, False))
        which = ctypes.c_int(0)
        h = self._handle()
        _lib.check(self._lib.pbf_run_frames(h, *args, n, int(first_frame), int(nframes),
                                            ctypes.byref(sweep.c) if sweep else None,
                                            _stream_handle(stream), ctypes.byref(which)), h)
        return int(which.value)

    def density_error(self, pos, nparticle: int, stream=None):
        """(mean, max) of |rho_i/pho0 - 1| on a fresh grid over pos; synchronises."""
        n = int(nparticle)
        out = (ctypes.c_double * 2)()
        h = self._handle()
        _lib.check(self._lib.pbf_density_error(h, _ptr(pos, self.layout, "pos", n, True), n,
                                               _stream_handle(stream), out), h)
        return float(out[0]), float(out[1])

    def set_timing(self, enable: bool) -> None:
        h = self._handle()
        _lib.check(self._lib.pbf_set_timing(h, 1 if enable else 0), h)

    def get_timing(self):
        """{phase: (total_ms, count)} since the last call (synchronises)."""
        ms = (ctypes.c_double * _lib.PBF_NUM_PHASES)()
        cnt = (ctypes.c_longlong * _lib.PBF_NUM_PHASES)()
        h = self._handle()
        _lib.check(self._lib.pbf_get_timing(h, ms, cnt), h)
        return {name: (float(ms[k]), int(cnt[k])) for k, name in enumerate(_lib.PHASES)}

    def set_graph(self, enable: bool) -> None:
        """Graph mode: record each (arrays, n, stream) step once as a HIP graph
        and replay it (one graph launch per step)."""
        h = self._handle()
        _lib.check(self._lib.pbf_set_graph(h, 1 if enable else 0), h)

    def graph_info(self) -> dict:
        """Graph bookkeeping since creation: recorded, instantiated, updated
        (in place after a parameter / box / shape change), replayed, live."""
        gi = _lib.PbfGraphInfo()
        h = self._handle()
        _lib.check(self._lib.pbf_get_graph_info(h, ctypes.byref(gi)), h)
        return {k: int(getattr(gi, k)) for k, _ in gi._fields_}

    def set_shapes(self, light_block: int, chunk_blocks: int) -> None:
        """Launch shapes of the neighbour passes (results never depend on them)."""
        sh = PbfShapes(int(light_block), int(chunk_blocks))
        h = self._handle()
        _lib.check(self._lib.pbf_set_shapes(h, ctypes.byref(sh)), h)

    def get_shapes(self):
        sh = PbfShapes()
        h = self._handle()
        _lib.check(self._lib.pbf_get_shapes(h, ctypes.byref(sh)), h)
        return int(sh.light_block), int(sh.chunk_blocks)

    def set_heavy_threshold(self, threshold: int) -> None:
        """Candidate count above which a target's chunks go to the chunk
        kernel; 0 = adaptive (the default).  Results never depend on it."""
        h = self._handle()
        _lib.check(self._lib.pbf_set_heavy_threshold(h, int(threshold)), h)

    def get_heavy_threshold(self) -> int:
        v = ctypes.c_int(0)
        h = self._handle()
        _lib.check(self._lib.pbf_get_heavy_threshold(h, ctypes.byref(v)), h)
        return int(v.value)

    def autotune(self, pos, vel, iid, nparticle: int, reps: int = 3, stream=None):
        """Times one step from (pos, vel, iid) per candidate launch shape on
        private copies, checks every shape gives the same bits, installs the
        fastest.  Returns ((light_block, chunk_blocks), median ms).
        Synchronises; the arrays are not modified."""
        n = int(nparticle)
        L = self.layout
        best = PbfShapes()
        ms = ctypes.c_double(0.0)
        h = self._handle()
        _lib.check(self._lib.pbf_autotune(h, _ptr(pos, L, "pos", n, True),
                                          _ptr(vel, L, "vel", n, True),
                                          _ptr(iid, L, "iid", n, False), n, int(reps),
                                          _stream_handle(stream), ctypes.byref(best),
                                          ctypes.byref(ms)), h)
        return (int(best.light_block), int(best.chunk_blocks)), float(ms.value)


TAPE_STORE = 1
TAPE_GATHER = 2
TAPE_SCATTER = 4
TAPE_RESCALE = 8


class Tape:
    """Checkpointed run of `nframes` frames of a single-scene Simulator and
    its backward pass (pbf_tape_*).  store=True keeps every frame's solver
    iterations, sort and grid on the device (pbf_tape_create_ex with
    PBF_TAPE_STORE), so the backward pass recomputes nothing; the gradients
    are the same bits either way.  adjoint: None (automatic), "gather" or
    "scatter" (PBF_TAPE_GATHER / PBF_TAPE_SCATTER: how the backward pass
    collects the pair terms the stale grid makes one-sided).  rescale=True
    (PBF_TAPE_RESCALE) renormalises the adjoint by a power of two before each
    frame; backward() then fills the true gradients x 2^-gradient_log2_scale()."""

    def __init__(self, sim: "Simulator", nparticle: int, nframes: int, store: bool = False,
                 adjoint: Optional[str] = None, rescale: bool = False):
        self._lib = _lib.load()
        self.sim = sim
        self.n = int(nparticle)
        self.nframes = int(nframes)
        self.store = bool(store)
        modes = {None: 0, "gather": TAPE_GATHER, "scatter": TAPE_SCATTER}
        if adjoint not in modes:
            raise ValueError(f"adjoint must be None, 'gather' or 'scatter', not {adjoint!r}")
        self.adjoint = adjoint
        self._t = ctypes.c_void_p()
        h = sim._handle()
        self.rescale = bool(rescale)
        flags = ((TAPE_STORE if self.store else 0) | modes[adjoint]
                 | (TAPE_RESCALE if self.rescale else 0))
        _lib.check(self._lib.pbf_tape_create_ex(h, self.n, self.nframes, flags,
                                                ctypes.byref(self._t)), h)

    def close(self) -> None:
        if self._t:
            self._lib.pbf_tape_destroy(self._t)
            self._t = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, pos, vel, iid, stream=None, first_frame: int = 0,
                sweep: Optional[Sweep] = None) -> None:
        """Runs the frames from (pos, vel, iid) (not modified); with a sweep,
        frame k runs in sweep.box(first_frame + k)."""
        L, n = self.sim.layout, self.n
        h = self.sim._handle()
        _lib.check(self._lib.pbf_tape_forward(h, self._t, _ptr(pos, L, "pos", n, True),
                                              _ptr(vel, L, "vel", n, True),
                                              _ptr(iid, L, "iid", n, False), int(first_frame),
                                              ctypes.byref(sweep.c) if sweep else None,
                                              _stream_handle(stream)), h)

    def read(self, frame: int, pos=None, vel=None, iid=None, stream=None) -> None:
        """Copies checkpoint `frame` (0 = inputs, nframes = final state)."""
        L, n = self.sim.layout, self.n
        _lib.check(self._lib.pbf_tape_read(
            self._t, int(frame), _ptr(pos, L, "pos", n, True) if pos is not None else None,
            _ptr(vel, L, "vel", n, True) if vel is not None else None,
            _ptr(iid, L, "iid", n, False) if iid is not None else None, _stream_handle(stream)),
            None)

    def backward(self, gpos, gvel, gpos0, gvel0, stream=None) -> float:
        """gpos/gvel: dL/d(final state) in the final order (None = 0); fills
        gpos0/gvel0 = dL/d(initial state); returns dL/dpho0.  Synchronises."""
        L, n = self.sim.layout, self.n
        g = ctypes.c_double(0.0)
        h = self.sim._handle()
        _lib.check(self._lib.pbf_tape_backward(
            h, self._t, _ptr(gpos, L, "gpos", n, True) if gpos is not None else None,
            _ptr(gvel, L, "gvel", n, True) if gvel is not None else None,
            _ptr(gpos0, L, "gpos0", n, True), _ptr(gvel0, L, "gvel0", n, True), ctypes.byref(g),
            _stream_handle(stream)), h)
        return float(g.value)

    def gradient_log2_scale(self) -> int:
        """e such that the last backward() wrote the true gradients x 2^-e
        (0 unless rescale=True)."""
        e = ctypes.c_int(0)
        _lib.check(self._lib.pbf_tape_gradient_scale(self._t, ctypes.byref(e)))
        return int(e.value)


class BatchedSimulator(Simulator):
    """Many independent box scenes stepped together (pbf_create_batched).

    Scene k owns the index range [offsets[k], offsets[k] + counts[k]) of
    every particle array and its own box (ulims[k], llims[k]); all scenes
    share `params`.  Each scene's result equals stepping it alone in a
    Simulator, bit for bit."""

    def __init__(self, params: FluidParams, ulims, llims, counts, max_cells: int = 0,
                 layout: int = _lib.PBF_LAYOUT_FLOAT4):
        self._lib = _lib.load()
        self._h = ctypes.c_void_p()
        self.layout = int(layout)
        k = len(counts)
        if len(ulims) != k or len(llims) != k:
            raise ValueError("ulims, llims and counts must have one entry per scene")
        self.counts = [int(c) for c in counts]
        self.offsets = [0] * k
        for q in range(1, k):
            self.offsets[q] = self.offsets[q - 1] + self.counts[q - 1]
        self.capacity = sum(self.counts)
        u = (PbfVec3 * k)(*[_vec3(v) for v in ulims])
        l = (PbfVec3 * k)(*[_vec3(v) for v in llims])
        c = (ctypes.c_int * k)(*self.counts)
        p = params.to_c()
        rc = self._lib.pbf_create_batched(ctypes.byref(p), k, u, l, c, int(max_cells), self.layout,
                                          ctypes.byref(self._h))
        _lib.check(rc, None)

    @property
    def nscenes(self) -> int:
        return len(self.counts)

    def step_batched(self, pos, npos, vel, nvel, iid, stream=None) -> None:
        """One frame of every scene (pbf_step_batched)."""
        n = self.capacity
        L = self.layout
        args = (_ptr(pos, L, "pos", n, True), _ptr(npos, L, "npos", n, True),
                _ptr(vel, L, "vel", n, True), _ptr(nvel, L, "nvel", n, True),
                _ptr(iid, L, "iid", n, False))
        h = self._handle()
        _lib.check(self._lib.pbf_step_batched(h, *args, _stream_handle(stream)), h)

    def scenes(self):
        """Per-scene counts as the library reports them."""
        k = ctypes.c_int(0)
        h = self._handle()
        _lib.check(self._lib.pbf_get_scenes(h, ctypes.byref(k), None), h)
        counts = (ctypes.c_int * k.value)()
        _lib.check(self._lib.pbf_get_scenes(h, ctypes.byref(k), counts), h)
        return list(counts)

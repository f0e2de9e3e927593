"""torch.autograd.Function over the differentiable rollout of the C ABI.

rollout(sim, tape, pos0, vel0, rho0, iid0) runs tape.nframes frames of the
single-scene Simulator `sim` from (pos0, vel0) with rest density rho0 and
returns the final positions indexed by particle identity (N x 3), so a loss
can compare them with per-particle targets.  Backward calls the library's
adjoint kernels (pbf_tape_backward): gradients with respect to pos0, vel0
(in their given order) and rho0, per-frame checkpoints, inner iterations
recomputed.  The frame map is differentiated with its discrete choices
(sort order, grid, candidate lists, cutoffs, active clamps) held fixed.

The solver's own parameters are left as they were: the forward run and the
backward pass each set the rest density to rho0 for their duration and put
the caller's parameters back afterwards (the library keeps its recorded
graphs across such changes).  rho0 is read on the host (the step constants
are formed there), so a device rho0 costs one synchronisation; pass a CPU
tensor to avoid it.
"""
from __future__ import annotations

import torch


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos0, vel0, rho0, sim, tape, iid0):
        n, L = tape.n, sim.layout
        saved = sim.saveParams()
        run = saved.replace(pho0=float(rho0))
        stream = torch.cuda.current_stream()
        sim.loadParams(run)
        try:
            tape.forward(pos0.contiguous(), vel0.contiguous(), iid0, stream)
        finally:
            sim.loadParams(saved)
        xf = torch.empty(n * L, device=pos0.device)
        iidf = torch.empty(n, dtype=torch.int32, device=pos0.device)
        tape.read(tape.nframes, pos=xf, iid=iidf, stream=stream)
        x_by_iid = torch.empty(n, 3, device=pos0.device)
        x_by_iid[iidf.long()] = xf.view(n, L)[:, :3]
        ctx.save_for_backward(iidf)
        ctx.sim, ctx.tape, ctx.params = sim, tape, run
        ctx.rho_dtype = rho0.dtype
        return x_by_iid

    @staticmethod
    def backward(ctx, g_by_iid):
        (iidf,) = ctx.saved_tensors
        sim, tape = ctx.sim, ctx.tape
        n, L = tape.n, sim.layout
        dev = g_by_iid.device
        gpos = torch.zeros(n, L, device=dev)
        gpos[:, :3] = g_by_iid.float()[iidf.long()]
        gpos0 = torch.empty(n * L, device=dev)
        gvel0 = torch.empty(n * L, device=dev)
        saved = sim.saveParams()
        sim.loadParams(ctx.params)  # the parameters of the forward run
        try:
            grho = tape.backward(gpos.view(-1), None, gpos0, gvel0, torch.cuda.current_stream())
        finally:
            sim.loadParams(saved)
        return (gpos0.view(n, L), gvel0.view(n, L), torch.tensor(grho, dtype=ctx.rho_dtype,
                                                                  device=dev), None, None, None)


def rollout(sim, tape, pos0: torch.Tensor, vel0: torch.Tensor, rho0: torch.Tensor,
            iid0: torch.Tensor) -> torch.Tensor:
    """Final positions (N x 3, by identity) after tape.nframes frames; pos0,
    vel0 are (N, layout) float32 device tensors, rho0 a 0-d tensor."""
    return _Rollout.apply(pos0, vel0, rho0, sim, tape, iid0)

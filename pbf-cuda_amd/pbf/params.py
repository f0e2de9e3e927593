"""Fluid parameter sets (the solver fields of GUIParams, fluids/GUIParams.h:6-17)."""
from __future__ import annotations

import dataclasses

import numpy as np

from ._lib import PbfParams


@dataclasses.dataclass
class FluidParams:
    """Same fields, names and meaning as GUIParams' fluid block."""

    niter: int = 4
    pho0: float = 8000.0
    g: float = 9.8
    h: float = 0.1
    dt: float = 0.0083
    lambda_eps: float = 1000.0
    delta_q: float = 0.03
    k_corr: float = 0.001
    n_corr: float = 4.0
    k_boundaryDensity: float = 0.0
    c_XSPH: float = 0.5

    def to_c(self) -> PbfParams:
        return PbfParams(int(self.niter), *(float(getattr(self, f)) for f in _FLOAT_FIELDS))

    @classmethod
    def from_c(cls, p: PbfParams) -> "FluidParams":
        return cls(niter=int(p.niter), **{f: float(getattr(p, f)) for f in _FLOAT_FIELDS})

    def replace(self, **kw) -> "FluidParams":
        return dataclasses.replace(self, **kw)


_FLOAT_FIELDS = ("pho0", "g", "h", "dt", "lambda_eps", "delta_q", "k_corr", "n_corr",
                 "k_boundaryDensity", "c_XSPH")


def reference_defaults() -> FluidParams:
    """Defaults the reference's FluidSystem writes (fluids/FluidSystem.cpp:15-25).

    delta_q = 0.3 * h is formed in double there and stored as float."""
    h = float(np.float32(0.1))
    return FluidParams(niter=4, pho0=8000.0, g=9.8, h=h, dt=0.0083, lambda_eps=1000.0,
                       delta_q=float(np.float32(0.3 * h)), k_corr=0.001, n_corr=4.0,
                       k_boundaryDensity=0.0, c_XSPH=0.5)


def dam_params(niter: int = 4) -> FluidParams:
    """The constants of the dam-break workloads: h=0.1, rest density 6378,
    epsilon 600, s_corr k=0.1 n=4 dq=0.2h, g=9.8 along -z, dt=1/60; XSPH and
    wall density keep the reference defaults (0.5 and 0)."""
    h = 0.1
    return FluidParams(niter=niter, pho0=6378.0, g=9.8, h=h, dt=1.0 / 60.0, lambda_eps=600.0,
                       delta_q=0.2 * h, k_corr=0.1, n_corr=4.0, k_boundaryDensity=0.0,
                       c_XSPH=0.5)

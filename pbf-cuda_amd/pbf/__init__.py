"""MI355X (gfx950) Position Based Fluids step: Python side of include/pbf.h.

`Simulator` mirrors the reference's Simulator class (fluids/Simulator.h:7-61)
over the C ABI in libpbf.so; `scenes` builds seeded particle scenes;
`params` holds the fluid parameter sets; `viewer` exports frames as point
clouds for an external viewer.
"""
from ._lib import PBFError, LIB_PATH, load  # noqa: F401
from .params import FluidParams, dam_params, reference_defaults  # noqa: F401
from .simulator import BatchedSimulator, Simulator, Sweep, Tape  # noqa: F401
from . import scenes, viewer  # noqa: F401

"""sparc_flows_amd: MI355X (gfx950) implementation of SPARC's 2D time-domain
solver for the linearized Euler equations with background flows, its adjoint
and the forward x adjoint sensitivity kernels.

Hot path in HIP (libsparc_hip.so, C ABI in include/sparc_hip.h); torch is used
for device memory and streams only.
"""
from .params import SparcParams, read_params_i, f32  # noqa: F401

__all__ = ["SparcParams", "read_params_i", "f32", "Sparc"]


def __getattr__(name):
    if name == "Sparc":
        from .solver import Sparc
        return Sparc
    raise AttributeError(name)

"""openqarp_amd — MI355X statevector backend for OpenQARP.

Hand-written HIP kernels for gfx950 behind the C ABI of include/qarp_hip.h
(libqarp_hip.so), with a Python host that mirrors the reference's
simulator / engine / primitive interface.
"""
from .errors import BackendUnavailable, CapabilityError
from .ir import (Command, GateType, Param, SamplingResult, affine_coefficients, concrete,
                 substitute_all, symbol)

__all__ = [
    "BackendUnavailable", "CapabilityError", "Command", "GateType", "Param", "SamplingResult",
    "affine_coefficients", "concrete", "substitute_all", "symbol",
]

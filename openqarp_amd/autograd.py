"""torch.autograd boundary of the adjoint gradient.

``expectation(theta, sim, commands, n_qubits, observable, symbols)`` returns
⟨ψ(θ)|H|ψ(θ)⟩ as a 0-d tensor whose backward is the device adjoint sweep.
The forward only applies the circuit and evaluates ⟨H⟩; the backward makes
one qhip_run_gradient call, which re-applies the circuit from the root state
and un-applies it gate by gate on ψ and λ = Hψ while every generator inner
product accumulates in FP64 — states are recomputed, no per-layer
checkpoints are kept, and a forward that is never differentiated costs no
gradient work.

Tensors cross the boundary on the CPU: the device work happens in
libqarp_hip.so on its own HIP stream, and only θ (the parameters) and the
scalar / gradient results move, as float64.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from .simulator import QarpSimulator


class AdjointExpectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, sim: QarpSimulator, commands, n_qubits: int, observable,
                symbols: Sequence[str]) -> torch.Tensor:
        if theta.dim() != 1 or theta.numel() != len(symbols):
            raise ValueError(f"theta must be a vector of {len(symbols)} parameters, got shape {tuple(theta.shape)}")
        values = theta.detach().to("cpu", torch.float64).numpy()
        params = {s: float(v) for s, v in zip(symbols, values)}
        concrete = [c.substitute(params) if c.is_parametric() else c for c in commands]
        energy = sim.expectation(concrete, n_qubits, observable)
        ctx.adjoint = (sim, commands, n_qubits, observable, list(symbols), params)
        return torch.tensor(energy, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        sim, commands, n_qubits, observable, symbols, params = ctx.adjoint
        grad, _ = sim.run_gradient_energy(commands, n_qubits, observable, params, symbols)
        grad = torch.from_numpy(np.ascontiguousarray(grad)).to(g.device, g.dtype)
        return g * grad, None, None, None, None, None


def expectation(theta: torch.Tensor, sim: QarpSimulator, commands, n_qubits: int, observable,
                symbols: Sequence[str]) -> torch.Tensor:
    return AdjointExpectation.apply(theta, sim, commands, n_qubits, observable, list(symbols))

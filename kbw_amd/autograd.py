"""torch.autograd wrapper of the adjoint-differentiation entry point.

`expval(params, sim, ops, hamiltonian)` returns <H> as a 0-d float64 tensor
whose backward is the gradient computed by kbwh_adjoint_gradient (two extra
state vectors, gates un-applied in reverse).  `parameter_shift` is the plain
per-gate path's gradient by the two-term shift rule, which libket uses for
KBW (quantum_execution.rs:237-241); it is exact for uncontrolled RX/RY/RZ/P.
"""
import math

import numpy as np
import torch

from ._lib import RUN_PER_GATE


class AdjointExpval(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, sim, ops, hamiltonian):
        p = params.detach().to("cpu", torch.float64).numpy()
        e, g = sim.adjoint_gradient(ops, p, hamiltonian)
        ctx.save_for_backward(torch.from_numpy(g).to(device=params.device, dtype=params.dtype))
        return params.new_tensor(e)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None, None


def expval(params, sim, ops, hamiltonian):
    return AdjointExpval.apply(params, sim, ops, hamiltonian)


def energy(sim, ops, params, hamiltonian, mode=RUN_PER_GATE):
    sim.clear()
    sim.run(ops, params=params, mode=mode)
    return sim.exp_value(hamiltonian)


def parameter_shift(sim, ops, params, hamiltonian, indices=None, mode=RUN_PER_GATE):
    """dE/dparams[j] = (E(p + pi/2 e_j) - E(p - pi/2 e_j)) / 2 for j in indices."""
    params = np.asarray(params, dtype=np.float64)
    idx = range(len(params)) if indices is None else indices
    out = np.zeros(len(list(idx)) if indices is not None else len(params))
    for k, j in enumerate(range(len(params)) if indices is None else indices):
        p = params.copy()
        p[j] += math.pi / 2
        ep = energy(sim, ops, p, hamiltonian, mode)
        p[j] -= math.pi
        em = energy(sim, ops, p, hamiltonian, mode)
        out[k] = 0.5 * (ep - em)
    return out

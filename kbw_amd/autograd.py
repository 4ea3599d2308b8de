"""torch.autograd wrapper of the adjoint-differentiation entry point.

`expval(params, sim, ops, hamiltonian)` returns <H> as a 0-d float64 tensor
whose backward is the gradient computed by kbwh_adjoint_gradient (two extra
state vectors, gates un-applied in reverse).  `parameter_shift` is the plain
per-gate path's gradient by the two-term shift rule, which libket uses for
KBW (quantum_execution.rs:237-241); it is exact for a parameter that drives
one uncontrolled RX/RY/RZ/P gate, and rejects any other.
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
    """dE/dparams[j] = (E(p + pi/2 e_j) - E(p - pi/2 e_j)) / 2 for j in indices.

    The two-term rule is exact when parameter j drives exactly one gate and
    that gate is an uncontrolled RX, RY, RZ or P (generator eigenvalues
    +-1/2, or 0 and 1 for P): a parameter shared by several gates, or a
    controlled rotation, needs more shift terms, so such parameters are
    rejected rather than given a wrong value.  `indices` may be any iterable
    (it is read once)."""
    params = np.asarray(params, dtype=np.float64)
    idx = list(range(len(params))) if indices is None else [int(j) for j in indices]
    a = np.asarray(ops)
    users = {}
    for o in a:
        if int(o["param"]) >= 0:
            users.setdefault(int(o["param"]), []).append(o)
    for j in idx:
        us = users.get(j, [])
        if len(us) > 1 or any(int(o["ctrl_mask"]) & ~(1 << int(o["target"])) for o in us):
            raise ValueError(f"parameter {j} drives {len(us)} gates or a controlled gate: the two-term shift rule is not exact")
    out = np.zeros(len(idx))
    for k, j in enumerate(idx):
        p = params.copy()
        p[j] += math.pi / 2
        ep = energy(sim, ops, p, hamiltonian, mode)
        p[j] -= math.pi
        em = energy(sim, ops, p, hamiltonian, mode)
        out[k] = 0.5 * (ep - em)
    return out

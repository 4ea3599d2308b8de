"""Host logic of kbw_amd.autograd.parameter_shift, without a device.

A stand-in simulator returns E(params) = sum_j cos(params[j]) w_j, whose
exact derivative -sin(params[j]) w_j the two-term rule reproduces (the rule
is exact for sinusoids of period 2 pi).  Checks that indices may be any
iterable, read once, and that parameters the rule does not cover exactly are
rejected.
"""
import math

import numpy as np
import pytest

from kbw_amd import RY, RZ, autograd as kag, circuits


class FakeSim:
    def __init__(self, weights):
        self.w = np.asarray(weights, dtype=np.float64)
        self.params = None

    def clear(self):
        self.params = None

    def run(self, ops, params=None, mode=None):
        self.params = np.asarray(params, dtype=np.float64)

    def exp_value(self, hamiltonian):
        return float(np.sum(np.cos(self.params) * self.w))


def _ops(n_params, shared=False, controlled=False):
    ol = circuits.OpList()
    for j in range(n_params):
        ol.add(RY if j % 2 else RZ, j % 3, control=[(j + 1) % 3] if (controlled and j == 1) else (), param=j)
    if shared:
        ol.add(RY, 2, param=0)
    return ol.array()


def test_parameter_shift_generator_indices_equal_list():
    rng = np.random.default_rng(3)
    params = rng.uniform(-math.pi, math.pi, 6)
    sim = FakeSim(rng.normal(size=6))
    ops = _ops(6)
    want = -np.sin(params) * sim.w
    a = kag.parameter_shift(sim, ops, params, None, indices=[1, 4, 5])
    b = kag.parameter_shift(sim, ops, params, None, indices=(j for j in (1, 4, 5)))
    c = kag.parameter_shift(sim, ops, params, None)
    assert np.array_equal(a, b)
    assert np.allclose(a, want[[1, 4, 5]], atol=1e-14)
    assert np.allclose(c, want, atol=1e-14)


def test_parameter_shift_rejects_inexact_parameters():
    params = np.zeros(4)
    sim = FakeSim(np.ones(4))
    with pytest.raises(ValueError):
        kag.parameter_shift(sim, _ops(4, shared=True), params, None, indices=[0])
    with pytest.raises(ValueError):
        kag.parameter_shift(sim, _ops(4, controlled=True), params, None, indices=iter([1]))
    # the other parameters of the same circuits are fine
    kag.parameter_shift(sim, _ops(4, shared=True), params, None, indices=[1, 2])

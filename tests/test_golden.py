"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
and cross-checked there against the Kronecker model and closed forms): the
oracle must reproduce them on the CPU, the engine on the GPU."""
import os

import numpy as np
import pytest

import oracle
from kbw_amd import OP_DTYPE

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(HERE, name), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["ops"] = np.frombuffer(d["ops"].tobytes(), dtype=OP_DTYPE).copy()
    return d


def ctrl(m, n):
    return [q for q in range(n) if int(m) >> q & 1]


def test_oracle_reproduces_grover12():
    d = load("grover12.npz")
    s = oracle.DenseV2(12)
    s.run(d["ops"])
    assert np.array_equal(s.state(), d["amps"])
    assert s.measure(list(range(12)), d["uniforms"]) == int(d["measured"])
    assert np.array_equal(s.state(), d["post"])


def test_oracle_reproduces_brickwork10():
    d = load("brickwork10.npz")
    s = oracle.DenseV2(10)
    s.run(d["ops"])
    assert np.array_equal(s.state(), d["amps"])
    l = oracle.load()
    e = l.orc_exp_value(s._s, d["paulis"].ctypes.data, d["qubits"].ctypes.data, d["offsets"].ctypes.data,
                        d["coeffs"].ctypes.data, len(d["coeffs"]))
    assert e == d["exp_value"]


def test_oracle_reproduces_sample9():
    d = load("sample9.npz")
    s = oracle.DenseV2(9)
    s.run(d["ops"])
    out = s.sample_hier(d["qubits"].tolist(), len(d["uniforms"]), d["uniforms"])
    assert np.array_equal(out, d["outcomes"])


def test_oracle_reproduces_block8():
    d = load("block8.npz")
    b = oracle.Block(8, int(d["block_size"]))
    for o in d["ops"]:
        b.gate(int(o["kind"]), float(o["angle"]), int(o["target"]), ctrl(o["ctrl_mask"], 8))
    assert np.array_equal(b.physical_state(), d["physical"])
    assert np.array_equal(b.qubit_map(), d["l2p"]) and b.swap_count() == int(d["swaps"])


@pytest.mark.gpu
def test_engine_reproduces_fixtures():
    from kbw_amd import DenseHip, Hamiltonian, RUN_TILED

    d = load("grover12.npz")
    g = DenseHip(12)
    g.run(d["ops"])
    assert np.array_equal(g.state(), d["amps"])
    assert g.measure(list(range(12)), d["uniforms"]) == int(d["measured"])
    assert np.array_equal(g.state(), d["post"])

    d = load("brickwork10.npz")
    for mode in (0, RUN_TILED):
        g = DenseHip(10)
        g.run(d["ops"], mode=mode)
        assert np.array_equal(g.state(), d["amps"])
    strings, k = [], 0
    letters = {1: "X", 2: "Y", 3: "Z"}
    for i in range(len(d["coeffs"])):
        lo, hi = int(d["offsets"][i]), int(d["offsets"][i + 1])
        strings.append([(letters[int(d["paulis"][j])], int(d["qubits"][j])) for j in range(lo, hi)])
    e = g.exp_value(Hamiltonian(strings, d["coeffs"].tolist()))
    assert abs(e - float(d["exp_value"])) <= 1e-10 * abs(float(d["exp_value"])) + 1e-14

    d = load("sample9.npz")
    g = DenseHip(9)
    g.run(d["ops"])
    assert np.array_equal(g.sample(d["qubits"].tolist(), len(d["uniforms"]), d["uniforms"]), d["outcomes"])

    d = load("block8.npz")
    g = DenseHip(8, block=True, block_size=int(d["block_size"]))
    g.run(d["ops"])
    assert np.array_equal(g.state(), d["physical"])
    assert np.array_equal(g.qubit_map(), d["l2p"]) and g.swap_count() == int(d["swaps"])

#!/usr/bin/env python3
"""Throughput of the dense c128 gate path on one MI355X.

Workload: the 30-qubit random brickwork circuit (16 GiB state, 20 layers,
922 gates, seed 1) followed by the expectation value of 64 random Pauli
strings.  A step is one pass of the circuit over the resident state; the
state stays in HBM across steps (gates are unitary, so the circuit can be
re-applied).  value = gates/s over the timed steps.

roofline: the dominant kernel is the non-diagonal pair pass (X/Y/H/RX/RY with
no control): per launch it reads and writes every amplitude once,
2^29 pairs x 64 B = 34.36 GB; `achieved` divides that by the average
duration of those launches, timed with HIP events on the engine's own
stream inside the timed region.

cpu_baseline: the oracle (the C restatement of the reference's DenseV2 with
OpenMP in place of rayon) on the same 30-qubit state and the same gates,
on a bounded prefix of the circuit.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "30-qubit c128 gates/s + HBM TB/s; adjoint grads/s vs plain at 1e-9"
HBM_PEAK_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=20)
    ap.add_argument("--mode", choices=["per_gate", "tiled", "best"], default="best")
    ap.add_argument("--cpu-gates", type=int, default=12, help="oracle sample size (gates); 0 = skip")
    ap.add_argument("--pmc", action="store_true",
                    help="counter run: a 1 GiB torch copy then 4 pair passes (for rocprofv3 --pmc)")
    return ap.parse_args()


def pmc_run(n):
    import torch

    from kbw_amd import DenseHip

    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
    a.fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    s = DenseHip(n)
    for t in (3, 11, 20, 29):
        s.hadamard(t)
    s.synchronize()
    print(json.dumps({"pmc_run": True, "copy_bytes": 1 << 30, "pair_pass_bytes": 32 << n}), flush=True)


def cpu_baseline(n, ops, ngates):
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sub = ops[:ngates]
    s = oracle.DenseV2(n)
    t0 = time.perf_counter()
    s.run(sub)
    dt = time.perf_counter() - t0
    del s
    return {"value": len(sub) / dt, "unit": "gates/s", "cores": threads, "kind": "port",
            "sample": f"first {len(sub)} gates of the same {n}-qubit brickwork circuit on the oracle "
                      f"(C restatement of dense_v2.rs, OpenMP {threads} threads), {dt:.2f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    n = args.qubits
    if args.pmc:
        pmc_run(n)
        return

    import numpy as np
    import torch

    from kbw_amd import DenseHip, _lib, circuits

    ops, ham = circuits.brickwork(n, args.layers, seed=1)
    nops = len(ops)
    kinds = ops["kind"].astype(int)
    tgt = ops["target"].astype(np.uint64)
    ctrl = ops["ctrl_mask"].astype(np.uint64) & ~(np.uint64(1) << tgt)
    is_pair = np.isin(kinds, [_lib.X, _lib.Y, _lib.H, _lib.RX, _lib.RY]) & (ctrl == 0)
    alg_bytes = circuits.circuit_bytes(n, ops)

    sim = DenseHip(n)
    stream = torch.cuda.ExternalStream(sim.stream())
    lib, h = sim._lib, sim.handle
    mode = args.mode
    run_mode = _lib.RUN_TILED if mode == "tiled" else _lib.RUN_PER_GATE

    def step_events():
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(nops + 1)]
        ev[0].record(stream)
        for i in range(nops):
            o = ops[i]
            rc = lib.kbwh_gate_mask(h, int(o["kind"]), float(o["angle"]), int(o["target"]), int(o["ctrl_mask"]))
            if rc:
                _lib.check(rc)
            ev[i + 1].record(stream)
        return ev

    for _ in range(args.warmup):
        sim.run(ops, mode=run_mode)
    sim.synchronize()
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    all_ev = []
    for _ in range(args.steps):
        if run_mode == _lib.RUN_PER_GATE:
            all_ev.append(step_events())
        else:
            sim.run(ops, mode=run_mode)
    sim.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    gates_per_s = nops * args.steps / elapsed
    result = {
        "metric": METRIC, "value": gates_per_s, "unit": "gates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"brickwork-{n}q-c128-{args.layers}layers-seed1", "qubits": n,
                   "layers": args.layers, "gates_per_step": nops,
                   "state_GiB": (16 << n) / 2 ** 30,
                   "run_mode": "per_gate" if run_mode == _lib.RUN_PER_GATE else "tiled",
                   "parallelism": "1 GPU"},
        "hbm_TBps_algorithmic": alg_bytes * args.steps / elapsed / 1e12,
    }
    if all_ev:
        dur = []
        for ev in all_ev:
            d = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(nops)])  # ms
            dur.append(d)
        dur = np.stack(dur)
        pair_ms = float(dur[:, is_pair].mean())
        per_launch = 32.0 * (1 << n)
        achieved = per_launch / (pair_ms * 1e-3) / 1e9
        result["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                              "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                              "kernel": "k_pair (uncontrolled X/Y/H/RX/RY pass)",
                              "launch_ms": pair_ms, "bytes_per_launch": per_launch,
                              "launches": int(is_pair.sum()) * args.steps}
        all_bytes = np.array([circuits.gate_bytes(n, int(o["kind"]), int(o["ctrl_mask"]), int(o["target"])) for o in ops])
        result["per_gate_GBps_all_kernels"] = float(all_bytes.sum() / (dur.mean(axis=0).sum() * 1e-3) / 1e9)

    # expectation value of the 64-string Hamiltonian
    t1 = time.perf_counter()
    e = sim.exp_value(ham)
    t_exp = time.perf_counter() - t1
    result["exp_value"] = {"strings": len(ham.coefficients), "ms": 1e3 * t_exp,
                           "GBps": len(ham.coefficients) * (16 << n) / t_exp / 1e9, "value": e}
    sim.close()
    del sim
    torch.cuda.synchronize()

    if args.cpu_gates > 0:
        result["cpu_baseline"] = cpu_baseline(n, ops, args.cpu_gates)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark: 30-qubit complex128 random-circuit throughput on one MI355X.

Workload (one step): |0...0> then 20 layers of [U(theta, phi, lambda) on every
qubit, angles uniform in [0, 2pi), then CX on a random perfect matching] --
900 gates -- with the reference's greedy single-qubit fusion on, evolved in
place in HBM by the engine's register-tile passes. By default each pass runs
on a kernel specialized to its structure and compiled at run time during the
warmup step (reported as first_apply_s, never timed); --tile-kernel register
runs the interpreting kernel instead. The register stays resident across
steps; the timed region excludes host transfers.

Prints one JSON line: gates/s (whole job), the roofline of the dominant
kernel (tiled pass, HIP-event timed), and the reference's OpenMP path timed
on this host on a bounded sample of the same workload. A second measurement
covers the metric's other half: batched adjoint gradients of 256 Pauli-
rotation ansatze (16-24 qubits, 400 rotations, 200-term Hamiltonians) against
parameter-shift gradients on the same GPU (timed on a bounded sample). Two
more objects report the noisy-trajectory workload (24 qubits, 2048 shots,
trajectories/s) and the learned planner (ms per step of the cost model's plan
next to greedy's, on the headline workload, on a deep circuit over 5 of the
30 qubits where dense blocks win, and on random U3/CX/CP layers of depth 40
at 32 qubits, with the final states compared). Further objects
time the QFT workload (20 qubits, 1024 shots, one run per unit), fused dense
blocks of 5 and 6 qubits (complex GEMMs on FP64 MFMA) at 30 qubits against
the FP64 peak, and the noisy workload's density matrix at
12, 14 and 16 qubits, each with the reference beside it.

`traffic` (HBM bytes per tiled pass) comes from the rocprofv3 counter summary
committed under profiles/ for the same kernel and workload, when one exists.
"""
from __future__ import annotations

import os

# The CPU legs run the reference's OpenMP kernels in this process: thread
# binding as the reference's own benchmark plan asks (one thread per core,
# spread), set before any OpenMP runtime is loaded.
os.environ.setdefault("OMP_PROC_BIND", "spread")
os.environ.setdefault("OMP_PLACES", "cores")

import argparse  # noqa: E402
import functools  # noqa: E402
import json  # noqa: E402
import subprocess  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402

PEAK_HBM_GBPS = 8000.0
ROOT = os.path.dirname(os.path.abspath(__file__))


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


@functools.lru_cache(maxsize=1)
def host_cpu() -> dict:
    """The host the CPU legs run on: physical cores and model from lscpu, and
    the OpenMP threads used (one per physical core available to this
    process, capped by OMP_NUM_THREADS when the environment sets it)."""
    model, cores = "unknown", 0
    try:
        info = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in info.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
        table = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
        allowed = os.sched_getaffinity(0)
        phys, phys_allowed = set(), set()
        for line in table.splitlines():
            if line.startswith("#"):
                continue
            cpu, core, sock = (int(x) for x in line.split(",")[:3])
            phys.add((sock, core))
            if cpu in allowed:
                phys_allowed.add((sock, core))
        cores = len(phys)
        threads = len(phys_allowed) or len(allowed)
    except (OSError, ValueError, subprocess.SubprocessError):
        threads = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
    return {"threads": threads, "physical_cores": cores, "cpu_model": model,
            "omp": f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')} OMP_PLACES={os.environ.get('OMP_PLACES')}"}


def cpu_fields(cores: int) -> dict:
    """Common fields of every cpu_baseline object."""
    h = host_cpu()
    return {"cores": cores, "physical_cores": h["physical_cores"], "cpu_model": h["cpu_model"], "omp": h["omp"]}


def cpu_baseline(n: int, layers: int, seed: int, budget_s: float) -> dict | None:
    """Reference (oracle/_ref, compiled from its sources) or, without it, the
    oracle restatement, on a reduced-width sample of the same generator; the
    time is scaled by 2^(n - n_sample) to the full register (both paths are
    memory-bound and linear in 2^n)."""
    try:
        from oracle import oracle as O
        from litequantumcomputingsimulator_amd import workloads as W
    except Exception as e:  # noqa: BLE001
        log(f"cpu baseline unavailable: {e}")
        return None
    cores = host_cpu()["threads"]
    kind = "reference" if O.ref_available() else "port"
    n_s = min(n, 26)
    c = O.Circuit(n_s)
    W.random_u3_cx_layers(c, n_s, 1, seed)
    gates = len(c.insts)
    if kind == "reference":
        O.ref_set_threads(cores)
        run = lambda: O.ref_run_statevector(c, fuse=True)  # noqa: E731
    else:
        run = lambda: O.run_statevector(c, fuse=True)  # noqa: E731
    reps, t_total = 0, 0.0
    while t_total < budget_s and reps < 50:
        t0 = time.perf_counter()
        run()
        t_total += time.perf_counter() - t0
        reps += 1
        log(f"cpu baseline rep {reps}: {t_total:.2f} s")
    per_gate_sample = t_total / (reps * gates)
    per_gate_full = per_gate_sample * (1 << (n - n_s))
    return {
        "value": 1.0 / per_gate_full,
        "unit": "gates/s",
        **cpu_fields(cores),
        "kind": kind,
        "sample": (f"{reps} x run_statevector of 1 layer ({gates} gates) of the same generator at "
                   f"{n_s} qubits, fusion on, OMP threads={cores}; seconds per gate scaled by "
                   f"2^{n - n_s} to {n} qubits"),
    }


def measured_traffic(kernel: str, n: int, layers: int, plan: str, tile: int, low: int) -> tuple[float | None, str | None]:
    """HBM bytes per launch of `kernel` on this workload, from the counter
    summary profiles/<kernel>_random_u3cx_<n>q.json: FETCH_SIZE and WRITE_SIZE
    per dispatch (KiB), each scaled by its copy-kernel calibration factor."""
    name = f"{kernel}_random_u3cx_{n}q.json"
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            pm = json.load(f)
    except (OSError, ValueError):
        return None, None
    want = {"qubits": n, "layers": layers, "plan": plan, "tile_qubits": tile, "low_qubits": low}
    if any(pm.get("workload", {}).get(k) != v for k, v in want.items()):
        return None, None
    kib = pm["fetch_kib_per_dispatch"] * pm["fetch_scale"] + pm["write_kib_per_dispatch"] * pm["write_scale"]
    return kib * 1024.0, f"profiles/{name}"


def adjoint_bench(lq, W, seed: int, n_circuits: int, ps_seconds: float, cpu: bool = True) -> dict:
    """Batched adjoint gradients of n_circuits Pauli-rotation ansatze of uneven
    width (16-24 qubits, 400 rotations each, 200-term Hamiltonians) against
    the parameter-shift gradient, both on this GPU. Parameter shift is timed
    on whole circuits of the batch, in order, until ps_seconds have passed;
    adjoint is timed on the whole batch and on that same sample."""
    import numpy as np

    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(16, 25, size=n_circuits)]
    pcs, thetas, hams = [], [], []
    for i, n in enumerate(sizes):
        rots, theta = W.random_pauli_ansatz(n, 400, seed + 1 + i)
        pc = lq.ParametricCircuit(n)
        for p, qs in rots:
            pc.rp(p, qs)
        pcs.append(pc)
        thetas.append([float(t) for t in theta])
        hams.append(W.random_hamiltonian(n, 200, seed + 100_000 + i))
    lq.algorithms.adjoint_gradient_batched(hams[:2], pcs[:2], thetas[:2])  # warm-up
    lq.synchronize()
    t0 = time.perf_counter()
    full = lq.algorithms.adjoint_gradient_batched(hams, pcs, thetas)
    lq.synchronize()
    t_adj = time.perf_counter() - t0
    n_grads = sum(len(t) for t in thetas)
    log(f"adjoint: {n_circuits} circuits, {n_grads} gradients in {t_adj:.3f} s")
    ps, t_ps, k = [], 0.0, 0
    while k < n_circuits and (k == 0 or t_ps < ps_seconds):
        t0 = time.perf_counter()
        ps.append(lq.algorithms.parameter_shift_gradient(hams[k], pcs[k], thetas[k]))
        lq.synchronize()
        t_ps += time.perf_counter() - t0
        k += 1
        log(f"parameter shift: circuit {k} ({sizes[k - 1]} qubits) done, {t_ps:.2f} s")
    t0 = time.perf_counter()
    same = lq.algorithms.adjoint_gradient_batched(hams[:k], pcs[:k], thetas[:k])
    lq.synchronize()
    t_adj_sample = time.perf_counter() - t0
    err = max(float(np.max(np.abs(np.array(g) - np.array(p)))) for (_, g), p in zip(same, ps))
    err_full = max(float(np.max(np.abs(np.array(g) - np.array(h)))) for (_, g), (_, h) in zip(same, full[:k]))
    sample_grads = sum(len(t) for t in thetas[:k])
    out = {
        "workload": (f"{n_circuits} Pauli-rotation ansatze, 16-24 qubits (uneven), 400 rotations of weight 2-8, "
                     "200-term Hamiltonians, theta ~ U(-pi, pi)"),
        "adjoint_gradient_vectors_per_s": round(n_circuits / t_adj, 2),
        "adjoint_components_per_s": round(n_grads / t_adj, 1),
        "adjoint_s": round(t_adj, 3),
        "param_shift_gradient_vectors_per_s": round(k / t_ps, 3),
        "param_shift_components_per_s": round(sample_grads / t_ps, 2),
        "param_shift_sample": f"first {k} circuit(s) of the batch, widths {sizes[:k]}, {sample_grads} components",
        "adjoint_components_per_s_on_sample": round(sample_grads / t_adj_sample, 1),
        "speedup_on_sample": round((sample_grads / t_adj_sample) / (sample_grads / t_ps), 1),
        "max_abs_diff_vs_param_shift": err,
        "max_abs_diff_batch_vs_sample_run": err_full,
    }
    if cpu:
        out["cpu_baseline"] = _ps_cpu_baseline(W, sizes, seed)
    return out


def adjoint_plan_bench(lq, W, seed: int, reps: int = 20) -> dict:
    """Captured adjoint plans against the uncaptured calls at widths where
    launches and host preparation matter: per-call wall time (host work, the
    launches and the copies back) of one 16-qubit circuit (400 rotations of
    weight 2-8, 200 terms: an optimiser loop's step) and of a batch of 16
    circuits of 12-16 qubits (100 rotations, 50 terms each), at new angles
    every call, results compared bit for bit."""
    import numpy as np

    def batch(widths, n_rot, n_terms, s0):
        pcs, hams = [], []
        for i, n in enumerate(widths):
            rots, _ = W.random_pauli_ansatz(n, n_rot, s0 + i)
            pc = lq.ParametricCircuit(n)
            for p, qs in rots:
                pc.rp(p, qs)
            pcs.append(pc)
            hams.append(W.random_hamiltonian(n, n_terms, s0 + 1000 + i))
        return pcs, hams

    rng = np.random.default_rng(seed)
    cases = {"one_16q_400rot": batch([16], 400, 200, seed + 7),
             "batch16_12to16q_100rot": batch([int(x) for x in rng.integers(12, 17, size=16)], 100, 50, seed + 70)}
    out = {}
    for name, (pcs, hams) in cases.items():
        angles = [[[float(v) for v in rng.uniform(-np.pi, np.pi, pc.num_parameters())] for pc in pcs]
                  for _ in range(reps + 1)]
        t0 = time.perf_counter()
        plan = lq.algorithms.AdjointPlan(hams, pcs)
        lq.synchronize()
        t_create = time.perf_counter() - t0
        plan.run(angles[0])
        lq.algorithms.adjoint_gradient_batched(hams, pcs, angles[0])
        lq.synchronize()
        t0 = time.perf_counter()
        eager = [lq.algorithms.adjoint_gradient_batched(hams, pcs, a) for a in angles[1:]]
        t_eager = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        captured = [plan.run(a) for a in angles[1:]]
        t_plan = (time.perf_counter() - t0) / reps
        same = all(eg == ew and np.array_equal(np.array(gg), np.array(gw))
                   for rg, rw in zip(captured, eager) for (eg, gg), (ew, gw) in zip(rg, rw))
        out[name] = {"widths": [pc.num_qubits() for pc in pcs], "calls": reps,
                     "uncaptured_ms_per_call": round(t_eager * 1e3, 3),
                     "captured_ms_per_call": round(t_plan * 1e3, 3),
                     "speedup": round(t_eager / t_plan, 2), "plan_create_ms": round(t_create * 1e3, 2),
                     "bitwise_equal": bool(same)}
        log(f"adjoint plan {name}: {t_eager * 1e3:.2f} -> {t_plan * 1e3:.2f} ms/call, equal={same}")
        del plan
    return out


def _ps_cpu_baseline(W, sizes, seed: int, budget_s: float = 30.0) -> dict | None:
    """The reference's parameter_shift_gradient (vqe.cpp:68-94, OpenMP on the
    host cores) for the batch's first 16-qubit circuit: 2 x 400 evaluations.
    Two evaluations (run_statevector + expectation) are timed first; when the
    full gradient would exceed budget_s, those stand in, scaled by 400."""
    try:
        from oracle import oracle as O
    except Exception as e:  # noqa: BLE001
        log(f"parameter-shift cpu baseline unavailable: {e}")
        return None
    i = next((j for j, n in enumerate(sizes) if n == 16), 0)
    n = sizes[i]
    rots, theta = W.random_pauli_ansatz(n, 400, seed + 1 + i)
    h = W.random_hamiltonian(n, 200, seed + 100_000 + i)
    t = O.Circuit(n)
    for (p, qs), th in zip(rots, theta):
        t.rp(float(th), p, qs)
    ref = O.ref_available()
    kind = "reference" if ref else "port"
    cores = host_cpu()["threads"]
    if ref:
        O.ref_set_threads(cores)
    run = O.ref_run_statevector if ref else O.run_statevector
    expect = O.ref_expectation if ref else (lambda st, nq, terms: O.expectation(st, nq, terms))
    t0 = time.perf_counter()
    for _ in range(2):
        expect(run(t, fuse=False), n, h)
    t_two = time.perf_counter() - t0
    if 400 * t_two <= budget_s:
        theta_l = [float(v) for v in theta]
        t0 = time.perf_counter()
        if ref:
            O.ref_parameter_shift_gradient(t, list(range(400)), theta_l, h)
        else:
            O.parameter_shift_gradient(t, list(range(400)), theta_l, h)
        per_vector = time.perf_counter() - t0
        sample = f"parameter_shift_gradient of circuit {i} ({n} qubits, 400 rotations, 200 terms): {per_vector:.2f} s"
    else:
        per_vector = 400 * t_two
        sample = (f"2 evaluations (run_statevector + 200-term expectation) of circuit {i} ({n} qubits, 400 "
                  f"rotations): {t_two:.3f} s, scaled by 400 to the 800 evaluations of one gradient")
    log(f"parameter-shift cpu baseline: {per_vector:.2f} s per gradient vector at {n} qubits ({kind})")
    return {"value": round(1.0 / per_vector, 5), "unit": "gradient vectors/s (400 parameters, one circuit)",
            **cpu_fields(cores), "kind": kind, "sample": sample}


def tile_roofline(prof: dict, traffic_json: str | None = None, kernel: str = "k_rtile") -> dict | None:
    """Roofline of the tiled passes in a launch profile: every pass reads and
    writes each amplitude of the registers it covers once (32 bytes per
    amplitude), timed by HIP events around each launch on the engine stream."""
    if not prof.get("tile_launches") or not prof.get("tile_ms"):
        return None
    by = 32.0 * prof["tile_amps"]
    achieved = by / (prof["tile_ms"] * 1e6)
    traffic, source = None, None
    if traffic_json:
        path = os.path.join(ROOT, "profiles", traffic_json)
        try:
            with open(path) as f:
                pm = json.load(f)
            kib = pm["fetch_kib_per_dispatch"] * pm["fetch_scale"] + pm["write_kib_per_dispatch"] * pm["write_scale"]
            traffic, source = kib * 1024.0, f"profiles/{traffic_json}"
        except (OSError, ValueError, KeyError):
            pass
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": traffic, "traffic_source": source,
            "kernel": kernel, "bytes_per_launch": by / prof["tile_launches"],
            "avg_launch_ms": round(prof["tile_ms"] / prof["tile_launches"], 4),
            "launches": prof["tile_launches"], "share_of_kernel_ms":
                round(prof["tile_ms"] / (prof["tile_ms"] + prof["other_ms"]), 4)}


def trajectories_bench(lq, W, seed: int, n: int, shots: int, cpu: bool = True, cpu_shots: int = 1) -> dict:
    """The noisy-trajectory workload: GHZ prep + 10 random U3/CZ layers at n
    qubits, depolarizing(1e-3) then amplitude_damping(1e-3) after every gate,
    measure_all, `shots` trajectories in batched launches."""
    qc = lq.QuantumCircuit(n, n)
    W.noisy_ghz_layers(qc, n, 10, seed)
    qc.measure_all()
    noise = lq.NoiseModel()
    noise.add_all_qubit_channel(lq.KrausChannel.depolarizing(1e-3))
    noise.add_all_qubit_channel(lq.KrausChannel.amplitude_damping(1e-3))
    sim = lq.StatevectorSimulator(seed=seed, noise=noise)
    sim.run(qc, 8)  # warm-up
    lq.synchronize()
    t0 = time.perf_counter()
    counts = sim.run(qc, shots).counts()
    lq.synchronize()
    dt = time.perf_counter() - t0
    # the same run again with every launch bracketed by events (not the timed one)
    lq.profile_enable(True)
    lq.profile_read()
    sim.run(qc, shots)
    prof = lq.profile_read()
    lq.profile_enable(False)
    gates = qc.num_instructions() - n
    log(f"trajectories: {shots} shots of {n} qubits in {dt:.3f} s")
    out = {"workload": (f"GHZ prep + 10 random U3/CZ layers, {n} qubits, {gates} gates, depolarizing(1e-3) + "
                        f"amplitude_damping(1e-3) after every gate, measure_all"),
           "shots": shots, "seconds": round(dt, 3), "trajectories_per_s": round(shots / dt, 2),
           "trajectory_gates_per_s": round(shots * gates / dt, 1), "distinct_outcomes": len(dict(counts)),
           "roofline": tile_roofline(prof, f"k_rtile_trajectories_{n}q.json")}
    if cpu:
        try:
            from oracle import oracle as O
            c = O.Circuit(n, n)
            W.noisy_ghz_layers(c, n, 10, seed)
            c.measure_all()
            chans = [O.depolarizing(1e-3), O.amplitude_damping(1e-3)]
            kind = "reference" if O.ref_available() else "port"
            cores = host_cpu()["threads"]
            t0 = time.perf_counter()
            if kind == "reference":
                O.ref_set_threads(cores)
                O.ref_run_counts(c, seed, cpu_shots, channels=chans)
            else:
                O.run_counts(c, seed, cpu_shots, channels=chans)
            dtc = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": round(cpu_shots / dtc, 4), "unit": "trajectories/s", **cpu_fields(cores),
                                   "kind": kind,
                                   "sample": f"run() of {cpu_shots} shot(s) of the same circuit and noise, {dtc:.2f} s"}
            log(f"trajectories cpu baseline: {cpu_shots} shot(s) in {dtc:.2f} s ({kind})")
        except Exception as e:  # noqa: BLE001
            log(f"trajectories cpu baseline unavailable: {e}")
    return out


def trajectories_small_bench(lq, W, seed: int, shots: int = 2048, widths=(8, 12)) -> dict:
    """The noisy-trajectory workload at widths whose register fits one
    workgroup's LDS (the 12-qubit variant is the one cross-checked against the
    density matrix): the shared tiled runs over a batch in HBM against the
    persistent kernel (one launch, a workgroup per shot), wall time of run()
    for `shots` trajectories, counts compared."""
    out = {}
    noise = lq.NoiseModel()
    noise.add_all_qubit_channel(lq.KrausChannel.depolarizing(1e-3))
    noise.add_all_qubit_channel(lq.KrausChannel.amplitude_damping(1e-3))
    for n in widths:
        qc = lq.QuantumCircuit(n, n)
        W.noisy_ghz_layers(qc, n, 10, seed)
        qc.measure_all()
        row, got = {}, {}
        for kernel in ("shared", "persistent"):
            sim = lq.StatevectorSimulator(seed=seed, noise=noise, trajectory_kernel=kernel)
            sim.run(qc, 8)  # warm-up
            lq.synchronize()
            t0 = time.perf_counter()
            got[kernel] = dict(sim.run(qc, shots).counts())
            lq.synchronize()
            dt = time.perf_counter() - t0
            row[f"{kernel}_trajectories_per_s"] = round(shots / dt, 1)
            log(f"trajectories {n}q {kernel}: {shots} shots in {dt * 1e3:.1f} ms")
        row["speedup"] = round(row["persistent_trajectories_per_s"] / row["shared_trajectories_per_s"], 2)
        row["counts_equal"] = got["shared"] == got["persistent"]
        out[f"{n}q"] = row
    return out


def density_matrix_bench(lq, W, seed: int, widths=(12, 14, 16), cpu: bool = True, cpu_width: int = 10,
                         tile_kernel: str = "specialized") -> dict:
    """The noisy workload's exact form: rho of GHZ prep + 10 random U3/CZ
    layers with depolarizing(1e-3) then amplitude_damping(1e-3) after every
    gate, on the device (a 2n-qubit register, tiled passes), at the
    reference's 12-qubit cross-check width and beyond its 13-qubit cap. The
    reference's DensityMatrixSimulator runs the same circuit on the host cores
    at `cpu_width` qubits, scaled by 4^(12 - cpu_width) (its passes are linear
    in 4^n)."""
    noise = lq.NoiseModel()
    noise.add_all_qubit_channel(lq.KrausChannel.depolarizing(1e-3))
    noise.add_all_qubit_channel(lq.KrausChannel.amplitude_damping(1e-3))
    out = {"workload": "rho of GHZ prep + 10 random U3/CZ layers, depolarizing(1e-3) + amplitude_damping(1e-3) "
                       "after every gate", "tile_kernel": tile_kernel}
    kname = "lqcs_pass" if tile_kernel == "specialized" else "k_rtile"
    for n in widths:
        qc = lq.QuantumCircuit(n, n)
        W.noisy_ghz_layers(qc, n, 10, seed)
        sim = lq.DensityMatrixSimulator(noise=noise, max_qubits=16, tile_kernel=tile_kernel)
        t0 = time.perf_counter()
        sim.run_density_matrix(qc)  # warm-up (compiles specialized passes)
        lq.synchronize()
        first_s = time.perf_counter() - t0
        lq.synchronize()
        lq.profile_enable(True)
        lq.profile_read()
        t0 = time.perf_counter()
        rho = sim.run_density_matrix(qc)
        lq.synchronize()
        dt = time.perf_counter() - t0
        prof = lq.profile_read()
        lq.profile_enable(False)
        passes = prof["tile_launches"] + prof["other_launches"]
        kms = prof["tile_ms"] + prof["other_ms"]
        out[f"n{n}"] = {"seconds": round(dt, 4), "first_run_s": round(first_s, 3), "gates": qc.num_instructions(),
                        "launches": passes,
                        "kernel_ms": round(kms, 3), "trace": rho.trace(),
                        "GB/s": round(passes * 32 * 4**n / (kms * 1e6), 1) if kms else None,
                        "roofline": tile_roofline(prof, f"{kname}_density_matrix_{n}q.json", kname)}
        log(f"density matrix {n} qubits: {dt:.3f} s, {passes} launches, {kms:.1f} ms of kernels")
        del rho
    out["roofline"] = out[f"n{widths[-1]}"]["roofline"]
    if cpu:
        try:
            from oracle import oracle as O
            c = O.Circuit(cpu_width, cpu_width)
            W.noisy_ghz_layers(c, cpu_width, 10, seed)
            chans = [O.depolarizing(1e-3), O.amplitude_damping(1e-3)]
            kind = "reference" if O.ref_available() else "port"
            cores = host_cpu()["threads"]
            t0 = time.perf_counter()
            if kind == "reference":
                O.ref_set_threads(cores)
                O.ref_density_probabilities(c, chans)
            else:
                O.density_probabilities(c, chans)
            dt = time.perf_counter() - t0
            scaled = dt * 4 ** (12 - cpu_width)
            out["cpu_baseline"] = {"value": round(scaled, 3), "unit": "seconds for rho at 12 qubits", **cpu_fields(cores),
                                   "kind": kind,
                                   "sample": f"run_density_matrix of the same generator at {cpu_width} qubits "
                                             f"({dt:.2f} s), scaled by 4^{12 - cpu_width}"}
            if "n12" in out:
                out["n12"]["speedup_vs_cpu"] = round(scaled / out["n12"]["seconds"], 1)
            log(f"density matrix cpu baseline: {dt:.2f} s at {cpu_width} qubits ({kind})")
        except Exception as e:  # noqa: BLE001
            log(f"density matrix cpu baseline unavailable: {e}")
    return out


def qft_bench(lq, W, seed: int, n: int = 20, shots: int = 1024, reps: int = 20, cpu: bool = True) -> dict:
    """The QFT workload: X on the bits of a seeded x, then algorithms::qft(n)
    (220 gates at 20 qubits), measure_all, fusion off, `shots` shots. One run
    (op upload, evolve, sampling, counts) is the unit; the fastest of `reps`.
    The counts are compared with the reference's run on the same seed (or the
    oracle's) and the reference's run is timed beside it."""
    import numpy as np

    x = int(np.random.default_rng(seed).integers(0, 1 << n))
    qc = lq.QuantumCircuit(n, n)
    W.basis_prep(qc, n, x)
    qc.compose(lq.algorithms.qft(n))
    gates = qc.num_instructions()
    qc.measure_all()
    sim = lq.StatevectorSimulator(seed=seed, fuse_gates=False)
    counts = dict(sim.run(qc, shots).counts())  # warm-up
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        sim.run(qc, shots)
        best = min(best, time.perf_counter() - t0)
    out = {"workload": f"QFT from a basis state, {n} qubits, {gates} gates, measure_all, {shots} shots, fusion off",
           "run_ms": round(best * 1e3, 3), "runs_per_s": round(1.0 / best, 1),
           "gates_per_s": round(gates / best, 1)}
    try:
        from oracle import oracle as O
    except Exception as e:  # noqa: BLE001
        log(f"qft reference unavailable: {e}")
        return out
    ref = O.Circuit(n)
    W.basis_prep(ref, n, x)
    W.qft_into(ref, n)
    ref.measure_all()
    if O.ref_available():
        cores = host_cpu()["threads"]
        O.ref_set_threads(cores)
        t0 = time.perf_counter()
        want = O.ref_run_counts(ref, seed, shots, fuse=False)
        t_ref = time.perf_counter() - t0
        out["counts_equal_reference"] = counts == want
        if cpu:
            out["cpu_baseline"] = {"value": round(1.0 / t_ref, 3), "unit": "runs/s", **cpu_fields(cores),
                                   "kind": "reference", "sample": f"one run() of the same circuit and seed, "
                                                                  f"OMP threads={cores}"}
    else:
        out["counts_equal_oracle"] = counts == O.run_counts(ref, seed, shots, fuse=False)
    log(f"qft{n}: {best * 1e3:.2f} ms per run of {shots} shots")
    return out


def deep_local_bench(lq, sv, n: int, seed: int) -> dict:
    """A deep circuit over a window of 5 of the n qubits (the cost model's
    deep_local family): greedy tiled passes, dense 5-qubit blocks and the
    learned plan, all on specialized passes, each the fastest of 3 applies;
    the learned plan's state against greedy's."""
    from litequantumcomputingsimulator_amd import costmodel as CM

    qc = CM.random_circuit(lq, n, "deep_local", 6, seed)
    comp = lq.CompiledCircuit(qc, True)
    out = {"workload": f"{qc.num_instructions()} U3/CX on a window of 5 of {n} qubits, greedy 1q fusion"}
    for name, kw in (("greedy_12q_low4", dict(plan="greedy", tile_qubits=12, low_qubits=4)),
                     ("dense_5q", dict(plan="dense", tile_qubits=5)), ("learned", dict(plan="learned"))):
        kw["tile_kernel"] = "specialized"
        comp.apply(sv, **kw)  # compiles its passes
        lq.synchronize()
        best = float("inf")
        for _ in range(3):
            sv.set_basis(0)
            lq.synchronize()
            t0 = time.perf_counter()
            comp.apply(sv, **kw)
            lq.synchronize()
            best = min(best, (time.perf_counter() - t0) * 1e3)
        out[name] = {"ms": round(best, 3), "passes": lq.last_plan_stats()["passes"]}
        if name == "learned":
            out[name]["pick"] = lq.last_plan_choice()
    ref = lq.statevector(n)
    comp.apply(ref, plan="greedy", tile_qubits=12, low_qubits=4, tile_kernel="specialized")
    out["max_abs_diff_learned_vs_greedy"] = ref.max_abs_diff(sv)
    del ref
    out["learned_vs_greedy"] = round(out["learned"]["ms"] / out["greedy_12q_low4"]["ms"], 4)
    log(f"deep local: greedy {out['greedy_12q_low4']['ms']:.2f} ms, dense {out['dense_5q']['ms']:.2f} ms, "
        f"learned {out['learned']['ms']:.2f} ms ({out['learned']['pick']})")
    return out


def planner_32q(lq, W, seed: int, n: int = 32, depth: int = 40) -> dict:
    """The learned-planner workload: random U3/CX/CP layers of depth 40 at 32
    qubits (64 GiB per register); applies with greedy grouping and with the
    learned plan on two registers, three of each in alternating order, the
    fastest of each reported, and the two states compared on the device."""
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_cp(qc, n, depth, seed)
    comp = lq.CompiledCircuit(qc, True)
    a = lq.statevector(n)
    b = lq.statevector(n)
    comp.apply(a, plan="greedy")  # untimed: stages both plans on the device
    comp.apply(b, plan="learned")
    a.set_basis(0)
    b.set_basis(0)
    lq.synchronize()
    t_greedy = t_learned = float("inf")
    for rep in range(3):  # alternating order: greedy first, then learned first, ...
        if rep:
            a.set_basis(0)
            b.set_basis(0)
            lq.synchronize()
        for plan in (("greedy", "learned") if rep % 2 == 0 else ("learned", "greedy")):
            reg = a if plan == "greedy" else b
            t0 = time.perf_counter()
            comp.apply(reg, plan=plan)
            lq.synchronize()
            dt = time.perf_counter() - t0
            if plan == "greedy":
                t_greedy = min(t_greedy, dt)
            else:
                t_learned = min(t_learned, dt)
    choice = lq.last_plan_choice()
    diff = a.max_abs_diff(b)
    log(f"planner at {n} qubits: greedy {t_greedy:.3f} s, learned {t_learned:.3f} s ({choice}), max diff {diff:.2e}")
    out = {"workload": f"random U3/CX/CP layers, {n} qubits, depth {depth}, {qc.num_instructions()} gates",
           "kernel": "k_rtile", "greedy_s": round(t_greedy, 3), "learned_s": round(t_learned, 3), "learned_pick": choice,
           "learned_vs_greedy": round(t_learned / t_greedy, 4),
           "same_plan_as_greedy": choice["plan"] == "greedy" and choice["low_qubits"] == 4,
           "max_abs_diff_states": diff}
    # The same on specialized passes (the specialized-pass model picks among
    # 12- and 13-qubit tiles); first applies compile, then alternate.
    ks = dict(tile_kernel="specialized")
    t0 = time.perf_counter()
    comp.apply(a, plan="greedy", **ks)
    comp.apply(b, plan="learned", **ks)
    lq.synchronize()
    compile_s = time.perf_counter() - t0
    schoice = lq.last_plan_choice()  # the learned pick (later applies of the same program reuse it)
    sg = sl = float("inf")
    for rep in range(2):
        for plan in (("greedy", "learned") if rep % 2 == 0 else ("learned", "greedy")):
            reg = a if plan == "greedy" else b
            reg.set_basis(0)
            lq.synchronize()
            t0 = time.perf_counter()
            comp.apply(reg, plan=plan, **ks)
            lq.synchronize()
            dt = time.perf_counter() - t0
            if plan == "greedy":
                sg = min(sg, dt)
            else:
                sl = min(sl, dt)
    sdiff = a.max_abs_diff(b)
    log(f"planner at {n} qubits, specialized: greedy {sg:.3f} s, learned {sl:.3f} s ({schoice}), max diff {sdiff:.2e}")
    out["specialized"] = {"greedy_s": round(sg, 3), "learned_s": round(sl, 3), "learned_pick": schoice,
                          "learned_vs_greedy": round(sl / sg, 4), "first_applies_s": round(compile_s, 3),
                          "max_abs_diff_states": sdiff}
    return out


PEAK_FP64_TFLOPS = 78.6


def dense_blocks_bench(lq, seed: int, n: int = 30, reps: int = 4, cpu_seconds: float = 8.0,
                       cpu: bool = True) -> dict:
    """Fused dense blocks: the reference's generic controlled-unitary path
    (src/backend/kernels/gate_kernels.cpp:143-167) for k = 5 and 6 targets,
    here complex GEMMs on FP64 MFMA. `reps` Haar-random k-qubit unitaries on
    qubits 8..8+k of an n-qubit register, one launch each, HIP-event timed.
    Per launch: 32 * 2^n bytes (one read and one write of every amplitude) and
    (8 * 2^k - 2) * 2^n flop; k = 6 is bound by FP64 MFMA, k = 5 by HBM. The
    reference runs the same circuit on the host cores at 24 qubits, scaled by
    2^(n - 24)."""
    import numpy as np

    rng = np.random.default_rng(seed)

    def haar(k):
        d = 1 << k
        q, r = np.linalg.qr(rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
        return q * (np.diag(r) / np.abs(np.diag(r)))

    out = {"workload": f"{reps} Haar-random k-qubit unitaries on qubits 8..8+k, {n} qubits, one launch each",
           "kernel": "k_cu_mfma"}
    sv = lq.statevector(n)
    mats = {}
    for k in (5, 6):
        mats[k] = [haar(k) for _ in range(reps)]
        qc = lq.QuantumCircuit(n)
        for m in mats[k]:
            qc.unitary(m, list(range(8, 8 + k)))
        comp = lq.CompiledCircuit(qc, False)
        comp.apply(sv, plan="gate_by_gate")  # untimed
        lq.synchronize()
        lq.profile_enable(True)
        lq.profile_read_launches()
        comp.apply(sv, plan="gate_by_gate")
        ms, _ = lq.profile_read_launches()
        lq.profile_enable(False)
        t = float(np.mean(ms))
        gbs = 32 * 2**n / t / 1e6
        tflops = (8 * 2**k - 2) * 2**n / t / 1e9
        out[f"k{k}"] = {"avg_launch_ms": round(t, 4), "GB/s": round(gbs, 1), "TFLOP/s": round(tflops, 2),
                        "hbm_frac": round(gbs / PEAK_HBM_GBPS, 4), "fp64_frac": round(tflops / PEAK_FP64_TFLOPS, 4)}
        log(f"dense k={k}: {t:.3f} ms per launch, {gbs:.0f} GB/s, {tflops:.1f} TFLOP/s")
    k6 = out["k6"]
    traffic, src = None, None
    try:
        with open(os.path.join(ROOT, "profiles", f"k_cu_mfma_dense_{n}q.json")) as f:
            pm = json.load(f)
        if pm.get("workload", {}).get("qubits") == n and pm.get("workload", {}).get("blocks") == reps:
            traffic, src = pm["per_k"]["6"]["traffic_bytes_per_dispatch"], f"profiles/k_cu_mfma_dense_{n}q.json"
    except (OSError, ValueError, KeyError):
        pass
    out["roofline"] = {"bound": "mfma", "achieved": k6["TFLOP/s"], "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                       "frac": k6["fp64_frac"], "traffic": traffic, "traffic_source": src, "kernel": "k_cu_mfma",
                       "flop_per_launch": (8 * 64 - 2) * 2**n, "avg_launch_ms": k6["avg_launch_ms"]}
    del sv
    if cpu:
        try:
            from oracle import oracle as O
            n_s = 24
            c = O.Circuit(n_s)
            for m in mats[6]:
                c.unitary(m, list(range(8, 14)))
            cores = host_cpu()["threads"]
            kind = "reference" if O.ref_available() else "port"
            if kind == "reference":
                O.ref_set_threads(cores)
                run = lambda: O.ref_run_statevector(c, fuse=False)  # noqa: E731
            else:
                run = lambda: O.run_statevector(c, fuse=False)  # noqa: E731
            reps_done, t_total = 0, 0.0
            while t_total < cpu_seconds and reps_done < 20:
                t0 = time.perf_counter()
                run()
                t_total += time.perf_counter() - t0
                reps_done += 1
            per_gate = t_total / (reps_done * reps) * (1 << (n - n_s))
            out["cpu_baseline"] = {"value": round(1e3 * per_gate, 3), "unit": "ms per k=6 block at 30 qubits",
                                   **cpu_fields(cores), "kind": kind,
                                   "sample": f"{reps_done} x run_statevector of the {reps} k=6 blocks at {n_s} "
                                             f"qubits; seconds per block scaled by 2^{n - n_s}"}
            log(f"dense cpu baseline: {1e3 * per_gate:.1f} ms per k=6 block at {n} qubits ({kind})")
        except Exception as e:  # noqa: BLE001
            log(f"dense cpu baseline unavailable: {e}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=20)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--plan", default="greedy")
    ap.add_argument("--tile-qubits", type=int, default=0)
    ap.add_argument("--low-qubits", type=int, default=0)
    ap.add_argument("--tile-kernel", default="specialized",
                    help="kernel of the tiled passes: specialized (one kernel per pass structure, compiled at run "
                         "time during the warmup step) or register (the interpreting register tile)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--adjoint-circuits", type=int, default=256)
    ap.add_argument("--ps-seconds", type=float, default=20.0)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--traj-qubits", type=int, default=24)
    ap.add_argument("--traj-shots", type=int, default=2048)
    ap.add_argument("--no-trajectories", action="store_true")
    ap.add_argument("--no-density-matrix", action="store_true",
                    help="skip the density-matrix workload (12, 14 and 16 qubits)")
    ap.add_argument("--no-qft", action="store_true", help="skip the QFT workload (20 qubits, 1024 shots)")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the dense-block workload (k = 5, 6 targets on FP64 MFMA)")
    ap.add_argument("--no-planner", action="store_true",
                    help="skip timing the learned plan on the same workload")
    ap.add_argument("--no-planner-32q", action="store_true",
                    help="skip the learned-planner workload at 32 qubits (two 64 GiB registers)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="accepted for the counter-calibration commands; the device-to-device copies of the "
                         "register that calibrate the byte counters run in every invocation")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(the statevector would shard by its top log2(N) qubits; see DESIGN.md)", file=sys.stderr)
        return 2

    from litequantumcomputingsimulator_amd import pylqcs as lq
    from litequantumcomputingsimulator_amd import workloads as W

    n, layers = args.qubits, args.layers
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_layers(qc, n, layers, args.seed)
    gates = qc.num_instructions()
    comp = lq.CompiledCircuit(qc, True)  # reference greedy single-qubit fusion
    log(f"device: {lq.device_info()}; {gates} gates -> {comp.num_ops()} ops after fusion")
    sv = lq.statevector(n)

    def step():
        sv.set_basis(0)
        comp.apply(sv, plan=args.plan, tile_qubits=args.tile_qubits, low_qubits=args.low_qubits,
                   tile_kernel=args.tile_kernel)

    # The first apply plans the circuit and, for specialized passes, compiles
    # one kernel per distinct pass structure; it is reported, not timed.
    first_apply_s = None
    for w in range(args.warmup):
        t0 = time.perf_counter()
        step()
        lq.synchronize()
        if first_apply_s is None:
            first_apply_s = time.perf_counter() - t0
        log(f"warmup {w + 1}/{args.warmup} done ({time.perf_counter() - t0:.3f} s)")
    stats = lq.last_plan_stats()

    lq.profile_enable(True)
    lq.profile_read()
    lq.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
        if k + 1 < args.steps:
            lq.synchronize()
            log(f"step {k + 1}/{args.steps}: {time.perf_counter() - t0:.3f} s")
    lq.synchronize()
    elapsed = time.perf_counter() - t0
    prof = lq.profile_read()
    lq.profile_enable(False)
    log(f"timed {args.steps} steps in {elapsed:.3f} s; launches {prof}")

    norm = sv.norm()
    if abs(norm - 1.0) > 1e-9:
        log(f"norm check failed: {norm}")
        return 1

    # The on-box copy rate: device-to-device copies of the register (one read
    # and one write of every amplitude, like a pass), after one untimed copy.
    # Under --copy-calibration the same copies calibrate the byte counters.
    pass_bytes = 32.0 * (1 << n)  # one read + one write of 2^n complex128
    sv2 = lq.statevector(n)
    sv2.copy_from(sv)
    lq.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        sv2.copy_from(sv)
    lq.synchronize()
    copy_s = (time.perf_counter() - t0) / 3
    copy_gbps = pass_bytes / copy_s / 1e9
    log(f"device-to-device copy of the register: {copy_s * 1e3:.3f} ms ({copy_gbps:.0f} GB/s read + write)")
    del sv2
    if prof["tile_launches"]:
        avg_s = prof["tile_ms"] / 1e3 / prof["tile_launches"]
        # every tiled pass of this plan runs on the selected register-tile kernel
        kernel = "lqcs_pass" if args.tile_kernel == "specialized" else "k_rtile"
    else:
        avg_s = prof["other_ms"] / 1e3 / max(1, prof["other_launches"])
        kernel = "gate-by-gate kernels"
    achieved = pass_bytes / avg_s / 1e9
    tile_q = args.tile_qubits or 12
    low_q = args.low_qubits or 4
    traffic, traffic_src = measured_traffic(kernel, n, layers, args.plan, tile_q, low_q)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBPS, 4),
                "traffic": round(traffic) if traffic is not None else None, "traffic_source": traffic_src,
                "kernel": kernel, "bytes_per_launch": pass_bytes, "avg_launch_ms": round(avg_s * 1e3, 4),
                "launches_per_step": prof["tile_launches"] // max(1, args.steps),
                "copy_gbps": round(copy_gbps, 1), "frac_of_copy": round(achieved / copy_gbps, 4)}

    planner = None
    if not args.no_planner:
        comp.apply(sv, plan="learned")  # warm-up (and the cost model's pick)
        choice = lq.last_plan_choice()
        lq.synchronize()
        # Alternating steps of the two plans, three of each; the faster of each.
        greedy_ms = learned_ms = float("inf")
        for _ in range(3):
            for plan in ("greedy", "learned"):
                t0 = time.perf_counter()
                sv.set_basis(0)
                comp.apply(sv, plan=plan)
                lq.synchronize()
                ms = (time.perf_counter() - t0) * 1e3
                if plan == "greedy":
                    greedy_ms = min(greedy_ms, ms)
                else:
                    learned_ms = min(learned_ms, ms)
        planner = {"kernel": "k_rtile (interpreting register tile: the kernel the cost model is trained on)",
                   "greedy_ms_per_step": round(greedy_ms, 3),
                   "learned_ms_per_step": round(learned_ms, 3), "learned_pick": choice,
                   "learned_vs_greedy": round(learned_ms / greedy_ms, 4)}
        log(f"planner: learned {learned_ms:.1f} ms/step ({choice})")
        # Other plans on the specialized passes: gain grouping at 3 low qubits
        # on 12- and 13-qubit tiles (13 runs only on specialized passes).
        spec = {}
        for name, L, plan, low in (("greedy_12q_low4", 12, "greedy", 4), ("gain_12q_low3", 12, "gain", 3),
                                   ("gain_13q_low3", 13, "gain", 3), ("learned", 0, "learned", 0)):
            kw = dict(plan=plan, tile_qubits=L, low_qubits=low, tile_kernel="specialized")
            t0 = time.perf_counter()
            comp.apply(sv, **kw)  # compiles its passes
            lq.synchronize()
            compile_s = time.perf_counter() - t0
            best = float("inf")
            for _ in range(3):
                sv.set_basis(0)
                lq.synchronize()
                t0 = time.perf_counter()
                comp.apply(sv, **kw)
                lq.synchronize()
                best = min(best, (time.perf_counter() - t0) * 1e3)
            spec[name] = {"ms_per_step": round(best, 3), "gates_per_s": round(gates / best * 1e3, 1),
                          "passes": lq.last_plan_stats()["passes"], "first_apply_s": round(compile_s, 3)}
            if plan == "learned":
                spec[name]["pick"] = lq.last_plan_choice()
            log(f"specialized {name}: {best:.1f} ms/step")
        spec["learned_vs_greedy"] = round(spec["learned"]["ms_per_step"] / spec["greedy_12q_low4"]["ms_per_step"], 4)
        planner["specialized_plans"] = spec
        planner["deep_local"] = deep_local_bench(lq, sv, n, args.seed)

    adjoint = None
    trajectories = None
    del sv
    if not args.no_adjoint:
        adjoint = adjoint_bench(lq, W, args.seed, args.adjoint_circuits, args.ps_seconds, cpu=not args.no_cpu_baseline)
        adjoint["captured_plans"] = adjoint_plan_bench(lq, W, args.seed)
    if not args.no_trajectories:
        trajectories = trajectories_bench(lq, W, args.seed, args.traj_qubits, args.traj_shots,
                                          cpu=not args.no_cpu_baseline)
        trajectories["small_widths"] = trajectories_small_bench(lq, W, args.seed, args.traj_shots)
    if planner is not None and not args.no_planner_32q:
        planner["at_32_qubits"] = planner_32q(lq, W, args.seed)
    dm = None
    if not args.no_density_matrix:
        dm = density_matrix_bench(lq, W, args.seed, cpu=not args.no_cpu_baseline, tile_kernel=args.tile_kernel)
    qft = None if args.no_qft else qft_bench(lq, W, args.seed, cpu=not args.no_cpu_baseline)
    dense = None
    if not args.no_dense:
        dense = dense_blocks_bench(lq, args.seed, cpu=not args.no_cpu_baseline)

    cpu = None if args.no_cpu_baseline else cpu_baseline(n, layers, args.seed, args.cpu_seconds)

    ms_per_step = elapsed / args.steps * 1e3
    value = gates * args.steps / elapsed
    line = {
        "metric": "30q complex128 gates/s + % HBM peak; batched adjoint vs param-shift grads/s",
        "value": round(value, 2),
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic",
        "config": {"workload": f"random U3+CX layers, {n} qubits, {layers} layers, greedy 1q fusion",
                   "qubits": n, "layers": layers, "gates": gates, "plan": args.plan,
                   "tile_qubits": tile_q, "low_qubits": low_q, "tile_kernel": args.tile_kernel,
                   "first_apply_s": None if first_apply_s is None else round(first_apply_s, 3),
                   "passes_per_step": stats["passes"], "seed": args.seed},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "adjoint": adjoint,
        "trajectories": trajectories,
        "planner": planner,
        "qft": qft,
        "dense_blocks": dense,
        "density_matrix": dm,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

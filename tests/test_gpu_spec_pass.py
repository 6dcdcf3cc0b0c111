"""GPU parity of the specialized register-tile passes (tile_kernel =
"specialized"): kernels generated per pass structure and compiled at run time
(spec_pass.hip) against the oracle (reference
src/backend/statevector_simulator.cpp:197-265 on the fused circuit) and against
the interpreting register-tile kernel of the same plan."""
import numpy as np
import pytest

from litequantumcomputingsimulator_amd import workloads as W
from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu


def _spread(c, n, seed):
    rng = np.random.default_rng(seed)
    full = O.Circuit(n)
    for q in range(n):
        full.u(*(float(v) for v in rng.uniform(0, 6.28, 3)), q)
    full.insts += c.insts
    full.calls += c.calls
    return full


def _rotations(n, seed):
    rng = np.random.default_rng(seed)
    c = O.Circuit(n)
    for _ in range(60):
        w = int(rng.integers(1, 9))
        qs = [int(v) for v in rng.permutation(n)[:w]]
        p = "".join(str(rng.choice(list("XYZ"))) for _ in range(w))
        c.rp(float(rng.uniform(-3.14, 3.14)), p, qs)
    return c


def _diag_rich(n, seed):
    rng = np.random.default_rng(seed)
    c = O.Circuit(n)
    for _ in range(90):
        q = [int(v) for v in rng.permutation(n)[:3]]
        a = float(rng.uniform(0, 6.28))
        k = int(rng.integers(0, 11))
        if k == 0: c.rz(a, q[0])
        elif k == 1: c.p(a, q[0])
        elif k == 2: c.t(q[0]).s(q[1])
        elif k == 3: c.cz(q[0], q[1])
        elif k == 4: c.cp(a, q[0], q[1])
        elif k == 5: c.crz(a, q[0], q[1])
        elif k == 6: c.mcp(a, q[:2], q[2])
        elif k == 7: c.cy(q[0], q[1])
        elif k == 8: c.cx(q[0], q[1])
        elif k == 9: c.cry(a, q[0], q[1])
        else: c.u(a, float(rng.uniform(0, 6.28)), float(rng.uniform(0, 6.28)), q[0])
    return c


def _u3cx(n, layers, seed):
    c = O.Circuit(n)
    W.random_u3_cx_layers(c, n, layers, seed=seed)
    return c


CASES = [
    ("u3cx", 16, lambda: _u3cx(16, 8, 3), dict(plan="greedy")),
    ("u3cx_gain", 18, lambda: _u3cx(18, 8, 4), dict(plan="gain", low_qubits=3)),
    ("u3cx_lane_split", 15, lambda: _u3cx(15, 6, 16), dict(plan="greedy")),
    ("u3cx_13q_tiles", 17, lambda: _u3cx(17, 8, 5), dict(plan="greedy", tile_qubits=13, low_qubits=5)),
    ("random16kinds_13q_tiles", 15, lambda: _spread(cases.random_circuit(15, 200, 9), 15, 9),
     dict(plan="gain", tile_qubits=13)),
    ("random16kinds", 14, lambda: _spread(cases.random_circuit(14, 200, 4), 14, 5), dict(plan="greedy")),
    ("every_gate", 13, lambda: _spread(cases.every_gate_circuit(13), 13, 6), dict(plan="greedy")),
    ("rotations", 15, lambda: _spread(_rotations(15, 5), 15, 7), dict(plan="greedy")),
    ("diagonals", 14, lambda: _spread(_diag_rich(14, 6), 14, 8), dict(plan="greedy", low_qubits=2)),
]


@pytest.mark.parametrize("name,n,make,plan", CASES, ids=[c[0] for c in CASES])
def test_specialized_passes_match_oracle_and_interpreter(lq, name, n, make, plan):
    print(f"specialized passes: {name} n={n} {plan}", flush=True)
    circ = make()
    comp = lq.CompiledCircuit(circ.to_product(lq), True)
    want = O.run_statevector(circ, fuse=True)
    plan = dict(plan)
    L = plan.pop("tile_qubits", 12)
    spec = lq.statevector(n)
    comp.apply(spec, tile_qubits=L, tile_kernel="specialized", **plan)
    got = np.array(spec.to_numpy())
    assert np.max(np.abs(got - want)) <= 1e-12
    interp = lq.statevector(n)  # the interpreting kernel stops at 12-qubit tiles
    comp.apply(interp, tile_qubits=min(L, 12), tile_kernel="register", **plan)
    assert spec.max_abs_diff(interp) <= 1e-12


def test_specialized_passes_on_a_4gib_register(lq):
    """28 qubits: non-temporal loads and stores and tile numbers above 2^16,
    against the interpreting kernel (and the norm)."""
    n = 28
    print(f"specialized passes: u3cx n={n}", flush=True)
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_layers(qc, n, 6, seed=28)
    comp = lq.CompiledCircuit(qc, True)
    a = lq.statevector(n)
    comp.apply(a, plan="greedy", tile_kernel="specialized")
    b = lq.statevector(n)
    comp.apply(b, plan="greedy", tile_kernel="register")
    assert a.max_abs_diff(b) <= 1e-12
    assert abs(a.norm() - 1.0) <= 1e-10


def test_benchmark_circuit_at_full_size_agrees_across_kernels_and_plans(lq):
    """The 30-qubit benchmark circuit (20 U3/CX layers, seed 1234, 16 GiB per
    register) at full size, where the oracle does not reach: the bench's plan
    on specialized passes, the same plan on the interpreting kernel, and the
    learned plan on specialized passes give the same state within 1e-12, of
    norm 1."""
    n = 30
    print(f"specialized passes: benchmark circuit n={n}", flush=True)
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_layers(qc, n, 20, seed=1234)
    comp = lq.CompiledCircuit(qc, True)
    a = lq.statevector(n)
    comp.apply(a, plan="greedy", tile_qubits=12, low_qubits=4, tile_kernel="specialized")
    b = lq.statevector(n)
    comp.apply(b, plan="greedy", tile_qubits=12, low_qubits=4, tile_kernel="register")
    assert a.max_abs_diff(b) <= 1e-12
    b.set_basis(0)
    comp.apply(b, plan="learned", tile_kernel="specialized")
    assert a.max_abs_diff(b) <= 1e-12
    assert abs(a.norm() - 1.0) <= 1e-10


def test_13_qubit_tiles_on_a_4gib_register(lq):
    """28 qubits on 13-qubit tiles (512 threads, 64 KiB of LDS per tile) with
    five contiguous low qubits, against the interpreting kernel on 12-qubit
    tiles."""
    n = 28
    print(f"specialized passes: 13-qubit tiles n={n}", flush=True)
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_layers(qc, n, 6, seed=29)
    comp = lq.CompiledCircuit(qc, True)
    a = lq.statevector(n)
    comp.apply(a, plan="greedy", tile_qubits=13, low_qubits=5, tile_kernel="specialized")
    b = lq.statevector(n)
    comp.apply(b, plan="greedy", tile_kernel="register")
    assert a.max_abs_diff(b) <= 1e-12


def test_13_qubit_tiles_need_specialized_passes(lq):
    qc = lq.QuantumCircuit(16)
    W.random_u3_cx_layers(qc, 16, 2, seed=1)
    comp = lq.CompiledCircuit(qc, True)
    with pytest.raises(Exception, match="tile width"):
        comp.apply(lq.statevector(16), plan="greedy", tile_qubits=13, tile_kernel="register")


def test_specialized_passes_replay_in_a_captured_graph(lq):
    """Captured once, replayed three times with eager work in between; every
    replay equals an eager apply on the interpreting kernel."""
    n = 20
    print(f"specialized passes: graph n={n}", flush=True)
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_layers(qc, n, 8, seed=20)
    comp = lq.CompiledCircuit(qc, True)
    a, b = lq.statevector(n), lq.statevector(n)
    g = lq.CircuitGraph(comp, a, plan="greedy", tile_qubits=12, low_qubits=4, tile_kernel="specialized")
    for k in range(3):
        g.launch()
        comp.apply(b, plan="greedy", tile_qubits=12, low_qubits=4, tile_kernel="register")
        lq.synchronize()
        assert abs(b.norm() - 1.0) <= 1e-10
        assert a.max_abs_diff(b) <= 1e-12, k
    del g


def test_rebound_angles_reuse_the_compiled_kernels(lq):
    """Two circuits with the same structure and different angles: the second
    apply compiles nothing new (same kernel text), and both states equal the
    oracle's."""
    n = 16
    print(f"specialized passes: rebinding n={n}", flush=True)
    c1, c2 = _u3cx(n, 6, 40), _u3cx(n, 6, 40)
    # same gates and qubits, new angles
    rng = np.random.default_rng(41)
    for k, (name, args) in enumerate(c2.calls):
        if name == "u":
            new = tuple(float(v) for v in rng.uniform(0, 6.28, 3))
            c2.calls[k] = ("u", new + (args[3],))
            c2.insts[k]["params"] = list(new)
    comp1 = lq.CompiledCircuit(c1.to_product(lq), True)
    comp2 = lq.CompiledCircuit(c2.to_product(lq), True)
    s1, s2 = lq.statevector(n), lq.statevector(n)
    comp1.apply(s1, plan="greedy", tile_qubits=12, tile_kernel="specialized")
    before = lq.specialized_kernels_cached()
    comp2.apply(s2, plan="greedy", tile_qubits=12, tile_kernel="specialized")
    assert lq.specialized_kernels_cached() == before
    assert np.max(np.abs(np.array(s1.to_numpy()) - O.run_statevector(c1, fuse=True))) <= 1e-12
    assert np.max(np.abs(np.array(s2.to_numpy()) - O.run_statevector(c2, fuse=True))) <= 1e-12

"""Engine extensions around the hot path: the autotuner, captured launch
sequences (HIP graphs) and the per-call tile-kernel choice. Results must be
the same state as the plain greedy plan (device comparison, 1e-12)."""
import numpy as np
import pytest

from litequantumcomputingsimulator_amd import workloads as W


@pytest.fixture(scope="module")
def lq():
    from litequantumcomputingsimulator_amd import pylqcs

    return pylqcs


def _circuit(lq, n, depth=10, seed=3):
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_cp(qc, n, depth, seed)
    return qc


def test_tile_kernel_names_are_checked(lq):
    qc = _circuit(lq, 8, 2)
    comp = lq.CompiledCircuit(qc, True)
    with pytest.raises(ValueError):
        comp.describe(plan="warp")
    with pytest.raises(ValueError):
        lq.set_tile_kernel("warp")


@pytest.mark.parametrize("k", [1, 3, 5, 6])
def test_dense_plan_blocks_cover_every_op_within_k_qubits(lq, k):
    """Dense plans (planned on the host, no device): every block spans
    exactly k qubits (padded up), ops wider than k stand alone, the steps hold
    every op once, and a deep circuit over few qubits takes far fewer passes
    than ops."""
    n = 12
    qc = _circuit(lq, n, 8, 4)
    qc.mcx([0, 1, 2, 3, 4, 5, 6], 11)  # 8 qubits: wider than any block
    comp = lq.CompiledCircuit(qc, False)
    steps = comp.describe(plan="dense", tile_qubits=k)
    assert sum(s["ops"] for s in steps) == comp.num_ops()
    for s in steps:
        assert s["phases"] == 0
        assert s["tile_mask"] == 0 or bin(s["tile_mask"]).count("1") == k
    assert any(s["tile_mask"] == 0 for s in steps)  # the 8-qubit MCX
    deep = lq.QuantumCircuit(n)
    W.random_u3_cx_cp(deep, 4, 40, 2)  # 4 of the 12 qubits only
    dsteps = lq.CompiledCircuit(deep, False).describe(plan="dense", tile_qubits=k)
    if k >= 4:
        assert len(dsteps) == 1
    with pytest.raises(ValueError):
        comp.describe(plan="dense", tile_qubits=7)
    tiny = lq.QuantumCircuit(1)
    tiny.h(0).t(0).h(0)
    assert [s["ops"] for s in lq.CompiledCircuit(tiny, False).describe(plan="dense")] == [3]


@pytest.mark.gpu
def test_autotuner_picks_the_fastest_candidate_and_keeps_the_state(lq):
    n = 24
    comp = lq.CompiledCircuit(_circuit(lq, n, 12, 5), True)
    scratch = lq.statevector(n)
    r = comp.autotune(scratch, reps=2)
    # 6 register-tile shapes, 3 other kernels, 3 dense widths, 3 specialized-pass plans
    assert len(r["trials"]) == 15
    assert r["best_ms"] == min(ms for _, ms in r["trials"])
    assert comp.tuned() == r["best"]
    want, got = lq.statevector(n), lq.statevector(n)
    comp.apply(want, plan="greedy", tile_qubits=12, low_qubits=4)
    comp.apply(got)  # plan="auto": the tuned options
    assert got.max_abs_diff(want) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["register", "lds"])
def test_captured_graph_replays_match_eager_applies(lq, kernel):
    """A circuit's launches captured once and replayed three times, with the
    engine's own eager work and synchronisation in between; every replay
    agrees with the eager result."""
    n = 20
    comp = lq.CompiledCircuit(_circuit(lq, n, 8, 7), True)
    a, b = lq.statevector(n), lq.statevector(n)
    g = lq.CircuitGraph(comp, a, plan="greedy", tile_qubits=12, low_qubits=4, tile_kernel=kernel)
    for k in range(3):
        g.launch()
        comp.apply(b, plan="greedy", tile_qubits=12, low_qubits=4, tile_kernel=kernel)
        lq.synchronize()
        assert abs(b.norm() - 1.0) <= 1e-10  # eager reduction between replays
        assert a.max_abs_diff(b) <= 1e-12, k
    del g


@pytest.mark.gpu
@pytest.mark.parametrize("eng", [dict(plan="greedy", tile_qubits=12, low_qubits=4), dict(plan="dense", tile_qubits=5)],
                         ids=["greedy", "dense"])
def test_captured_graph_with_table_steps_matches_eager_applies(lq, eng):
    """Stand-alone steps that stage tables per call (a 3-qubit unitary, a
    4-controlled X, a 3-qubit permutation, a Toffoli; on the dense plan every
    block): captured, their tables live in memory the graph owns. Three
    replays agree with eager applies, and the first with the oracle."""
    from oracle import oracle as O

    n = 16
    rng = np.random.default_rng(12)
    c = O.Circuit(n)
    W.random_u3_cx_cp(c, n, 3, 12)
    m = np.linalg.qr(rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8)))[0]
    c.unitary(m, [0, 9, 15]).mcx([1, 2, 3, 4], 14).permutation([int(v) for v in rng.permutation(8)], [5, 11, 7])
    c.ccx(6, 13, 8)
    W.random_u3_cx_cp(c, n, 2, 13)
    want = O.run_statevector(c, fuse=True)
    comp = lq.CompiledCircuit(c.to_product(lq), True)
    a, b = lq.statevector(n), lq.statevector(n)
    g = lq.CircuitGraph(comp, a, **eng)
    for k in range(3):
        g.launch()
        comp.apply(b, **eng)
        lq.synchronize()
        if k == 0:
            assert np.max(np.abs(np.asarray(a.to_numpy()) - want)) <= 1e-12
        assert a.max_abs_diff(b) <= 1e-12, k
    del g


@pytest.mark.gpu
def test_host_view_then_two_applies_compose(lq):
    """A host view taken before device work is stale afterwards: a register
    exported with to_numpy() and then applied to twice holds U^2|0>, not U
    applied twice to the view's old contents; a view taken again shows the
    device result, and writes through a fresh view reach the device."""
    n = 14
    comp = lq.CompiledCircuit(_circuit(lq, n, 3, 21), True)
    a, b = lq.statevector(n), lq.statevector(n)
    view = a.to_numpy()
    comp.apply(a)
    comp.apply(a)
    comp.apply(b)
    comp.apply(b)
    assert a.max_abs_diff(b) <= 1e-12
    assert np.max(np.abs(np.asarray(a.to_numpy()) - np.asarray(b.to_numpy()))) <= 1e-12
    fresh = a.to_numpy()
    fresh[:] = 0
    fresh[3] = 1.0
    b.set_basis(3)
    assert a.max_abs_diff(b) == 0.0
    del view


@pytest.mark.gpu
def test_host_writes_reach_device_reductions(lq):
    """Writes through a host view reach the device before its next use: a
    state written on the host gives its exact norm and Pauli expectations;
    after an apply the view is refreshed, not replayed."""
    n = 10
    a = lq.statevector(n)
    v = a.to_numpy()
    v[:] = 0
    v[0] = v[3] = 1 / np.sqrt(2)  # (|0> + |3>)/sqrt2 on qubits 0 and 1: <XX> = 1, <ZZ> = 1, <Z0> = 0
    assert abs(a.norm() - 1.0) <= 1e-14
    assert abs(a.expectation_value("I" * (n - 2) + "XX") - 1.0) <= 1e-14
    assert abs(a.expectation_value("I" * (n - 2) + "ZZ") - 1.0) <= 1e-14
    assert abs(a.expectation_value("I" * (n - 1) + "Z")) <= 1e-14
    qc = lq.QuantumCircuit(n)
    qc.x(0)
    lq.CompiledCircuit(qc, False).apply(a)  # X on qubit 0: (|1> + |2>)/sqrt2
    w = np.asarray(a.to_numpy())
    assert abs(w[1] - 1 / np.sqrt(2)) <= 1e-14 and abs(w[2] - 1 / np.sqrt(2)) <= 1e-14
    assert abs(w[0]) == 0.0 and abs(w[3]) == 0.0


@pytest.mark.gpu
def test_per_call_tile_kernels_agree(lq):
    n = 22
    comp = lq.CompiledCircuit(_circuit(lq, n, 10, 9), True)
    ref = lq.statevector(n)
    comp.apply(ref, plan="greedy", tile_qubits=12, low_qubits=4, tile_kernel="register")
    for kernel in ("lds", "stream", "land"):
        sv = lq.statevector(n)
        comp.apply(sv, plan="greedy", tile_qubits=12, low_qubits=4, tile_kernel=kernel)
        assert sv.max_abs_diff(ref) <= 1e-12, kernel


@pytest.mark.gpu
def test_register_tile_on_a_1gib_register_matches_gate_by_gate(lq):
    """At 26 qubits (1 GiB) the register tile streams the state with
    non-temporal loads and stores; its state equals the gate-by-gate kernels'
    (a separate set of kernels) within 1e-12, with U3/CX layers whose CX merge
    into select ops."""
    n = 26
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_layers(qc, n, 4, seed=26)
    comp = lq.CompiledCircuit(qc, True)
    a = lq.statevector(n)
    comp.apply(a, plan="greedy", tile_qubits=12, low_qubits=4, tile_kernel="register")
    b = lq.statevector(n)
    comp.apply(b, plan="gate_by_gate")
    assert a.max_abs_diff(b) <= 1e-12
    assert abs(a.norm() - 1.0) <= 1e-10


def _diag_rich_circuit(lq, n, seed):
    rng = np.random.default_rng(seed)
    qc = lq.QuantumCircuit(n)
    for _ in range(120):
        k = int(rng.integers(0, 11))
        a, b, c = (int(v) for v in rng.choice(n, 3, replace=False))
        th = float(rng.uniform(0, 6.28))
        if k == 0:
            qc.rz(th, a)
        elif k == 1:
            qc.p(th, a)
        elif k == 2:
            qc.t(a).s(b)
        elif k == 3:
            qc.cz(a, b)
        elif k == 4:
            qc.cp(th, a, b)
        elif k == 5:
            qc.crz(th, a, b)
        elif k == 6:
            qc.mcp(th, [a, b], c)
        elif k == 7:
            qc.cx(a, b)
        elif k == 8:
            qc.cy(a, b)
        else:
            qc.u(*(float(v) for v in rng.uniform(0, 6.28, 3)), a)
    return qc


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_diagonal_and_select_forms_match_gate_by_gate(lq, seed):
    """Circuits dense in single-qubit diagonals (RZ, P, T, S), controlled
    diagonals (CZ, CP, CRZ, MCP) and CX/CY next to U3 gates, at 16 qubits
    (tiles of 12): the register tile's diagonal, select and controlled forms,
    on register, thread and tile-number bits, with and without controls,
    give the gate-by-gate kernels' state within 1e-12, with and without the
    host's single-qubit fusion."""
    n = 16
    qc = _diag_rich_circuit(lq, n, seed)
    for fuse in (False, True):
        comp = lq.CompiledCircuit(qc, fuse)
        a = lq.statevector(n)
        comp.apply(a, plan="greedy", tile_qubits=12, low_qubits=4, tile_kernel="register")
        b = lq.statevector(n)
        comp.apply(b, plan="gate_by_gate")
        assert a.max_abs_diff(b) <= 1e-12, fuse
        assert abs(a.norm() - 1.0) <= 1e-10

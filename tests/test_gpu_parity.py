"""Parity of the HIP engine (through pylqcs -> C++20 host -> C-ABI) with the
CPU oracle on the same seeded inputs. Amplitudes: max-abs <= 1e-12.
Sampled indices: identical except draws whose scaled value lies within a
rounding margin of a cumulative boundary ("ambiguous", expected none)."""
import math

import numpy as np
import pytest

from litequantumcomputingsimulator_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def random_circuit(n, n_gates, seed):
    """16 gate kinds as in the reference's tests/test_equivalence.cpp:47-93
    (kinds and arities; the draw sequence is numpy's)."""
    rng = np.random.default_rng(seed)
    c = O.Circuit(n)

    def pick(k):
        return [int(q) for q in rng.permutation(n)[:k]]

    for _ in range(n_gates):
        k = int(rng.integers(0, 16))
        a = float(rng.uniform(0, 6.28))
        if k == 0: c.h(pick(1)[0])
        elif k == 1: c.x(pick(1)[0])
        elif k == 2: c.t(pick(1)[0])
        elif k == 3: c.sx(pick(1)[0])
        elif k == 4: c.rx(a, pick(1)[0])
        elif k == 5: c.u(a, float(rng.uniform(0, 6.28)), float(rng.uniform(0, 6.28)), pick(1)[0])
        elif k == 6: c.cx(*pick(2))
        elif k == 7: c.cp(a, *pick(2))
        elif k == 8: c.swap(*pick(2))
        elif k == 9: c.iswap(*pick(2))
        elif k == 10: c.rxx(a, *pick(2))
        elif k == 11: c.rzz(a, *pick(2))
        elif k == 12: c.ccx(*pick(3))
        elif k == 13: c.cswap(*pick(3))
        elif k == 14:
            q = pick(3)
            c.mcp(a, q[:2], q[2])
        else:
            c.permutation([int(v) for v in rng.permutation(4)], pick(2))
    return c


def product_state(lq, circ, fuse=True, **eng):
    sim = lq.StatevectorSimulator(fuse_gates=fuse, **eng)
    return np.array(sim.run_statevector(circ.to_product(lq)).to_numpy())


PLANS = [dict(plan="gate_by_gate"), dict(plan="greedy"), dict(plan="greedy", tile_qubits=6, low_qubits=2)]
# Dense blocks of 1 to 6 qubits (up to 3 on the generic kernel, 4..6 on FP64
# MFMA where the register has enough groups).
DENSE_PLANS = [dict(plan="dense", tile_qubits=k) for k in (1, 2, 3, 4, 5, 6)]


@pytest.mark.parametrize("seed", [11, 22, 33])
@pytest.mark.parametrize("eng", PLANS + DENSE_PLANS)
def test_random_circuits_match_oracle(lq, seed, eng):
    n = 7
    c = random_circuit(n, 60, seed)
    want = O.run_statevector(c, fuse=True)
    got = product_state(lq, c, True, **eng)
    assert np.max(np.abs(got - want)) <= TOL


@pytest.mark.parametrize("n", [7, 9, 12, 14])
def test_extra_gates_match_oracle(lq, n):
    rng = np.random.default_rng(n)
    c = O.Circuit(n)
    for q in range(n):
        c.h(q)
    c.rp(0.7, "XYZ", [0, n - 1, 2]).rp(-1.1, "ZZ", [1, 3]).rp(0.4, "YIX", [n - 2, 0, 4])
    c.ch(0, n - 1).cry(0.3, n - 1, 1).crz(0.9, 2, 3).crx(0.2, 1, 0).cy(3, 4)
    c.ryy(0.5, 0, n - 1).rzz(0.25, 1, 2).rxx(0.75, n - 1, 3)
    c.mcx([0, 1, 2], n - 1).mcp(0.6, [n - 1, 1], 0).s(1).sdg(2).tdg(3).y(4).z(5).i(0).p(0.3, 2).ry(0.8, 1)
    m = np.linalg.qr(rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8)))[0]
    c.unitary(m, [0, n - 1, 2])
    m2 = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]
    c.controlled_unitary(m2, [n - 1], [0, 1])
    c.controlled_permutation([1, 2, 3, 0], [n - 1], [2, 3])
    c.permutation([int(v) for v in rng.permutation(8)], [1, n - 1, 3])
    want = O.run_statevector(c, fuse=False)
    for eng in PLANS[:2] + DENSE_PLANS:
        got = product_state(lq, c, False, **eng)
        assert np.max(np.abs(got - want)) <= TOL, eng


def test_qft20_from_basis_matches_oracle_and_dft(lq):
    n = 20
    c = O.Circuit(n)
    x = W.qft_from_basis(c, n, seed=7)
    want = O.run_statevector(c, fuse=False)
    got = product_state(lq, c, False)
    assert np.max(np.abs(got - want)) <= TOL
    k = np.arange(1 << n)
    dft = np.exp(2j * np.pi * ((x * k) % (1 << n)) / (1 << n)) / math.sqrt(1 << n)
    assert np.max(np.abs(got - dft)) <= TOL


@pytest.mark.parametrize("n,layers", [(16, 3), (20, 4)])
def test_random_layers_greedy_vs_oracle(lq, n, layers):
    c = O.Circuit(n)
    W.random_u3_cx_layers(c, n, layers, seed=n)
    want = O.run_statevector(c, fuse=True)
    for eng in (dict(plan="greedy"), dict(plan="greedy", tile_qubits=10, low_qubits=3),
                dict(plan="gate_by_gate")):
        got = product_state(lq, c, True, **eng)
        assert np.max(np.abs(got - want)) <= TOL, eng


@pytest.mark.parametrize("k", [4, 5, 6])
def test_dense_blocks_on_mfma_match_oracle(lq, k):
    """Dense blocks of 4..6 qubits at 18 qubits, where they run on the FP64
    MFMA kernel: random U3/CX/CP layers plus Pauli rotations (an X part and
    Z-only), a Toffoli and a swap, against the oracle; each block holds at
    most k qubits and the blocks hold every op once."""
    n = 18
    c = O.Circuit(n)
    W.random_u3_cx_cp(c, n, 6, seed=k)
    c.rp(0.3, "XZY", [0, 9, 16]).rp(0.9, "ZIZ", [2, 11, 4]).ccx(3, 15, 7).swap(1, 14).ryy(0.4, 5, 12)
    c.rp(-0.7, "YY", [17, 16]).rp(1.3, "Z", [8])
    want = O.run_statevector(c, fuse=False)
    got = product_state(lq, c, False, plan="dense", tile_qubits=k)
    assert np.max(np.abs(got - want)) <= TOL


WAVE_PLANS = [dict(plan="greedy", tile_qubits=10, low_qubits=3), dict(plan="gain", tile_qubits=10, low_qubits=2),
              dict(plan="gain", tile_qubits=10, low_qubits=4), dict(plan="greedy", tile_qubits=10, low_qubits=6)]


@pytest.mark.parametrize("n,seed", [(13, 11), (16, 22)])
@pytest.mark.parametrize("eng", WAVE_PLANS)
def test_wave_tiles_match_oracle(lq, n, seed, eng):
    """Wave tiles (10-qubit tiles, one wave each, wave-private transposes):
    every gate kind, Pauli rotations with Z parts on tile and outside
    qubits, multi-controlled and diagonal gates, against the oracle."""
    c = random_circuit(n, 160, seed)
    c.rp(0.3, "XZY", [0, n // 2, n - 1]).rp(0.9, "ZIZ", [2, n - 3, 4]).rp(-0.5, "YY", [1, n - 2])
    c.crz(0.4, n - 1, 3).mcp(0.7, [0, n - 2], 5).cy(2, n - 4).cz(n - 1, 0).ccx(1, n - 1, 6)
    want = O.run_statevector(c, fuse=True)
    got = product_state(lq, c, True, **eng)
    assert np.max(np.abs(got - want)) <= TOL


def test_wave_tiles_noisy_trajectories_match_oracle_counts(lq):
    """Batched trajectories on wave tiles (the batch member is the
    workgroup's second grid index): counts equal the oracle's."""
    from tests import cases

    c, chans = cases.noisy_ghz(12)
    want = O.run_counts(c, seed=13, shots=256, channels=chans)
    got = lq.StatevectorSimulator(seed=13, noise=_noise_model(lq), plan="greedy", tile_qubits=10,
                                  low_qubits=3, trajectory_kernel="shared").run(c.to_product(lq), 256).counts()
    assert dict(got) == want


def _ambiguous(cum, draws, margin):
    total = cum[-1]
    r = draws * total
    j = np.searchsorted(cum, r, side="right")
    lo = np.where(j > 0, cum[np.maximum(j - 1, 0)], 0.0)
    hi = np.where(j < len(cum), cum[np.minimum(j, len(cum) - 1)], np.inf)
    return (np.abs(r - lo) < margin) | (np.abs(hi - r) < margin)


def test_qft20_sampling_indices_match_oracle(lq):
    n, shots, seed = 20, 1024, 2024
    c = O.Circuit(n)
    W.qft_from_basis(c, n, seed=3)
    c.measure_all()
    want_idx, cum = O.run_sampling(c, seed, shots, fuse=False)
    sim = lq.StatevectorSimulator(seed=seed, fuse_gates=False)
    got_idx = np.array(sim.sample_indices(c.to_product(lq), shots))
    draws = O.uniform_stream(seed, shots)
    amb = _ambiguous(cum, draws, 1e-12)
    assert int(amb.sum()) == 0
    assert np.array_equal(got_idx[~amb], want_idx[~amb])


@pytest.mark.parametrize("fuse", [True, False])
def test_run_counts_match_oracle(lq, fuse):
    c = O.Circuit(3, 3)
    c.h(0).t(0).cx(0, 1).rx(0.4, 2).sx(2).cx(1, 2).measure_all()
    want = O.run_counts(c, seed=5, shots=512, fuse=fuse)
    got = lq.StatevectorSimulator(seed=5, fuse_gates=fuse).run(c.to_product(lq), 512).counts()
    assert dict(got) == want


@pytest.mark.parametrize("kernel", ["shared", "persistent"])
def test_midcircuit_counts_match_oracle(lq, kernel):
    c = O.Circuit(2, 2)
    c.h(0).t(0).measure(0, 0).h(1).cx(0, 1).reset(0).h(0).measure(1, 1)
    want = O.run_counts(c, seed=9, shots=256)
    got = lq.StatevectorSimulator(seed=9, trajectory_kernel=kernel).run(c.to_product(lq), 256).counts()
    assert dict(got) == want


@pytest.mark.parametrize("kernel", ["shared", "persistent"])
def test_noisy_counts_match_oracle(lq, kernel):
    c = O.Circuit(3, 3)
    c.h(0).cx(0, 1).t(1).cx(1, 2).u(0.3, 0.2, 0.1, 2).measure_all()
    chans = [O.depolarizing(0.1), O.amplitude_damping(0.2)]
    want = O.run_counts(c, seed=99, shots=400, channels=chans)
    noise = lq.NoiseModel()
    noise.add_all_qubit_channel(lq.KrausChannel.depolarizing(0.1))
    noise.add_all_qubit_channel(lq.KrausChannel.amplitude_damping(0.2))
    got = lq.StatevectorSimulator(seed=99, noise=noise, trajectory_kernel=kernel).run(c.to_product(lq), 400).counts()
    assert dict(got) == want


def test_expectations_match_oracle(lq):
    n = 10
    c = random_circuit(n, 50, 5)
    st = O.run_statevector(c)
    sv = lq.StatevectorSimulator().run_statevector(c.to_product(lq))
    h = W.random_hamiltonian(n, 30, seed=1)
    for coeff, p in h[:10]:
        assert abs(sv.expectation_value(p) - O.expectation_value(st, n, p)) <= TOL
    assert abs(lq.algorithms.expectation(sv, h) - O.expectation(st, n, h)) <= 1e-11
    assert abs(sv.norm() - 1.0) <= 1e-12


def _ansatz(lq, n, n_rot, seed):
    rots, theta = W.random_pauli_ansatz(n, n_rot, seed)
    pc = lq.ParametricCircuit(n)
    t = O.Circuit(n)
    for q in range(n):
        pc.h(q)
        t.h(q)
    param_inst = []
    for p, qs in rots:
        pc.rp(p, qs)
        param_inst.append(len(t.insts))
        t.rp(0.0, p, qs)
    pc.cx(0, 1)
    t.cx(0, 1)
    pc.ry(2)
    param_inst.append(len(t.insts))
    t.ry(0.0, 2)
    theta = list(theta) + [0.37]
    return pc, t, param_inst, theta


def test_gradients_match_oracle(lq):
    n = 8
    pc, t, pinst, theta = _ansatz(lq, n, 12, seed=4)
    h = W.random_hamiltonian(n, 15, seed=2)
    want = O.parameter_shift_gradient(t, pinst, theta, h)
    ps = np.array(lq.algorithms.parameter_shift_gradient(h, pc, theta))
    e, adj = lq.algorithms.adjoint_gradient(h, pc, theta)
    assert np.max(np.abs(ps - want)) <= 1e-10
    assert np.max(np.abs(np.array(adj) - want)) <= 1e-10
    assert abs(e - O.bound_energy(t, pinst, theta, h)) <= 1e-10


def test_batched_adjoint_matches_single(lq):
    jobs = []
    for k, n in enumerate([5, 7, 6, 9]):
        pc, t, pinst, theta = _ansatz(lq, n, 10, seed=10 + k)
        h = W.random_hamiltonian(n, 8, seed=20 + k)
        jobs.append((pc, t, pinst, theta, h))
    out = lq.algorithms.adjoint_gradient_batched([j[4] for j in jobs], [j[0] for j in jobs],
                                                 [j[3] for j in jobs])
    for (pc, t, pinst, theta, h), (e, g) in zip(jobs, out):
        want = O.parameter_shift_gradient(t, pinst, theta, h)
        assert np.max(np.abs(np.array(g) - want)) <= 1e-10
        assert abs(e - O.bound_energy(t, pinst, theta, h)) <= 1e-10


def _pauli_only(lq, n, seed):
    """Pauli rotations only (no other gates), so that a batch of them shares
    launches, with the cases that shape fused rotation groups: Z-only
    strings (no X part), one rotation repeated, and three rotations whose
    third X mask is the XOR of the first two."""
    rots, theta = W.random_pauli_ansatz(n, 10, seed)
    rots = list(rots)
    theta = [float(x) for x in theta]
    special = [("ZZ", [0, 1]), ("XY", [0, 2]), ("XY", [0, 2]), ("X", [0]), ("X", [1]), ("XX", [0, 1]),
               ("Z", [n - 1]), ("YZX", [n - 1, 0, 1])]
    rng = np.random.default_rng(seed + 7)
    at = 3
    for p, qs in special:
        rots.insert(at, (p, qs))
        theta.insert(at, float(rng.uniform(-math.pi, math.pi)))
        at += 1
    pc = lq.ParametricCircuit(n)
    t = O.Circuit(n)
    pinst = []
    for p, qs in rots:
        pc.rp(p, qs)
        pinst.append(len(t.insts))
        t.rp(0.0, p, qs)
    return pc, t, pinst, theta


@pytest.mark.parametrize("widths", [(3, 4, 5, 8), (2, 5), (7,)])
def test_batched_adjoint_pauli_only_matches_oracle(lq, widths):
    jobs = []
    for k, n in enumerate(widths):
        if n >= 3:
            pc, t, pinst, theta = _pauli_only(lq, n, seed=40 + k)
        else:  # two qubits: rotations on both
            pc, t, pinst = lq.ParametricCircuit(n), O.Circuit(n), []
            theta = [0.3, -1.1, 0.7, 2.0, -0.4]
            for p, qs in [("XY", [0, 1]), ("Z", [1]), ("YY", [0, 1]), ("ZX", [0, 1]), ("X", [0])]:
                pc.rp(p, qs)
                pinst.append(len(t.insts))
                t.rp(0.0, p, qs)
        h = W.random_hamiltonian(n, 9, seed=60 + k)
        jobs.append((pc, t, pinst, theta, h))
    out = lq.algorithms.adjoint_gradient_batched([j[4] for j in jobs], [j[0] for j in jobs],
                                                 [j[3] for j in jobs])
    for (pc, t, pinst, theta, h), (e, g) in zip(jobs, out):
        want = O.parameter_shift_gradient(t, pinst, theta, h)
        assert np.max(np.abs(np.array(g) - want)) <= 1e-10
        assert abs(e - O.bound_energy(t, pinst, theta, h)) <= 1e-10


def test_adjoint_on_28_qubit_registers_matches_parameter_shift(lq):
    """Rotation groups on registers past 2^27 amplitudes (byte offsets past
    4 GiB), where the oracle is too slow: the adjoint against the engine's
    own parameter shift, which shares no kernel with it."""
    n = 28

    def pauli(letters):
        s = ["I"] * n
        for q, p in letters.items():
            s[q] = p
        return "".join(s)

    rots = [("XY", [0, 27]), ("ZX", [3, 26]), ("YYZ", [27, 1, 14]), ("Z", [27]), ("XZ", [25, 27]),
            ("YX", [0, 26]), ("XXY", [2, 13, 27])]
    theta = [0.3, -1.2, 0.8, 2.1, -0.5, 1.4, -2.6]
    h = [(0.7, pauli({27: "X", 0: "Y"})), (-1.1, pauli({26: "Z", 3: "Z"})), (0.4, pauli({14: "X", 27: "Z"})),
         (0.9, pauli({1: "Y", 25: "Y", 2: "X"}))]
    pc = lq.ParametricCircuit(n)
    for p, qs in rots:
        pc.rp(p, qs)
    ps = np.array(lq.algorithms.parameter_shift_gradient(h, pc, theta))
    e, adj = lq.algorithms.adjoint_gradient(h, pc, theta)
    assert np.max(np.abs(np.array(adj) - ps)) <= 1e-10
    assert np.max(np.abs(ps)) > 1e-3  # the check is not vacuous


def test_torch_expectation_backward_matches_oracle(lq):
    torch = pytest.importorskip("torch")
    n = 6
    pc, t, pinst, theta = _ansatz(lq, n, 10, seed=31)
    h = W.random_hamiltonian(n, 10, seed=32)
    want = O.parameter_shift_gradient(t, pinst, theta, h)
    e_want = O.bound_energy(t, pinst, theta, h)
    # theta stays on the host: the torch wheel carries its own HIP runtime,
    # and whichever of it and the engine's runtime opens the device first
    # holds it for the process, so this test never initialises torch's.
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    e = lq.expectation(th, pc, h)
    (2.0 * e).backward()
    assert np.max(np.abs(th.grad.numpy() - 2.0 * want)) <= 1e-10
    assert abs(e.item() - e_want) <= 1e-10
    # a float32 theta gets the float32-rounded gradient at its own values
    th32 = torch.tensor(theta, dtype=torch.float32, requires_grad=True)
    lq.expectation(th32, pc, h).backward()
    assert th32.grad.dtype == torch.float32
    want32 = O.parameter_shift_gradient(t, pinst, [float(x) for x in th32.detach().numpy()], h)
    assert np.max(np.abs(th32.grad.numpy().astype(np.float64) - want32)) <= 1e-5


def test_torch_expectation_forward_only_without_grad(lq):
    """Under torch.no_grad(), or for a theta that needs no gradient, the
    bridge evolves the state once (no adjoint sweep) and returns the same
    energy, with nothing to differentiate; a theta off the CPU is refused
    before any device work (a meta tensor stands for one here, so torch never
    opens the GPU in this process)."""
    torch = pytest.importorskip("torch")
    n = 6
    pc, t, pinst, theta = _ansatz(lq, n, 10, seed=33)
    h = W.random_hamiltonian(n, 10, seed=34)
    e_want = O.bound_energy(t, pinst, theta, h)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    with torch.no_grad():
        e = lq.expectation(th, pc, h)
    assert not e.requires_grad and abs(e.item() - e_want) <= 1e-10
    e2 = lq.expectation(torch.tensor(theta, dtype=torch.float64), pc, h)
    assert not e2.requires_grad and abs(e2.item() - e_want) <= 1e-10
    with pytest.raises(ValueError, match="CPU tensor"):
        lq.expectation(torch.empty(len(theta), dtype=torch.float64, device="meta"), pc, h)


def test_errors_keep_reference_types(lq):
    with pytest.raises(ValueError):
        lq.Statevector(31)
    with pytest.raises(ValueError):
        lq.Statevector(0)
    qc = lq.QuantumCircuit(2)
    qc.h(0)
    with pytest.raises(ValueError):
        lq.StatevectorSimulator().run(qc, 10)  # no measurement
    qm = lq.QuantumCircuit(1, 1)
    qm.measure(0, 0)
    with pytest.raises(ValueError):
        lq.StatevectorSimulator().run(qm, 0)
    sv = lq.Statevector(2)
    with pytest.raises(ValueError):
        sv.expectation_value("Z")
    with pytest.raises(IndexError):
        _ = sv[4]


@pytest.mark.parametrize("kernel", ["register", "stream", "land", "lds", "specialized"])
@pytest.mark.parametrize("plan,low", [("greedy", 4), ("learned", 3), ("learned", 2), ("greedy", 5)])
def test_tile_kernels_and_planners_match_oracle(lq, kernel, plan, low):
    n = 17
    c = O.Circuit(n)
    W.random_u3_cx_cp(c, n, 5, seed=low)
    c.rp(0.3, "XZY", [0, 9, 16]).rp(0.9, "ZIZ", [2, 11, 4]).ccx(3, 15, 7).swap(1, 14).ryy(0.4, 5, 12)
    want = O.run_statevector(c, fuse=True)
    lq.set_tile_kernel(kernel)
    try:
        got = product_state(lq, c, True, plan=plan, tile_qubits=12, low_qubits=low)
    finally:
        lq.set_tile_kernel("register")
    assert np.max(np.abs(got - want)) <= TOL


def _noise_model(lq, p=1e-3, gamma=1e-3):
    noise = lq.NoiseModel()
    noise.add_all_qubit_channel(lq.KrausChannel.depolarizing(p))
    noise.add_all_qubit_channel(lq.KrausChannel.amplitude_damping(gamma))
    return noise


@pytest.mark.parametrize("kernel", ["shared", "persistent"])
def test_noisy_trajectories_12q_match_oracle_counts(lq, kernel):
    """The noisy-trajectory workload at 12 qubits: device trajectories (shared
    tiled runs over a batch, or one workgroup per shot with the register in
    LDS) reproduce the oracle's per-shot counts exactly (same draws)."""
    from tests import cases

    c, chans = cases.noisy_ghz(12)
    want = O.run_counts(c, seed=11, shots=256, channels=chans)
    sim = lq.StatevectorSimulator(seed=11, noise=_noise_model(lq), trajectory_kernel=kernel)
    assert dict(sim.run(c.to_product(lq), 256).counts()) == want


def test_noisy_trajectories_on_specialized_passes_match_oracle_counts(lq):
    """Shared trajectory runs with every tiled pass on specialized kernels:
    batches of registers (grid rows), state-independent branches as
    single-register ops and default branches on all but one register (batch
    restrictions in the generated code); counts equal the oracle's."""
    from tests import cases

    c, chans = cases.noisy_ghz(12, p=0.02, gamma=0.05)
    want = O.run_counts(c, seed=13, shots=128, channels=chans)
    noise = _noise_model(lq, 0.02, 0.05)
    sim = lq.StatevectorSimulator(seed=13, noise=noise, trajectory_kernel="shared")
    lq.set_tile_kernel("specialized")
    try:
        got = dict(sim.run(c.to_product(lq), 128).counts())
    finally:
        lq.set_tile_kernel("register")
    assert got == want


@pytest.mark.parametrize("kernel", ["shared", "persistent"])
@pytest.mark.parametrize("fixture,n,rate", [("dm_noisy_ghz12.npy", 12, 1e-3), ("dm_noisy_ghz10_strong.npy", 10, 0.05)])
def test_noisy_trajectories_match_density_matrix(lq, fixture, n, rate, kernel):
    """2048 device trajectories of the noisy workload against the exact
    distribution from the reference's density-matrix simulator
    (tests/golden/make_dm_fixture.py): every single-qubit marginal and the
    Hamming-weight histogram within 5 sigma, and a chi-square over the
    bitstrings (pooled, in order of probability, into groups expecting at
    least 5 counts) within 5 sigma of its degrees of freedom. At rate 0.05 the
    check fails for trajectories without noise or without the damping
    channel (tried on the oracle); at 1e-3, 2048 shots cannot tell noise
    from none, so the 12-qubit case is a consistency check."""
    import os

    from tests import cases

    shots = 2048
    p = np.load(os.path.join(os.path.dirname(__file__), "golden", fixture))
    c, _ = cases.noisy_ghz(n, p=rate, gamma=rate)
    sim = lq.StatevectorSimulator(seed=2024, noise=_noise_model(lq, rate, rate), trajectory_kernel=kernel)
    counts = sim.run(c.to_product(lq), shots).counts()
    freq = np.zeros(1 << n)
    for key, v in dict(counts).items():
        freq[int(key, 2)] += v
    assert freq.sum() == shots
    idx = np.arange(1 << n)

    def within(obs, prob):
        sigma = np.sqrt(shots * prob * (1 - prob))
        return np.all(np.abs(obs - shots * prob) <= 5 * sigma + 1)

    for q in range(n):
        bit = (idx >> q) & 1 == 1
        assert within(freq[bit].sum(), p[bit].sum()), q
    weight = np.array([bin(i).count("1") for i in idx])
    for w in range(n + 1):
        assert within(freq[weight == w].sum(), p[weight == w].sum()), w
    order = np.argsort(p)
    chi2, groups, e_acc, o_acc = 0.0, 0, 0.0, 0.0
    for i in order:
        e_acc += shots * p[i]
        o_acc += freq[i]
        if e_acc >= 5.0:
            chi2 += (o_acc - e_acc) ** 2 / e_acc
            groups += 1
            e_acc = o_acc = 0.0
    dof = groups - 1
    assert chi2 <= dof + 5 * np.sqrt(2 * dof), (chi2, dof)


@pytest.mark.parametrize("kernel", ["stream", "land"])
def test_walking_tile_kernels_match_register_kernel(lq, kernel):
    """The tile-walking kernels (next tile prefetched into registers, or into
    an LDS landing buffer by DMA) at a width with more tiles than workgroups
    (22 qubits: 1024 tiles), on a circuit whose passes need several op-word
    segments: same state as the one-tile-per-workgroup kernel, which the
    oracle checks above."""
    n = 22
    qc = lq.QuantumCircuit(n)
    rng = np.random.default_rng(21)
    for i in range(200):  # one pass of 400 ops on qubits 0..7: op words span several segments
        qc.u(*(float(a) for a in rng.uniform(0, 6.28, 3)), i % 8)
        qc.cx(i % 8, (i + 3) % 8)
    W.random_u3_cx_cp(qc, n, 40, seed=21)
    for q in range(0, n - 2, 3):
        qc.ccx(q, q + 1, q + 2)
    comp = lq.CompiledCircuit(qc, True)
    want = lq.statevector(n)
    comp.apply(want, plan="greedy", tile_qubits=12, low_qubits=4)
    lq.set_tile_kernel(kernel)
    try:
        got = lq.statevector(n)
        comp.apply(got, plan="greedy", tile_qubits=12, low_qubits=4)
        got2 = lq.statevector(n)
        comp.apply(got2, plan="gain", tile_qubits=12, low_qubits=2)
    finally:
        lq.set_tile_kernel("register")
    ref2 = lq.statevector(n)
    comp.apply(ref2, plan="gain", tile_qubits=12, low_qubits=2)
    assert got.max_abs_diff(want) <= 1e-12
    assert got2.max_abs_diff(ref2) <= 1e-12
    assert abs(got.norm() - 1.0) <= 1e-10

"""Dense controlled unitaries with 4..6 targets on FP64 MFMA (the reference's
generic path, src/backend/kernels/gate_kernels.cpp:143-167, which fused k<=6
blocks and multi-target gates take).

Parity: the MFMA kernel against the CPU oracle (max-abs <= 1e-12) over
target orders that put the low, contiguous qubits inside and outside the
block, 0-2 controls, and register widths at and just above the smallest one
the kernel takes (16 column groups per wave iteration); against the vector
kernels on the device at 24 and 29 qubits; and on a trajectory batch, where
one launch covers every shot. Also: permutations on registers wide enough
that one workgroup per group exceeds a single launch."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def haar(k, rng):
    d = 1 << k
    q, r = np.linalg.qr(rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


def scrambled(n, seed):
    """A dense random state: U3 on every qubit, a CX ladder, U3 again."""
    rng = np.random.default_rng(seed)
    c = O.Circuit(n)
    for layer in range(2):
        for q in range(n):
            c.u(*(float(v) for v in rng.uniform(0, 2 * np.pi, 3)), q)
        if layer == 0:
            for q in range(n - 1):
                c.cx(q, q + 1)
    return c


def run_product(lq, c):
    sim = lq.StatevectorSimulator(fuse_gates=False, plan="gate_by_gate")
    return np.array(sim.run_statevector(c.to_product(lq)).to_numpy())


TARGET_ORDERS = {
    "low": lambda n, k: list(range(k)),                       # the block is the contiguous low qubits
    "high": lambda n, k: list(range(n - 1, n - 1 - k, -1)),    # top qubits, reversed local order
    "mixed": lambda n, k: [n - 1, 0, n // 2, 1, n - 3, 2][:k],  # interleaved with the column bits
}


@pytest.mark.parametrize("k", [4, 5, 6])
@pytest.mark.parametrize("order", sorted(TARGET_ORDERS))
@pytest.mark.parametrize("nc", [0, 1, 2])
def test_dense_block_matches_oracle(lq, k, order, nc):
    n = 13
    rng = np.random.default_rng(100 * k + 10 * nc + len(order))
    tg = TARGET_ORDERS[order](n, k)
    free = [q for q in range(n) if q not in tg]
    ctl = [int(q) for q in rng.permutation(free)[:nc]]
    c = scrambled(n, seed=k + nc)
    u = haar(k, rng)
    if nc:
        c.controlled_unitary(u, ctl, tg)
    else:
        c.unitary(u, tg)
    c.h(0).cx(n - 1, 1)
    want = O.run_statevector(c, fuse=False)
    got = run_product(lq, c)
    assert np.max(np.abs(got - want)) <= TOL


@pytest.mark.parametrize("k,extra", [(4, 5), (4, 6), (5, 4), (5, 5), (6, 3), (6, 4)])
def test_dense_block_at_smallest_widths(lq, k, extra):
    """n = k + extra: below 4 + log2(slabs per iteration) free qubits the
    vector kernel runs; at and above it, MFMA. Both agree with the oracle."""
    n = k + extra
    rng = np.random.default_rng(7 * k + extra)
    c = scrambled(n, seed=n)
    c.unitary(haar(k, rng), [int(q) for q in rng.permutation(n)[:k]])
    want = O.run_statevector(c, fuse=False)
    assert np.max(np.abs(run_product(lq, c) - want)) <= TOL


@pytest.mark.parametrize("k", [4, 5, 6])
def test_mfma_matches_vector_kernel_24q(lq, k):
    n = 24
    rng = np.random.default_rng(k)
    qc = lq.QuantumCircuit(n)
    for q in range(n):
        qc.h(q)
        qc.rz(float(rng.uniform(0, 6.28)), q)
    tg = [int(q) for q in rng.permutation(n)[:k]]
    u = haar(k, rng)
    qc.controlled_unitary(u, [int(q) for q in sorted(set(range(n)) - set(tg))[:1]], tg)
    qc.unitary(haar(k, rng), list(range(k)))
    comp = lq.CompiledCircuit(qc, False)
    a = lq.statevector(n)
    b = lq.statevector(n)
    comp.apply(a, plan="gate_by_gate")
    try:
        lq.set_dense_mfma_min_targets(7)
        comp.apply(b, plan="gate_by_gate")
    finally:
        lq.set_dense_mfma_min_targets(4)
    assert abs(a.norm() - 1.0) <= 1e-10
    assert a.max_abs_diff(b) <= TOL


def test_dense_block_on_trajectory_batch_matches_oracle_counts(lq):
    """Noisy runs apply shared ops to the whole batch of shots in one launch
    (blockIdx.y = shot); the counts equal the oracle's draw for draw."""
    n, shots, seed = 10, 64, 41
    rng = np.random.default_rng(3)
    c = O.Circuit(n, n)
    for q in range(n):
        c.h(q)
    c.unitary(haar(5, rng), [9, 0, 4, 2, 7])
    c.cx(0, 5).measure_all()
    chans = [O.depolarizing(0.05), O.amplitude_damping(0.05)]
    want = O.run_counts(c, seed=seed, shots=shots, channels=chans)
    noise = lq.NoiseModel()
    noise.add_all_qubit_channel(lq.KrausChannel.depolarizing(0.05))
    noise.add_all_qubit_channel(lq.KrausChannel.amplitude_damping(0.05))
    got = lq.StatevectorSimulator(seed=seed, noise=noise).run(c.to_product(lq), shots).counts()
    assert dict(got) == want


def test_dense_mfma_switch_rejects_fewer_than_4_targets(lq):
    with pytest.raises(ValueError):
        lq.set_dense_mfma_min_targets(3)


def test_vector_kernel_matches_mfma_at_2_24_groups(lq):
    """29 qubits, 5 targets: 2^24 column groups, the width at which one
    workgroup per group no longer fits a single launch; the vector kernel
    walks the groups and agrees with MFMA."""
    n, k = 29, 5
    rng = np.random.default_rng(29)
    qc = lq.QuantumCircuit(n)
    for q in range(n):
        qc.h(q)
        qc.rz(float(rng.uniform(0, 6.28)), q)
    qc.unitary(haar(k, rng), [3, 27, 0, 12, 20])
    comp = lq.CompiledCircuit(qc, False)
    a = lq.statevector(n)
    comp.apply(a, plan="gate_by_gate")
    b = lq.statevector(n)
    try:
        lq.set_dense_mfma_min_targets(7)
        comp.apply(b, plan="gate_by_gate")
    finally:
        lq.set_dense_mfma_min_targets(4)
    assert a.max_abs_diff(b) <= TOL
    assert abs(b.norm() - 1.0) <= 1e-10


@pytest.mark.parametrize("controls", [[], [25]])
def test_permutation_on_wide_register_moves_basis_state(lq, controls):
    """26 qubits, a 2-qubit permutation: 2^24 groups (2^23 with a control).
    |x> goes to |x with the target bits permuted>; read back through <Z_q>."""
    n = 26
    table = [2, 3, 1, 0]  # local index l -> table[l]
    x = (1 << 25) | (1 << 7) | (1 << 3)  # target bits (3, 7) = local index 1 | 2 = 3 -> 0
    qc = lq.QuantumCircuit(n)
    for q in range(n):
        if (x >> q) & 1:
            qc.x(q)
    if controls:
        qc.controlled_permutation(table, controls, [3, 7])
    else:
        qc.permutation(table, [3, 7])
    sv = lq.StatevectorSimulator(plan="gate_by_gate").run_statevector(qc)
    y = (1 << 25)  # local 3 -> table[3] = 0: bits 3 and 7 cleared
    for q in (0, 3, 7, 25):
        z = sv.expectation_value("".join("Z" if i == q else "I" for i in reversed(range(n))))
        assert z == (-1.0 if (y >> q) & 1 else 1.0)

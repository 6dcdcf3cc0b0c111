"""Controlled permutations (reference src/backend/kernels/gate_kernels.cpp:169-189,
|l> -> |perm[l]> on the control subspace), as Shor's modular exponentiation
uses them (src/algorithms/arithmetic.cpp:10-43: one control, the work register
as targets, y -> a^(2^j) y mod N below N and the identity above).

Two device paths: cycle rotation (one thread per group) when the lowest four
qubits are free, LDS rounds of several groups otherwise. Both against the CPU
oracle at max-abs <= 1e-12 (a permutation moves amplitudes, so the match is
exact), over widths of the permuted register from 1 to 11 qubits."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def scrambled(n, seed):
    rng = np.random.default_rng(seed)
    c = O.Circuit(n)
    for q in range(n):
        c.u(*(float(v) for v in rng.uniform(0, 2 * np.pi, 3)), q)
    for q in range(n - 1):
        c.cx(q, q + 1)
    return c


def run(lq, c):
    sim = lq.StatevectorSimulator(fuse_gates=False, plan="gate_by_gate")
    return np.array(sim.run_statevector(c.to_product(lq)).to_numpy())


def modmul_table(a, n_mod, k):
    """arithmetic.cpp:31-37: y -> a*y mod N for y < N, identity above."""
    return [(a * y) % n_mod if y < n_mod else y for y in range(1 << k)]


@pytest.mark.parametrize("k", [1, 2, 4, 7, 11])
@pytest.mark.parametrize("layout", ["high", "low", "mixed"])
@pytest.mark.parametrize("nc", [0, 1, 2])
def test_controlled_permutation_matches_oracle(lq, k, layout, nc):
    n = 15
    rng = np.random.default_rng(100 * k + nc)
    if layout == "high":      # lowest qubits free: cycle rotation (controls on 0..3 when no room above)
        tg = list(range(n - k, n))[::-1]
        free = list(range(4, n - k)) if n - k - 4 >= nc else list(range(n - k))
    elif layout == "low":     # targets on the low qubits: LDS rounds
        tg = list(range(k))
        free = list(range(k, n))
    else:                     # interleaved
        tg = [int(q) for q in rng.permutation(n)[:k]]
        free = [q for q in range(n) if q not in tg]
    ctl = [int(q) for q in rng.permutation(free)[:nc]]
    assert len(ctl) == nc
    c = scrambled(n, seed=k)
    table = [int(v) for v in rng.permutation(1 << k)]
    if nc:
        c.controlled_permutation(table, ctl, tg)
    else:
        c.permutation(table, tg)
    c.h(0)
    want = O.run_statevector(c, fuse=False)
    assert np.max(np.abs(run(lq, c) - want)) <= TOL


@pytest.mark.parametrize("n_mod,a", [(15, 7), (21, 2), (221, 5)])
def test_modular_exponentiation_permutations_match_oracle(lq, n_mod, a):
    """Shor's controlled modular multiplications: counting register on the
    low qubits (Hadamards), work register above it holding |1>."""
    k = int(np.ceil(np.log2(n_mod)))
    n_count = min(2 * k, 16 - k)
    n = n_count + k
    c = O.Circuit(n)
    for q in range(n_count):
        c.h(q)
    c.x(n_count)  # work register = |1>
    work = list(range(n_count, n))
    for j in range(n_count):
        factor = pow(a, 1 << j, n_mod)
        c.controlled_permutation(modmul_table(factor, n_mod, k), [j], work)
    want = O.run_statevector(c, fuse=False)
    assert np.max(np.abs(run(lq, c) - want)) <= TOL


def test_identity_permutation_leaves_register(lq):
    n = 12
    c = scrambled(n, seed=3)
    want = O.run_statevector(c, fuse=False)
    c.permutation(list(range(8)), [9, 10, 11])
    assert np.max(np.abs(run(lq, c) - want)) <= TOL

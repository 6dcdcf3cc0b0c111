"""Circuits shared by the golden-fixture generator and the tests."""
from __future__ import annotations

import math

import numpy as np

from oracle import oracle as O


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


def every_gate_circuit(n=6, seed=3):
    """Every builder method of the gate model once or more (n >= 6)."""
    rng = np.random.default_rng(seed)
    c = O.Circuit(n)
    for q in range(n):
        c.h(q)
    c.rp(0.7, "XYZ", [0, n - 1, 2]).rp(-1.1, "ZZ", [1, 3]).rp(0.4, "YIX", [n - 2, 0, 3])
    c.ch(0, n - 1).cry(0.3, n - 1, 1).crz(0.9, 2, 3).crx(0.2, 1, 0).cy(3, 4)
    c.ryy(0.5, 0, n - 1).rzz(0.25, 1, 2).rxx(0.75, n - 1, 3).iswap(2, 4)
    c.mcx([0, 1, 2], n - 1).mcp(0.6, [n - 1, 1], 0).s(1).sdg(2).t(0).tdg(3).y(4).z(5).i(0)
    c.p(0.3, 2).ry(0.8, 1).rz(1.3, 0).rx(-0.4, 3).sx(4).u(0.3, 1.2, -0.7, 5).x(2)
    m = np.linalg.qr(rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8)))[0]
    c.unitary(m, [0, n - 1, 2])
    m2 = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]
    c.controlled_unitary(m2, [n - 1], [0, 1])
    c.controlled_permutation([1, 2, 3, 0], [n - 1], [2, 3])
    c.permutation([int(v) for v in rng.permutation(8)], [1, n - 1, 3])
    c.cswap(0, 1, 2).ccx(3, 4, 5).swap(0, 5).cp(0.45, 2, 3).cz(1, 4).cx(5, 0)
    return c


def sampling_circuit():
    c = O.Circuit(3, 3)
    c.h(0).t(0).cx(0, 1).rx(0.4, 2).sx(2).cx(1, 2).measure_all()
    return c


def midcircuit_circuit():
    c = O.Circuit(2, 2)
    c.h(0).t(0).measure(0, 0).h(1).cx(0, 1).reset(0).h(0).measure(1, 1)
    return c


def noisy_circuit():
    c = O.Circuit(3, 3)
    c.h(0).cx(0, 1).t(1).cx(1, 2).u(0.3, 0.2, 0.1, 2).measure_all()
    return c


NOISE = [O.depolarizing(0.1), O.amplitude_damping(0.2)]


def gradient_case():
    """tests/test_pauli_rotation.cpp:178-203 of the reference."""
    h = [(1.0, "ZZ"), (0.5, "XI"), (-0.3, "IZ")]
    t = O.Circuit(2)
    t.ry(0.0, 0).ry(0.0, 1).cx(0, 1).rz(0.0, 0).rp(0.0, "XX", [0, 1])
    return t, [0, 1, 3, 4], [0.3, -0.8, 1.1, 0.5], h


def pauli_ansatz(n, n_rot, seed):
    """Random Pauli-rotation ansatz: H on all, n_rot rotations, CX(0,1), RY(2)."""
    from litequantumcomputingsimulator_amd import workloads as W

    rots, theta = W.random_pauli_ansatz(n, n_rot, seed)
    t = O.Circuit(n)
    for q in range(n):
        t.h(q)
    param_inst = []
    for p, qs in rots:
        param_inst.append(len(t.insts))
        t.rp(0.0, p, qs)
    t.cx(0, 1)
    param_inst.append(len(t.insts))
    t.ry(0.0, 2)
    return t, param_inst, list(theta) + [0.37], rots


def qft_basis(n, x):
    from litequantumcomputingsimulator_amd import workloads as W

    c = O.Circuit(n)
    W.basis_prep(c, n, x)
    W.qft_into(c, n)
    return c


def dft(n, x):
    k = np.arange(1 << n)
    return np.exp(2j * np.pi * ((x * k) % (1 << n)) / (1 << n)) / math.sqrt(1 << n)


def noisy_ghz(n, layers=10, seed=7, p=1e-3, gamma=1e-3):
    """The noisy-trajectory workload at width n: GHZ prep + `layers` random
    U3/CZ layers, depolarizing(p) then amplitude_damping(gamma) after every
    gate, measure_all. Returns (circuit, channels)."""
    from litequantumcomputingsimulator_amd import workloads as W

    c = O.Circuit(n, n)
    W.noisy_ghz_layers(c, n, layers, seed)
    c.measure_all()
    return c, [O.depolarizing(p), O.amplitude_damping(gamma)]

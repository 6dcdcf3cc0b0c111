"""Synthetic workloads of the engine's benchmark configurations.

Each builder writes into any object with the QuantumCircuit builder methods
(pylqcs.QuantumCircuit, or a recorder with the same method names), so the same
seeded circuit can be replayed into several backends. Randomness comes from
numpy's PCG64 seeded generator.
"""
from __future__ import annotations

import math

import numpy as np


def basis_prep(qc, n: int, x: int):
    """X on the set bits of x (|x> from |0>)."""
    for b in range(n):
        if (x >> b) & 1:
            qc.x(b)
    return qc


def qft_into(qc, n: int, do_swaps: bool = True):
    """QFT ladder: H(i), CP(pi/2^(i-j), j, i) for j < i, then swaps
    (the reference's src/algorithms/qft.cpp:7-23)."""
    for i in range(n - 1, -1, -1):
        qc.h(i)
        for j in range(i - 1, -1, -1):
            qc.cp(math.pi / float(1 << (i - j)), j, i)
    if do_swaps:
        for i in range(n // 2):
            qc.swap(i, n - 1 - i)
    return qc


def qft_from_basis(qc, n: int, seed: int):
    """QFT(n) on |x>, x drawn uniformly from [0, 2^n) by `seed`."""
    x = int(np.random.default_rng(seed).integers(0, 1 << n))
    basis_prep(qc, n, x)
    qft_into(qc, n)
    return x


def random_u3_cx_layers(qc, n: int, layers: int, seed: int):
    """`layers` x [U(theta, phi, lambda) on every qubit with angles uniform in
    [0, 2pi) drawn in that order, then CX on a random perfect matching]."""
    rng = np.random.default_rng(seed)
    for _ in range(layers):
        angles = rng.uniform(0.0, 2.0 * math.pi, size=(n, 3))
        for q in range(n):
            qc.u(float(angles[q, 0]), float(angles[q, 1]), float(angles[q, 2]), q)
        perm = rng.permutation(n)
        for k in range(0, n - 1, 2):
            qc.cx(int(perm[k]), int(perm[k + 1]))
    return qc


def gate_count_u3_cx(n: int, layers: int) -> int:
    return layers * (n + n // 2)


def noisy_ghz_layers(qc, n: int, layers: int, seed: int):
    """H(0) + CX chain, then `layers` x [U3 on all qubits + CZ on a random
    perfect matching]."""
    rng = np.random.default_rng(seed)
    qc.h(0)
    for q in range(n - 1):
        qc.cx(q, q + 1)
    for _ in range(layers):
        angles = rng.uniform(0.0, 2.0 * math.pi, size=(n, 3))
        for q in range(n):
            qc.u(float(angles[q, 0]), float(angles[q, 1]), float(angles[q, 2]), q)
        perm = rng.permutation(n)
        for k in range(0, n - 1, 2):
            qc.cz(int(perm[k]), int(perm[k + 1]))
    return qc


def random_pauli_ansatz(n: int, n_rot: int, seed: int):
    """(strings, qubit lists, thetas) of n_rot Pauli rotations with weight
    w ~ U{2..8} on w distinct qubits, letters ~ U{X,Y,Z}, theta ~ U(-pi, pi)."""
    rng = np.random.default_rng(seed)
    rots = []
    for _ in range(n_rot):
        w = int(rng.integers(2, min(8, n) + 1))
        qs = [int(q) for q in rng.choice(n, size=w, replace=False)]
        p = "".join("XYZ"[int(k)] for k in rng.integers(0, 3, size=w))
        rots.append((p, qs))
    theta = rng.uniform(-math.pi, math.pi, size=n_rot)
    return rots, theta


def random_hamiltonian(n: int, n_terms: int, seed: int):
    """n_terms full-width Pauli strings (letters ~ U{I,X,Y,Z}, all-I rejected)
    with N(0,1) coefficients: [(coeff, pauli), ...]."""
    rng = np.random.default_rng(seed)
    terms = []
    while len(terms) < n_terms:
        s = "".join("IXYZ"[int(k)] for k in rng.integers(0, 4, size=n))
        if set(s) == {"I"}:
            continue
        terms.append((float(rng.standard_normal()), s))
    return terms


def random_u3_cx_cp(qc, n: int, depth: int, seed: int):
    """depth x [U3 on every qubit, then a random matching of CX or CP(phi)]."""
    rng = np.random.default_rng(seed)
    for _ in range(depth):
        angles = rng.uniform(0.0, 2.0 * math.pi, size=(n, 3))
        for q in range(n):
            qc.u(float(angles[q, 0]), float(angles[q, 1]), float(angles[q, 2]), q)
        perm = rng.permutation(n)
        for k in range(0, n - 1, 2):
            a, b = int(perm[k]), int(perm[k + 1])
            if rng.integers(0, 2):
                qc.cx(a, b)
            else:
                qc.cp(float(rng.uniform(0.0, 2.0 * math.pi)), a, b)
    return qc

#!/usr/bin/env python3
"""Writes the exact outcome distributions of the noisy-trajectory workload
(GHZ prep + 10 random U3/CZ layers, depolarizing(p) then
amplitude_damping(gamma) after every gate on its qubits, measure_all), from
the reference's own DensityMatrixSimulator
(src/backend/density_matrix_simulator.cpp) compiled into oracle/_ref:

  dm_noisy_ghz12.npy         12 qubits, p = gamma = 1e-3 (the workload's rates)
  dm_noisy_ghz10_strong.npy  10 qubits, p = gamma = 0.05 (noise strong enough
                             that a mistake in the channels shows in 2048 shots)

Each array is data: 2^n probabilities, index = measured bitstring value.
Takes a few minutes on 8 cores."""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests import cases  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = {
    "dm_noisy_ghz12.npy": dict(n=12),
    "dm_noisy_ghz10_strong.npy": dict(n=10, p=0.05, gamma=0.05),
}


def main(names=None):
    if not O.ref_available():
        raise SystemExit("oracle/_ref not built: run oracle/build_ref.sh first")
    for name, kw in FIXTURES.items():
        if names and name not in names:
            continue
        c, chans = cases.noisy_ghz(**kw)
        p = O.ref_density_probabilities(c, chans)
        np.save(os.path.join(HERE, name), p)
        print("wrote", name, "sum", p.sum())


if __name__ == "__main__":
    main(sys.argv[1:])

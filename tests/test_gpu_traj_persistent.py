"""The persistent trajectory kernel (one launch for every shot, a workgroup
per shot with its register in LDS, branch decisions on the device) against
the shared tiled runs and the oracle (statevector_simulator.cpp:122-190):
the same counts from the same draws, its width and op limits refused with an
error when it is asked for, and registers narrower than its 6-qubit tile."""
import numpy as np
import pytest

from litequantumcomputingsimulator_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _noise(lq, p, gamma):
    noise = lq.NoiseModel()
    noise.add_all_qubit_channel(lq.KrausChannel.depolarizing(p))
    noise.add_all_qubit_channel(lq.KrausChannel.amplitude_damping(gamma))
    return noise


def _counts(lq, qc, seed, shots, noise, kernel):
    sim = lq.StatevectorSimulator(seed=seed, noise=noise, trajectory_kernel=kernel)
    return dict(sim.run(qc, shots).counts())


@pytest.mark.parametrize("n,rate", [(12, 1e-3), (10, 0.02), (6, 0.05)])
def test_persistent_equals_shared_runs_on_noisy_layers(lq, n, rate):
    """GHZ prep + 10 random U3/CZ layers with depolarizing and amplitude
    damping after every gate, 2048 shots: both kernels give the same counts."""
    qc = lq.QuantumCircuit(n, n)
    W.noisy_ghz_layers(qc, n, 10, 5)
    qc.measure_all()
    noise = _noise(lq, rate, rate)
    a = _counts(lq, qc, 77, 2048, noise, "shared")
    b = _counts(lq, qc, 77, 2048, noise, "persistent")
    assert a == b
    assert len(a) > 1


@pytest.mark.parametrize("n", [1, 2, 5])
def test_persistent_narrow_registers_match_oracle(lq, n):
    """Registers narrower than the kernel's 6-qubit tile (its extra qubits
    stay |0>), with a mid-circuit measurement and a reset."""
    rng = np.random.default_rng(n)
    c = O.Circuit(n, n)
    for q in range(n):
        c.u(*(float(v) for v in rng.uniform(0, 2 * np.pi, 3)), q)
    c.measure(0, 0).reset(0).h(0)
    for q in range(n - 1):
        c.cx(q, q + 1)
    c.measure_all()
    chans = [O.depolarizing(0.05), O.amplitude_damping(0.1)]
    want = O.run_counts(c, seed=3, shots=500, channels=chans)
    got = _counts(lq, c.to_product(lq), 3, 500, _noise(lq, 0.05, 0.1), "persistent")
    assert got == want


def test_persistent_refuses_what_it_cannot_run(lq):
    qc = lq.QuantumCircuit(13, 13)
    qc.h(0).cx(0, 12).measure_all()
    with pytest.raises(ValueError):
        _counts(lq, qc, 1, 8, _noise(lq, 0.01, 0.01), "persistent")
    assert sum(_counts(lq, qc, 1, 8, _noise(lq, 0.01, 0.01), "auto").values()) == 8
    wide = lq.QuantumCircuit(4, 4)
    wide.h(0).ccx(0, 1, 2).measure_all()  # two targets at most: CCX is one target with two controls
    assert sum(_counts(lq, wide, 1, 8, _noise(lq, 0.01, 0.01), "persistent").values()) == 8
    u = np.eye(8, dtype=complex)[[1, 2, 3, 4, 5, 6, 7, 0]]
    three = lq.QuantumCircuit(4, 4)
    three.h(0).unitary(u, [0, 1, 2]).measure_all()
    with pytest.raises(ValueError):
        _counts(lq, three, 1, 8, _noise(lq, 0.01, 0.01), "persistent")
    assert sum(_counts(lq, three, 1, 8, _noise(lq, 0.01, 0.01), "auto").values()) == 8
    with pytest.raises(ValueError):
        lq.StatevectorSimulator(trajectory_kernel="fast")

"""Trajectory path (reference statevector_simulator.cpp:122-190,
collapse_qubit gate_kernels.cpp:227-247). Trailing measurements:
the engine samples them from sums over the amplitudes consistent with the
outcomes so far instead of collapsing the register; the counts equal the
oracle's, which collapses, draw for draw. Orders that are not the low-qubit
prefix, repeated qubits, clbit remapping, and a measurement followed by
gates (so only the last ones are trailing) are covered. Events: branches of
state-independent channels (unitary default or not) and state-dependent
draws, at batch sizes that put one, few or many events on a site. Each
case runs through both trajectory kernels: the persistent one (a workgroup
per shot, register in LDS) and the shared tiled runs over a batch in HBM."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def noisy_body(n, seed):
    rng = np.random.default_rng(seed)
    c = O.Circuit(n, n)
    for q in range(n):
        c.u(*(float(v) for v in rng.uniform(0, 2 * np.pi, 3)), q)
    for q in range(n - 1):
        c.cx(q, q + 1)
    for q in range(n):
        c.ry(float(rng.uniform(0, 3)), q)
    return c


KERNELS = ["shared", "persistent"]


def counts(lq, c, seed, shots, chans, kernel):
    noise = lq.NoiseModel()
    for name, p in chans:
        noise.add_all_qubit_channel(getattr(lq.KrausChannel, name)(p))
    sim = lq.StatevectorSimulator(seed=seed, noise=noise, trajectory_kernel=kernel)
    return dict(sim.run(c.to_product(lq), shots).counts())


CH = [("depolarizing", 0.02), ("amplitude_damping", 0.03)]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("order", ["prefix", "reversed", "scrambled", "repeat"])
def test_trailing_measurements_match_oracle(lq, order, kernel):
    n, seed, shots = 9, 17, 300
    c = noisy_body(n, 5)
    qs = list(range(n))
    if order == "reversed":
        qs = qs[::-1]
    elif order == "scrambled":
        qs = [int(q) for q in np.random.default_rng(1).permutation(n)]
    elif order == "repeat":
        qs = [0, 3, 0, 5, 3, 8, 1]
    for k, q in enumerate(qs):
        c.measure(q, (k * 5) % n)
    want = O.run_counts(c, seed=seed, shots=shots, channels=[getattr(O, a)(b) for a, b in CH])
    assert counts(lq, c, seed, shots, CH, kernel) == want


@pytest.mark.parametrize("kernel", KERNELS)
def test_measurement_then_gates_then_trailing(lq, kernel):
    n, seed, shots = 8, 23, 300
    c = noisy_body(n, 9)
    c.measure(2, 0).h(2).cx(2, 5).measure(4, 1)
    c.rx(0.4, 4).measure(5, 2).measure(4, 3).measure(0, 4)
    want = O.run_counts(c, seed=seed, shots=shots, channels=[getattr(O, a)(b) for a, b in CH])
    assert counts(lq, c, seed, shots, CH, kernel) == want


@pytest.mark.parametrize("kernel", KERNELS)
def test_noiseless_midcircuit_trailing(lq, kernel):
    c = O.Circuit(3, 3)
    c.h(0).cx(0, 1).measure(0, 0).h(0).t(2).h(2).measure(2, 1).measure(1, 2)
    want = O.run_counts(c, seed=4, shots=400)
    sim = lq.StatevectorSimulator(seed=4, trajectory_kernel=kernel)
    assert dict(sim.run(c.to_product(lq), 400).counts()) == want


@pytest.mark.parametrize("kernel", KERNELS)
def test_mixture_with_unitary_default_branch(lq, kernel):
    """A state-independent channel whose default branch is not the identity
    (sqrt(1-p) S, sqrt(p) X): the default rides on every register and an
    event register takes X S^dagger after it, in the shared run."""
    import math

    p = 0.08
    a, b = math.sqrt(1 - p), math.sqrt(p)
    ops = [[a, 0, 0, 1j * a], [0, b, b, 0]]
    n, seed, shots = 7, 31, 300
    c = noisy_body(n, 3)
    c.measure_all()
    want = O.run_counts(c, seed=seed, shots=shots, channels=[ops, O.amplitude_damping(0.04)])
    noise = lq.NoiseModel()
    noise.add_all_qubit_channel(lq.KrausChannel([np.array(o, complex) for o in ops]))
    noise.add_all_qubit_channel(lq.KrausChannel.amplitude_damping(0.04))
    sim = lq.StatevectorSimulator(seed=seed, noise=noise, trajectory_kernel=kernel)
    got = dict(sim.run(c.to_product(lq), shots).counts())
    assert got == want


@pytest.mark.parametrize("max_batch", [1, 3, 64])
def test_event_handling_independent_of_batch_size(lq, max_batch):
    """Single-register ops, deferred defaults and skip launches depend on how
    many events share a site within a batch (shared tiled runs); the counts
    do not."""
    n, seed, shots = 8, 12, 200
    c = noisy_body(n, 4)
    c.measure_all()
    ch = [O.depolarizing(0.06), O.amplitude_damping(0.06)]
    want = O.run_counts(c, seed=seed, shots=shots, channels=ch)
    noise = lq.NoiseModel()
    noise.add_all_qubit_channel(lq.KrausChannel.depolarizing(0.06))
    noise.add_all_qubit_channel(lq.KrausChannel.amplitude_damping(0.06))
    sim = lq.StatevectorSimulator(seed=seed, noise=noise, max_batch=max_batch, trajectory_kernel="shared")
    assert dict(sim.run(c.to_product(lq), shots).counts()) == want

"""The learned planning cost model: the CPU restatement pinned by finite
differences, the device network's forward and gradients against it, training
on the device, plan features, and the learned plan's effect on results (none:
same state as greedy) and on time (measured against greedy)."""
import os

import numpy as np
import pytest

from oracle import mlp_oracle as M


@pytest.fixture(scope="module")
def lq():
    from litequantumcomputingsimulator_amd import pylqcs

    return pylqcs


def _data(F=7, rows=40, seed=3):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(rows, F))
    t = np.sin(X[:, 0]) + 0.5 * X[:, 1] ** 2 - 0.3 * X[:, 2]
    p = rng.normal(0, 0.3, M.num_params(F))
    return X, t, p


def test_oracle_gradient_matches_finite_differences():
    X, t, p = _data()
    g = M.grad(p, X, t)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(len(p), 60, replace=False), [len(p) - 1, 7 * 64, 7 * 64 + 64 + 5]])
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (M.loss(p + e, X, t) - M.loss(p - e, X, t)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), k


def test_oracle_gradient_check_catches_a_dropped_relu_mask():
    X, t, p = _data()
    W1, b1, W2, b2, w3, b3 = M.unpack(p, X.shape[1])
    Z1 = X @ W1 + b1
    A1 = np.maximum(Z1, 0.0)
    Z2 = A1 @ W2 + b2
    A2 = np.maximum(Z2, 0.0)
    dy = 2.0 * (A2 @ w3 + b3 - t) / len(t)
    dZ2 = np.outer(dy, w3) * (Z2 > 0)
    wrong_dW1 = X.T @ (dZ2 @ W2.T)  # mask of layer 1 dropped
    right = M.grad(p, X, t)[: X.shape[1] * 64].reshape(X.shape[1], 64)
    assert np.max(np.abs(wrong_dW1 - right)) > 1e-3


def test_model_file_round_trip(tmp_path, lq):
    from litequantumcomputingsimulator_amd import costmodel as CM

    F = lq.COST_FEATURES
    rng = np.random.default_rng(1)
    m = CM.Model(rng.normal(size=lq.mlp_num_params(F)), rng.normal(size=F), rng.uniform(0.5, 2, F), -1.25, 0.75)
    path = str(tmp_path / "m.txt")
    m.save(path)
    back = CM.Model.load(path)
    assert np.array_equal(back.params, m.params) and np.array_equal(back.x_scale, m.x_scale)
    assert back.t_mean == m.t_mean and back.t_scale == m.t_scale
    lq.costmodel_load(path)  # parses in C++; no device needed
    lq.costmodel_uninstall()
    with pytest.raises(ValueError):
        lq.costmodel_install([0.0] * 10, [0.0] * F, [1.0] * F, 0.0, 1.0)


def test_plan_features_describe_the_plan(lq):
    from litequantumcomputingsimulator_amd import workloads as W

    n = 20
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_cp(qc, n, 6, seed=2)
    comp = lq.CompiledCircuit(qc, True)
    for plan, low in (("greedy", 4), ("gain", 3)):
        f = comp.features(plan=plan, tile_qubits=12, low_qubits=low)
        d = comp.describe(plan=plan, tile_qubits=12, low_qubits=low)
        assert f.shape == (len(d), lq.COST_FEATURES)
        assert np.all(f[:, 0] == 1) and np.all(f[:, 1] == n)
        assert [int(x) for x in f[:, 3]] == [p["phases"] for p in d]
        assert [int(x) for x in f[:, 17]] == [p["ops"] for p in d]
        assert np.all(f[:, 16] >= low)  # contiguous low qubits of every tile
        # op counts by kind add up to the ops of each pass, less the
        # singly-controlled X ops merged into select ops (one per select op;
        # the merge count itself is checked exactly by the register-tile
        # emulator)
        kinds = f[:, 6:13].sum(axis=1)
        assert np.all(kinds <= f[:, 17]) and np.all(f[:, 17] - kinds <= kinds)
    # with no singly-controlled gate in the circuit nothing merges: the counts
    # add up exactly
    qd = lq.QuantumCircuit(n)
    rng = np.random.default_rng(3)
    for _ in range(6):
        for q in range(n):
            qd.u(*(float(v) for v in rng.uniform(0, 6.28, 3)), q)
        for q in range(0, n - 1, 2):
            qd.iswap(q, q + 1)
    fd = lq.CompiledCircuit(qd, True).features(plan="greedy", tile_qubits=12, low_qubits=4)
    assert np.all(fd[:, 6:13].sum(axis=1) == fd[:, 17])
    # without an installed model the learned plan is the greedy one
    lq.costmodel_uninstall()
    g = comp.features(plan="greedy", tile_qubits=12, low_qubits=4)
    ln = comp.features(plan="learned", tile_qubits=12, low_qubits=4)
    assert np.array_equal(g, ln)


def test_dense_plan_features_and_pick(lq):
    """Dense-block candidates as the cost model sees them: every block is a
    stand-alone 5-target controlled unitary without controls (kind, k,
    controls), and the pick rule takes a dense candidate only when it is
    predicted at least MARGIN faster than the greedy default."""
    from litequantumcomputingsimulator_amd import costmodel as CM
    from litequantumcomputingsimulator_amd import workloads as W

    n = 16
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_cp(qc, n, 4, seed=6)
    comp = lq.CompiledCircuit(qc, True)
    f = comp.features(plan="dense", tile_qubits=5)
    d = comp.describe(plan="dense", tile_qubits=5)
    assert f.shape == (len(d), lq.COST_FEATURES)
    assert np.all(f[:, 0] == 0) and np.all(f[:, 1] == n)
    assert np.all(f[:, 18] == 5) and np.all(f[:, 19] == 5) and np.all(f[:, 20] == 0)
    assert [int(x).bit_count() for x in [s["tile_mask"] for s in d]] == [5] * len(d)
    pred = {c: 100.0 for c in CM.CANDIDATES}
    pred[("dense", 5)] = 100.0 * (1 - CM.MARGIN) + 1e-6
    assert CM.pick(pred) == CM.GREEDY_DEFAULT
    pred[("dense", 5)] = 100.0 * (1 - CM.MARGIN) - 1.0
    assert CM.pick(pred) == ("dense", 5)


@pytest.mark.gpu
def test_device_network_matches_oracle(lq):
    X, t, p = _data(F=lq.COST_FEATURES, rows=300, seed=5)
    y = lq.mlp_forward(p, X)
    assert np.max(np.abs(y - M.forward(p, X))) <= 1e-12 * max(1.0, np.max(np.abs(y)))
    loss, g = lq.mlp_loss_grad(p, X, t)
    want = M.grad(p, X, t)
    assert abs(loss - M.loss(p, X, t)) <= 1e-12 * max(1.0, loss)
    assert np.max(np.abs(g - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))


@pytest.mark.gpu
def test_device_training_fits(lq):
    from litequantumcomputingsimulator_amd import costmodel as CM

    rng = np.random.default_rng(9)
    F = lq.COST_FEATURES
    X = rng.normal(size=(2048, F))
    t = np.sin(X[:, 0]) + 0.5 * X[:, 1] ** 2 - 0.3 * X[:, 2] * X[:, 3]
    p0 = CM.init_params(lq, F, 0)
    before = M.loss(p0, X, t)
    p, after = lq.mlp_train(p0, X, t, 60, 128, 2e-3, 7)
    assert after < 0.1 * before
    assert abs(after - M.loss(np.asarray(p), X, t)) <= 1e-9 * max(1.0, after)
    p2, after2 = lq.mlp_train(p0, X, t, 60, 128, 2e-3, 7)
    assert np.array_equal(np.asarray(p2), np.asarray(p))  # reproducible


@pytest.mark.gpu
def test_learned_plan_gives_the_greedy_state(lq):
    from litequantumcomputingsimulator_amd import costmodel as CM
    from litequantumcomputingsimulator_amd import workloads as W

    n = 22
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_cp(qc, n, 12, seed=4)
    comp = lq.CompiledCircuit(qc, True)
    a, b = lq.statevector(n), lq.statevector(n)
    comp.apply(a, plan="greedy", tile_qubits=12, low_qubits=4)
    installed = CM.install_default(lq)
    try:
        if not installed:  # any model will do for this check: a random one
            F = lq.COST_FEATURES
            lq.costmodel_install(list(CM.init_params(lq, F, 3)), [0.0] * F, [1.0] * F, 0.0, 1.0)
        comp.apply(b, plan="learned")
        choice = lq.last_plan_choice()
        assert (choice["plan"] in ("greedy", "gain") and choice["low_qubits"] in (2, 3, 4)) or \
            (choice["plan"] == "dense" and choice["tile_qubits"] in (4, 5, 6))
        assert choice["predicted_ms"] > 0
    finally:
        if not installed:
            lq.costmodel_uninstall()
    assert np.max(np.abs(a.to_numpy() - b.to_numpy())) <= 1e-12


@pytest.mark.gpu
def test_learned_plan_on_specialized_passes_gives_the_greedy_state(lq):
    """plan='learned' with specialized passes consults the specialized-pass
    model (a random one when none is packaged) among 12- and 13-qubit tiles
    at 3..5 low qubits and dense blocks of 4..6 qubits; the state equals
    greedy grouping's on the interpreting kernel."""
    from litequantumcomputingsimulator_amd import costmodel as CM
    from litequantumcomputingsimulator_amd import workloads as W

    n = 20
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_cp(qc, n, 12, seed=5)
    comp = lq.CompiledCircuit(qc, True)
    a, b = lq.statevector(n), lq.statevector(n)
    comp.apply(a, plan="greedy", tile_qubits=12, low_qubits=4, tile_kernel="register")
    packaged = os.path.exists(CM.default_model_path("specialized"))
    if packaged:
        lq.costmodel_load(CM.default_model_path("specialized"), tile_kernel="specialized")
    else:
        F = lq.COST_FEATURES
        lq.costmodel_install(list(CM.init_params(lq, F, 4)), [0.0] * F, [1.0] * F, 0.0, 1.0,
                             tile_kernel="specialized")
    try:
        comp.apply(b, plan="learned", tile_kernel="specialized")
        choice = lq.last_plan_choice()
        assert choice["predicted_ms"] > 0
        if choice["plan"] == "dense":
            assert choice["tile_qubits"] in (4, 5, 6)
        else:
            assert choice["plan"] in ("greedy", "gain") and choice["low_qubits"] in (3, 4, 5)
            assert choice["tile_qubits"] in (12, 13)
    finally:
        if not packaged:
            lq.costmodel_uninstall(tile_kernel="specialized")
    assert a.max_abs_diff(b) <= 1e-12


def _gain_circuits(lq):
    from litequantumcomputingsimulator_amd import costmodel as CM

    specs = [(26, "pauli", 4, 11), (26, "u3cx", 10, 12), (28, "pauli", 3, 13), (28, "u3cxcp", 6, 14),
             (28, "qft", 0, 15), (28, "u3cx", 8, 16), (26, "u3cxcp", 9, 17), (28, "pauli", 2, 18)]
    return [(n, CM.random_circuit(lq, n, fam, depth, seed)) for n, fam, depth, seed in specs]


@pytest.mark.gpu
def test_trained_model_gains_over_greedy(lq):
    """What the trained network buys: on circuits it was not trained on, the
    plan it picks runs faster in total than greedy grouping at 4 low qubits
    (the plan without it), and within 8% of the best candidate in hindsight.
    Each time is the minimum over 3 applies; every candidate gives the same
    state (checked on the device)."""
    from litequantumcomputingsimulator_amd import costmodel as CM

    if not os.path.exists(CM.default_model_path()):
        pytest.skip("no trained cost model packaged")
    model = CM.Model.load(CM.default_model_path())
    tot = {"learned": 0.0, "greedy": 0.0, "best": 0.0}
    for n, qc in _gain_circuits(lq):
        comp = lq.CompiledCircuit(qc, True)
        sv, ref = lq.statevector(n), lq.statevector(n)
        comp.apply(ref, plan="greedy", tile_qubits=12, low_qubits=4)
        meas, pred = {}, {}
        for cand in CM.CANDIDATES:  # (plan, low qubits), or ("dense", block width)
            dense = cand[0] == "dense"
            kw = dict(plan="dense", tile_qubits=cand[1], low_qubits=0) if dense else \
                dict(plan=cand[0], tile_qubits=12, low_qubits=cand[1])
            times = []
            for _ in range(3):
                f, ms, t = CM.time_plan(lq, comp, sv, kw["plan"], kw["low_qubits"], L=kw["tile_qubits"])
                times.append(t)
            meas[cand] = min(times)
            pred[cand] = float(model.predict_ms(lq, f).sum())
            sv.set_basis(0)
            comp.apply(sv, **kw)
            assert sv.max_abs_diff(ref) <= 1e-12, cand
        pick = CM.pick(pred)
        sv.set_basis(0)
        comp.apply(sv, plan="learned")  # the engine makes the same pick (barring a near tie)
        ranked = sorted(pred.values())
        best = min(pred, key=pred.get)
        margin_tie = abs(pred[best] - (1 - CM.MARGIN) * pred[CM.GREEDY_DEFAULT]) <= 1e-9 * pred[best]
        if ranked[1] - ranked[0] > 1e-9 * ranked[0] and not margin_tie:
            c = lq.last_plan_choice()
            assert ((c["plan"], c["tile_qubits"]) if c["plan"] == "dense" else (c["plan"], c["low_qubits"])) == pick
        tot["learned"] += meas[pick]
        tot["greedy"] += meas[("greedy", 4)]
        tot["best"] += min(meas.values())
    print("cost-model gain:", {k: round(v, 3) for k, v in tot.items()},
          f"learned/greedy = {tot['learned'] / tot['greedy']:.4f}")
    assert tot["learned"] <= 0.99 * tot["greedy"]
    # Near the best candidate in hindsight: the model's picks miss it by up to
    # 8% on single circuits, and the engine keeps greedy unless another plan
    # is predicted 2% faster.
    assert tot["learned"] <= 1.08 * tot["best"]


@pytest.mark.gpu
def test_learned_plan_32_qubits_matches_greedy_state(lq):
    """The learned-planner workload at full width: 32 qubits (64 GiB per
    register), random U3/CX/CP layers of depth 40; the learned plan's final
    state equals greedy's within 1e-12 (compared on the device)."""
    from litequantumcomputingsimulator_amd import costmodel as CM
    from litequantumcomputingsimulator_amd import workloads as W

    n = 32
    qc = lq.QuantumCircuit(n)
    W.random_u3_cx_cp(qc, n, 40, seed=32)
    comp = lq.CompiledCircuit(qc, True)
    a = lq.statevector(n)
    comp.apply(a, plan="greedy", tile_qubits=12, low_qubits=4)
    b = lq.statevector(n)
    installed = CM.install_default(lq)
    comp.apply(b, plan="learned")
    if installed:
        assert lq.last_plan_choice()["predicted_ms"] > 0
    assert abs(a.norm() - 1.0) <= 1e-10
    assert a.max_abs_diff(b) <= 1e-12


@pytest.mark.gpu
def test_learned_plan_matches_oracle_on_planner_workload(lq):
    """The learned-planner workload's shape below the reference's 30-qubit
    cap: random U3/CX/CP layers of depth 40 at 20 qubits, run with the plan
    the installed model picks, against the oracle's statevector (the
    reference's fused path restated) at 1e-12."""
    from litequantumcomputingsimulator_amd import costmodel as CM
    from litequantumcomputingsimulator_amd import workloads as W
    from oracle import oracle as O

    n = 20
    ref = O.Circuit(n)
    W.random_u3_cx_cp(ref, n, 40, seed=20)
    want = O.run_statevector(ref, fuse=True)
    CM.install_default(lq)
    comp = lq.CompiledCircuit(ref.to_product(lq), True)
    sv = lq.statevector(n)
    comp.apply(sv, plan="learned")
    choice = lq.last_plan_choice()
    assert choice["plan"] in ("greedy", "gain", "dense")
    assert np.max(np.abs(np.asarray(sv.to_numpy()) - want)) <= 1e-12


@pytest.mark.gpu
def test_learned_plan_picks_dense_blocks_on_a_deep_local_circuit(lq):
    """Specialized passes' learned plan on a deep circuit over 5 of 24 qubits
    (160 U3/CX): the packaged model picks dense blocks (one 5-qubit block
    holds the whole circuit, where a tiled pass pays for every gate), and the
    state equals the oracle's (the reference's fused path) at 1e-12."""
    import math

    from litequantumcomputingsimulator_amd import costmodel as CM
    from oracle import oracle as O

    if not os.path.exists(CM.default_model_path("specialized")):
        pytest.skip("no trained specialized-pass model packaged")
    n = 24
    rng = np.random.default_rng(77)
    c = O.Circuit(n)
    lo = int(rng.integers(0, n - 5))
    for _ in range(160):
        a, b = (int(x) for x in rng.choice(np.arange(lo, lo + 5), 2, replace=False))
        if rng.random() < 0.6:
            c.u(*(float(v) for v in rng.uniform(0, 2 * math.pi, 3)), a)
        else:
            c.cx(a, b)
    want = O.run_statevector(c, fuse=True)
    lq.costmodel_load(CM.default_model_path("specialized"), tile_kernel="specialized")
    comp = lq.CompiledCircuit(c.to_product(lq), True)
    sv = lq.statevector(n)
    comp.apply(sv, plan="learned", tile_kernel="specialized")
    choice = lq.last_plan_choice()
    assert choice["plan"] == "dense" and choice["tile_qubits"] in (5, 6), choice
    assert np.max(np.abs(np.asarray(sv.to_numpy()) - want)) <= 1e-12

"""Learned cost model for the tiled-pass planner.

The engine's LEARNED plan asks a small network how long each pass of a
candidate plan will take. The candidates are greedy and gain tile grouping at
2, 3 and 4 low qubits on the interpreting register tile, or at 3, 4 and 5 low
qubits on 12- and 13-qubit tiles for specialized passes (a model per kernel),
and from 12 qubits dense blocks of 4, 5 and 6 qubits.
It then runs the candidate with the smallest predicted total. The network is a multilayer perceptron COST_FEATURES -> 64 -> 64 -> 1
(ReLU, FP64) on per-pass features: register phases, LDS transposes, op counts
by kind, lane spread of the load layout, tile qubit positions, register width
and batch; for stand-alone ops, k, controls and qubit positions. It predicts
log(milliseconds).

Everything here runs on the device:
* the training data is the engine's own pass timings (HIP events around
  every launch) on random circuits;
* training is Adam on the device through the engine's MLP forward/backward
  kernels (lqcs_mlp_train).

The trained model is a small text file. `default_model_path()` ships with the
package and is installed when pylqcs is imported. Without a model, LEARNED
falls back to greedy grouping.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

from . import workloads as W

FORMAT = "lqcs-costmodel"


def default_model_path(kernel: str = "register") -> str:
    """The packaged model for passes on the interpreting register tile
    (costmodel.txt) or on specialized passes (costmodel_specialized.txt)."""
    name = "costmodel_specialized.txt" if kernel == "specialized" else "costmodel.txt"
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), name)


@dataclass
class Model:
    params: np.ndarray
    x_mean: np.ndarray
    x_scale: np.ndarray
    t_mean: float
    t_scale: float

    def save(self, path: str) -> None:
        F = len(self.x_mean)
        with open(path, "w") as f:
            f.write(f"{FORMAT} 1 {F} {len(self.params)}\n")
            f.write(" ".join(repr(float(v)) for v in self.x_mean) + "\n")
            f.write(" ".join(repr(float(v)) for v in self.x_scale) + "\n")
            f.write(f"{float(self.t_mean)!r} {float(self.t_scale)!r}\n")
            for k in range(0, len(self.params), 8):
                f.write(" ".join(repr(float(v)) for v in self.params[k:k + 8]) + "\n")

    @staticmethod
    def load(path: str) -> "Model":
        with open(path) as f:
            tok = f.read().split()
        if tok[0] != FORMAT or tok[1] != "1":
            raise ValueError(f"{path}: not a version-1 cost model")
        F, n = int(tok[2]), int(tok[3])
        vals = np.array([float(v) for v in tok[4:]])
        if len(vals) != 2 * F + 2 + n:
            raise ValueError(f"{path}: truncated")
        return Model(vals[2 * F + 2:], vals[:F], vals[F:2 * F], float(vals[2 * F]), float(vals[2 * F + 1]))

    def install(self, lq, kernel: str = "register") -> None:
        lq.costmodel_install(list(self.params), list(self.x_mean), list(self.x_scale), self.t_mean, self.t_scale,
                             tile_kernel=kernel)

    def predict_ms(self, lq, features: np.ndarray) -> np.ndarray:
        x = (np.asarray(features, float) - self.x_mean) / self.x_scale
        return np.exp(lq.mlp_forward(self.params, x) * self.t_scale + self.t_mean)


# ---- circuits the timings are taken on -------------------------------------------------

def random_circuit(lq, n: int, family: str, depth: int, seed: int):
    """One circuit of a training/evaluation family (widths n >= 12)."""
    qc = lq.QuantumCircuit(n)
    rng = np.random.default_rng(seed)
    if family == "u3cx":
        W.random_u3_cx_layers(qc, n, depth, seed)
    elif family == "u3cxcp":
        W.random_u3_cx_cp(qc, n, depth, seed)
    elif family == "pauli":
        rots, theta = W.random_pauli_ansatz(n, depth * n // 2, seed)
        for (p, qs), t in zip(rots, theta):
            qc.rp(float(t), p, qs)
    elif family == "local":  # gates concentrated on a random window of qubits
        lo = int(rng.integers(0, n - 8))
        for _ in range(depth * 8):
            a, b = (int(x) for x in rng.choice(np.arange(lo, lo + 8), 2, replace=False))
            if rng.random() < 0.5:
                qc.u(*(float(v) for v in rng.uniform(0, 2 * math.pi, 3)), a)
            else:
                qc.cx(a, b)
    elif family == "deep_local":  # many gates on a random window of 5 qubits: dense blocks' case
        lo = int(rng.integers(0, n - 5))
        for _ in range(depth * 40):
            a, b = (int(x) for x in rng.choice(np.arange(lo, lo + 5), 2, replace=False))
            if rng.random() < 0.6:
                qc.u(*(float(v) for v in rng.uniform(0, 2 * math.pi, 3)), a)
            else:
                qc.cx(a, b)
    elif family == "qft":
        x = int(rng.integers(0, 1 << n))
        W.basis_prep(qc, n, x)
        W.qft_into(qc, n)
    else:
        raise ValueError(family)
    return qc


FAMILIES = ("u3cx", "u3cxcp", "pauli", "local", "qft")
CANDIDATES = tuple((plan, low) for plan in ("greedy", "gain") for low in (2, 3, 4)) + \
    tuple(("dense", k) for k in (4, 5, 6))  # (plan, low qubits), or ("dense", block width)
# Specialized passes: (plan, tile width, low qubits), as the engine plans them
# for registers of 17 qubits and more.
SPEC_CANDIDATES = tuple((plan, L, low) for L in (12, 13) for plan in ("greedy", "gain") for low in (3, 4, 5))
# Dense blocks of k = 4..6 qubits (plan, k, 0): candidates of the specialized
# passes' learned plan from 12 qubits up.
DENSE_CANDIDATES = tuple(("dense", k, 0) for k in (4, 5, 6))
# The engine keeps greedy grouping at the default low qubits (4 at these
# widths) unless another candidate is predicted faster by this fraction.
MARGIN = 0.02
GREEDY_DEFAULT = ("greedy", 4)
SPEC_GREEDY_DEFAULT = ("greedy", 12, 4)


def pick(pred: dict) -> tuple:
    """The engine's LEARNED choice among predicted totals {candidate: ms}
    (candidates as in CANDIDATES, or SPEC_CANDIDATES + DENSE_CANDIDATES)."""
    best = min(pred, key=pred.get)
    base = SPEC_GREEDY_DEFAULT if SPEC_GREEDY_DEFAULT in pred else GREEDY_DEFAULT
    if base in pred and pred[best] > (1.0 - MARGIN) * pred[base]:
        return base
    return best


def time_plan(lq, comp, sv, plan: str, low: int, reps: int = 1, L: int = 12, kernel: str = "register"):
    """(features, per-pass ms, total ms) of one plan of a compiled circuit:
    one untimed apply (which compiles specialized passes), then `reps` timed
    applies (per-pass times averaged)."""
    kw = dict(plan=plan, tile_qubits=L, low_qubits=low, tile_kernel=kernel)
    feats = comp.features(**kw)
    comp.apply(sv, **kw)
    lq.synchronize()
    lq.profile_enable(True)
    lq.profile_read_launches()
    acc = None
    for _ in range(reps):
        comp.apply(sv, **kw)
        ms, _ = lq.profile_read_launches()
        if len(ms) != len(feats):
            raise RuntimeError(f"{len(ms)} launches timed for a {len(feats)}-pass plan")
        acc = ms if acc is None else acc + ms
    lq.profile_enable(False)
    ms = acc / reps
    return feats, ms, float(ms.sum())


def generate(lq, specs, reps: int = 1, log=print, kernel: str = "register", only=None):
    """Times every candidate plan of every (n, family, depth, seed) in specs
    (or the candidates in `only`). Returns (X, t, records): per-pass feature
    rows and t = log(ms), and per circuit a record {"spec", "cands":
    [(candidate, features, total_ms)]}, the candidate as in CANDIDATES
    (register tile) or SPEC_CANDIDATES + DENSE_CANDIDATES."""
    X, T, records = [], [], []
    cands_of = SPEC_CANDIDATES + DENSE_CANDIDATES if kernel == "specialized" else CANDIDATES
    if only is not None:
        cands_of = tuple(c for c in cands_of if c in only)
    base = SPEC_GREEDY_DEFAULT if kernel == "specialized" else GREEDY_DEFAULT
    for n, family, depth, seed in specs:
        qc = random_circuit(lq, n, family, depth, seed)
        comp = lq.CompiledCircuit(qc, True)
        sv = lq.statevector(n)
        cands = []
        for cand in cands_of:
            if cand[0] == "dense" and n < 12:
                continue
            if kernel == "specialized":
                plan, L, low = cand
                try:
                    f, ms, tot = time_plan(lq, comp, sv, plan, low, reps, L=L, kernel=kernel)
                except ValueError:
                    if L == 12:
                        raise
                    continue  # 13-qubit tiles refuse Pauli rotations wider than 4 qubits; so does the engine's pick
            elif cand[0] == "dense":
                f, ms, tot = time_plan(lq, comp, sv, "dense", 0, reps, L=cand[1])
            else:
                plan, low = cand
                f, ms, tot = time_plan(lq, comp, sv, plan, low, reps)
            X.append(f)
            T.append(np.log(np.maximum(ms, 1e-4)))
            cands.append((cand, f, tot))
        records.append({"spec": (n, family, depth, seed), "cands": cands})
        best = min(cands, key=lambda c: c[2])
        greedy = [c[2] for c in cands if c[0] == base]
        log(f"timed n={n} {family} depth={depth} seed={seed}: best {best[0]} {best[2]:.3f} ms"
            + (f", greedy {greedy[0]:.3f} ms" if greedy else ""))
        del sv
    return np.concatenate(X), np.concatenate(T), records


def evaluate(lq, model: "Model", records):
    """Per circuit: measured time of the candidate the model picks (the
    engine's LEARNED choice), of greedy at 4 low qubits (the plan without the
    model) and of the best candidate in hindsight."""
    out = []
    for rec in records:
        pred = {c: float(model.predict_ms(lq, f).sum()) for (c, f, _) in rec["cands"]}
        k = pick(pred)
        meas = {c: t for (c, _, t) in rec["cands"]}
        base = SPEC_GREEDY_DEFAULT if SPEC_GREEDY_DEFAULT in meas else GREEDY_DEFAULT
        out.append({"spec": list(rec["spec"]), "pick": list(k),
                    "learned_ms": meas[k], "greedy_ms": meas[base],
                    "best_ms": min(meas.values()), "predicted_ms": pred[k]})
    return out


def init_params(lq, F: int, seed: int) -> np.ndarray:
    """He-style initialisation of the packed parameters."""
    rng = np.random.default_rng(seed)
    H = 64
    w1 = rng.normal(0, math.sqrt(2.0 / F), (F, H))
    w2 = rng.normal(0, math.sqrt(2.0 / H), (H, H))
    w3 = rng.normal(0, math.sqrt(1.0 / H), H)
    p = np.concatenate([w1.ravel(), np.zeros(H), w2.ravel(), np.zeros(H), w3, np.zeros(1)])
    assert len(p) == lq.mlp_num_params(F)
    return p


def train(lq, X: np.ndarray, T: np.ndarray, epochs: int = 300, batch_rows: int = 256, lr: float = 2e-3,
          seed: int = 0):
    """Fits the network on the device; returns (Model, final loss in
    normalised units)."""
    x_mean = X.mean(axis=0)
    x_scale = X.std(axis=0)
    x_scale[x_scale < 1e-9] = 1.0
    t_mean, t_scale = float(T.mean()), float(T.std() or 1.0)
    xn = (X - x_mean) / x_scale
    tn = (T - t_mean) / t_scale
    p0 = init_params(lq, X.shape[1], seed)
    p, loss = lq.mlp_train(p0, xn, tn, epochs, min(batch_rows, len(X)), lr, seed)
    return Model(np.asarray(p), x_mean, x_scale, t_mean, t_scale), float(loss)


def install_default(lq) -> bool:
    """Installs the packaged models (register tile, specialized passes) that
    are present; returns whether the register tile's was."""
    spec = default_model_path("specialized")
    if os.path.exists(spec):
        lq.costmodel_load(spec, tile_kernel="specialized")
    path = default_model_path()
    if not os.path.exists(path):
        return False
    lq.costmodel_load(path)
    return True


def default_specs(seed: int = 0, widths=(24, 26, 28, 30)):
    """The training set: every family at several widths and depths, plus deep
    random U3/CX and U3/CX/CP circuits (depth 20-40) from 30 qubits up, the
    shape of the planner workload."""
    rng = np.random.default_rng(seed)
    specs = []
    for n in widths:
        for fam in FAMILIES:
            for _ in range(2 if fam != "qft" else 1):
                depth = int(rng.integers(3, 13)) if fam != "pauli" else int(rng.integers(2, 7))
                specs.append((n, fam, depth, int(rng.integers(1 << 30))))
        if n >= 30:
            for _ in range(3 if n >= 32 else 1):
                for fam in ("u3cx", "u3cxcp"):
                    specs.append((n, fam, int(rng.integers(20, 41)), int(rng.integers(1 << 30))))
    return specs


def deep_specs(seed: int = 0, widths=(24, 26, 28, 30)):
    """Deep circuits on a window of 5 qubits (one per width), where dense
    blocks win: part of both models' training sets."""
    rng = np.random.default_rng(seed + 1000)
    return [(n, "deep_local", int(rng.integers(3, 7)), int(rng.integers(1 << 30))) for n in widths]


def main(argv=None) -> int:
    import argparse
    import json
    import time

    from . import pylqcs as lq

    ap = argparse.ArgumentParser(description="time plans on this GPU, train the cost model on it, evaluate")
    ap.add_argument("--kernel", default="register", choices=("register", "specialized"),
                    help="the tile kernel whose pass times the model learns")
    ap.add_argument("--out", default=None)
    ap.add_argument("--summary", default=None, help="JSON file for the held-out evaluation")
    ap.add_argument("--widths", default="24,26,28,30,32")
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    widths = tuple(int(w) for w in args.widths.split(","))
    out = args.out or default_model_path(args.kernel)
    lq.costmodel_uninstall(tile_kernel=args.kernel)
    t0 = time.perf_counter()
    specs = default_specs(args.seed, widths) + deep_specs(args.seed, widths)
    X, T, _ = generate(lq, specs, kernel=args.kernel)
    print(f"training rows: {len(X)} ({time.perf_counter() - t0:.1f} s of timing)", flush=True)
    model, loss = train(lq, X, T, epochs=args.epochs, seed=args.seed)
    print(f"trained: final loss {loss:.4f} (normalised log-ms)", flush=True)
    model.save(out)
    held_specs = default_specs(args.seed + 1, widths) + deep_specs(args.seed + 1, widths)
    _, _, held = generate(lq, held_specs, kernel=args.kernel)
    ev = evaluate(lq, model, held)
    tot = {k: sum(e[k] for e in ev) for k in ("learned_ms", "greedy_ms", "best_ms")}
    print(json.dumps({"held_out_circuits": len(ev), **{k: round(v, 3) for k, v in tot.items()}}), flush=True)
    if args.summary:
        with open(args.summary, "w") as f:
            json.dump({"train_rows": len(X), "final_loss": loss, "totals": tot, "circuits": ev}, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

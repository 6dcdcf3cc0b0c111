"""TEST INFRASTRUCTURE ONLY: a restatement of the reference's redundant-gate
elimination (src/transpiler/redundant_elimination.cpp:60-104), the sweep form:
each sweep walks the instructions keeping, per qubit, the last surviving
instruction that touched it (a barrier forgets all of them, a cancellation
forgets the pair's qubits); an instruction cancels against that last one when
both touch the same qubit list and form an inverse pair
(:15-55, restated in `undoes`); sweeps repeat until nothing changes.

Instructions are tuples (name, params, qubits); names follow OpenQASM
(h, x, y, z, id, s, sdg, t, tdg, rx, ry, rz, u1, u3, cx, cz, swap, ccx, cu1,
crz, barrier, measure). Used by tests/test_transpiler_passes.py to pin the
engine's one-sweep stack form."""
from __future__ import annotations

SELF_INVERSE = {"id", "x", "y", "z", "h", "swap", "cx", "cy", "cz", "ch", "ccx", "cswap"}
NEGATED = {"rx", "ry", "rz", "u1", "cu1", "crx", "cry", "crz"}
PAIRS = {("s", "sdg"), ("sdg", "s"), ("t", "tdg"), ("tdg", "t")}


def undoes(a, b) -> bool:
    (na, pa, _), (nb, pb, _) = a, b
    if na in SELF_INVERSE:
        return nb == na
    if na in NEGATED:
        return nb == na and len(pa) == 1 and len(pb) == 1 and pa[0] == -pb[0]
    if (na, nb) in PAIRS:
        return True
    if na == "u3":
        return nb == "u3" and pb[0] == -pa[0] and pb[1] == -pa[2] and pb[2] == -pa[1]
    return False


def eliminate(insts, n_qubits):
    insts = list(insts)
    changed = True
    while changed:
        changed = False
        removed = [False] * len(insts)
        last = [-1] * n_qubits
        for i, cur in enumerate(insts):
            name, _, qubits = cur
            if name == "barrier":
                last = [-1] * n_qubits
                continue
            cancelled = False
            if name != "measure" and qubits:
                j = last[qubits[0]]
                if j >= 0 and not removed[j] and insts[j][2] == qubits and all(last[q] == j for q in qubits):
                    if undoes(insts[j], cur):
                        removed[i] = removed[j] = True
                        for q in qubits:
                            last[q] = -1
                        cancelled = changed = True
            if not cancelled:
                for q in qubits:
                    last[q] = i
        insts = [x for x, r in zip(insts, removed) if not r]
    return insts

"""The C++ drop-in: the reference's own callers compile and link against
include/ and liblqcs.so, unchanged.

The reference's CMake target keeps its out-of-scope algorithm sources
(src/algorithms/{arithmetic,bernstein_vazirani,deutsch_jozsa,grover,
number_theory,oracles,qpe,shor,simon}.cpp); they are compiled here against the
new headers, and the callers (benchmarks/bench_gates.cpp and every example)
are linked with them against liblqcs.so + liblqcs_device.so. Nothing is run.
Skipped where the reference sources are absent."""
import concurrent.futures as cf
import os
import subprocess

import pytest

from litequantumcomputingsimulator_amd import LIB_DIR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
REF = "/root/reference"

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference sources absent")

KEPT_ALGORITHMS = ["arithmetic", "bernstein_vazirani", "deutsch_jozsa", "grover", "number_theory", "oracles",
                   "qpe", "shor", "simon"]
CALLERS = ["benchmarks/bench_gates.cpp"] + [f"examples/{e}.cpp" for e in (
    "trajectory_noise", "vqe_gradient", "shor", "bell_state", "ghz_state", "qft", "grover", "algorithms_demo",
    "vqe_h2", "noisy_bell")]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return p.returncode, p.stdout


def test_every_reference_header_has_a_counterpart():
    base = os.path.join(REF, "include")
    missing = []
    for d, _, files in os.walk(base):
        for f in files:
            rel = os.path.relpath(os.path.join(d, f), base)
            if not os.path.exists(os.path.join(INCLUDE, rel)):
                missing.append(rel)
    assert not missing, missing


def test_reference_callers_compile_and_link(tmp_path):
    if not os.path.exists(os.path.join(LIB_DIR, "liblqcs.so")):
        pytest.skip("liblqcs.so not built")
    flags = ["g++", "-std=c++20", "-O0", "-w", f"-I{INCLUDE}"]
    objs, jobs = [], []
    for a in KEPT_ALGORITHMS:
        o = str(tmp_path / f"{a}.o")
        objs.append(o)
        jobs.append(flags + ["-fPIC", f"-I{REF}/src", "-c", f"{REF}/src/algorithms/{a}.cpp", "-o", o])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        res = list(ex.map(_run, jobs))
    bad = [(j[-3], out) for j, (rc, out) in zip(jobs, res) if rc != 0]
    assert not bad, bad[0][1][-3000:]
    links = []
    for c in CALLERS:
        exe = str(tmp_path / os.path.basename(c)[:-4])
        links.append(flags + [f"{REF}/{c}", *objs, f"-L{LIB_DIR}", "-llqcs", "-llqcs_device",
                              f"-Wl,-rpath,{LIB_DIR}", "-o", exe])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        res = list(ex.map(_run, links))
    bad = [(l[5], out) for l, (rc, out) in zip(links, res) if rc != 0]
    assert not bad, f"{bad[0][0]}:\n{bad[0][1][-3000:]}"
    for c in CALLERS:
        assert os.path.exists(str(tmp_path / os.path.basename(c)[:-4]))

"""CPU checks of the specialized register-tile passes (spec_pass.hip).

The engine generates one kernel per pass structure and compiles it at run time
for gfx950. Here the same generated text is compiled as host C++ (clang, with
tests/native/spec_shim.hpp standing in for the HIP device names: 256 threads
per workgroup as std::threads, __syncthreads as a std::barrier) and run pass
by pass on a small register. The result must equal the oracle's statevector
(reference src/backend/statevector_simulator.cpp:197-265 through the fused
circuit), so the generator's layouts, thread orders, in-place transposes and
op calls are checked without a device. A second check compiles the 30-qubit
benchmark's passes for gfx950 offline and reads their register use.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from litequantumcomputingsimulator_amd import pylqcs as lq
from litequantumcomputingsimulator_amd import workloads as W
from oracle import oracle as O
from tests import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEVICE_SRC = os.path.join(ROOT, "litequantumcomputingsimulator_amd", "csrc", "device")
CLANG = "/opt/rocm/llvm/bin/clang++"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
READELF = "/opt/rocm/llvm/bin/llvm-readelf"


def _passes(circ, n, L=12, **plan):
    comp = lq.CompiledCircuit(circ.to_product(lq), True)
    steps = comp.describe(tile_qubits=L, tile_kernel="specialized", **plan)
    return [comp.specialized_pass(i, tile_qubits=L, tile_kernel="specialized", **plan) for i in range(len(steps))]


def _emulate(tmp_path, passes, n, state, tag, L=12):
    """Runs every pass's generated kernel on the host, in order."""
    import ctypes
    # Device-only forms become plain host C++: the occupancy hint, the
    # global address space and the asm that keeps a base in scalar registers.
    def host_text(t):
        t = re.sub(r"__attribute__\(\(amdgpu_waves_per_eu\([^)]*\)\)\)", "", t)
        t = t.replace("__attribute__((address_space(1)))", "")
        return t.replace('asm volatile("" : "+s"(g));', "").replace('asm volatile("" : "+v"(x));', "")
    hdr = host_text(open(os.path.join(DEVICE_SRC, "spec_pass.hpp")).read())
    src = ['#include "spec_shim.hpp"', hdr]
    for i, (text, _) in enumerate(passes):
        assert text, f"pass {i} has no specialized kernel"
        src.append(host_text(text).replace("lqcs_pass(", f"lqcs_pass_{i}("))
        src.append(f'extern "C" void run_pass_{i}(double* s, const double* w, unsigned tiles) {{\n'
                   f'    shim_launch([&] {{ lqcs_pass_{i}(reinterpret_cast<sp_amp*>(s), '
                   f'reinterpret_cast<const sp_amp*>(w)); }}, tiles, 1, {1 << (L - 4)});\n}}')
    cpp = tmp_path / f"{tag}.cpp"
    so = tmp_path / f"{tag}.so"
    cpp.write_text("\n".join(src))
    subprocess.run([CLANG, "-std=c++20", "-O1", "-shared", "-fPIC", "-I", os.path.join(ROOT, "tests", "native"),
                    "-I", DEVICE_SRC, str(cpp), "-o", str(so), "-lpthread"], check=True, capture_output=True)
    lib = ctypes.CDLL(str(so))
    buf = np.ascontiguousarray(np.stack([state.real, state.imag], axis=-1).astype(np.float64))
    dp = ctypes.POINTER(ctypes.c_double)
    for i, (_, words) in enumerate(passes):
        w = np.ascontiguousarray(words, dtype=np.float64)
        getattr(lib, f"run_pass_{i}")(buf.ctypes.data_as(dp), w.ctypes.data_as(dp), ctypes.c_uint(1 << (n - L)))
    return buf[..., 0] + 1j * buf[..., 1]


def _u3cx(n, layers, seed):
    c = O.Circuit(n)
    W.random_u3_cx_layers(c, n, layers, seed=seed)
    return c


def _rotations(n, seed):
    """Pauli rotations of weight 1-4 (wider X parts run as stand-alone steps)."""
    rng = np.random.default_rng(seed)
    c = O.Circuit(n)
    for _ in range(60):
        w = int(rng.integers(1, 5))
        qs = [int(v) for v in rng.permutation(n)[:w]]
        p = "".join(str(rng.choice(list("XYZ"))) for _ in range(w))
        c.rp(float(rng.uniform(-3.14, 3.14)), p, qs)
    return c


def _diag_rich(n, seed):
    """RZ, P, T, S, CZ, CP, CRZ, MCP, CY and controlled U3s among U3 and CX."""
    rng = np.random.default_rng(seed)
    c = O.Circuit(n)
    for _ in range(90):
        q = [int(v) for v in rng.permutation(n)[:3]]
        a = float(rng.uniform(0, 6.28))
        k = int(rng.integers(0, 11))
        if k == 0: c.rz(a, q[0])
        elif k == 1: c.p(a, q[0])
        elif k == 2: c.t(q[0]).s(q[1])
        elif k == 3: c.cz(q[0], q[1])
        elif k == 4: c.cp(a, q[0], q[1])
        elif k == 5: c.crz(a, q[0], q[1])
        elif k == 6: c.mcp(a, q[:2], q[2])
        elif k == 7: c.cy(q[0], q[1])
        elif k == 8: c.cx(q[0], q[1])
        elif k == 9: c.cry(a, q[0], q[1])
        else: c.u(a, float(rng.uniform(0, 6.28)), float(rng.uniform(0, 6.28)), q[0])
    return c


def _spread(c, n, seed):
    """U3 on every qubit first, so that every amplitude of every tile is
    nonzero when the circuit proper starts."""
    rng = np.random.default_rng(seed)
    full = O.Circuit(n)
    for q in range(n):
        full.u(*(float(v) for v in rng.uniform(0, 6.28, 3)), q)
    full.insts += c.insts
    full.calls += c.calls
    return full


@pytest.mark.skipif(not os.path.exists(CLANG), reason="clang++ not available")
@pytest.mark.parametrize("name,n,make,plan", [
    ("u3cx_13q_tiles", 15, lambda: _u3cx(15, 6, 8), {"plan": "greedy", "L": 13, "low_qubits": 5}),
    ("random16kinds_13q_tiles", 14, lambda: cases.random_circuit(14, 150, 9), {"plan": "gain", "L": 13}),
    ("u3cx", 14, lambda: _u3cx(14, 6, 3), {"plan": "greedy"}),
    ("u3cx_gain", 14, lambda: _u3cx(14, 6, 4), {"plan": "gain", "low_qubits": 3}),
    # a pass whose first two layouts share only low qubits: its transpose is
    # split by a lane bit (half the lanes of every wave per round)
    ("u3cx_lane_split", 15, lambda: _u3cx(15, 6, 16), {"plan": "greedy"}),
    ("random16kinds", 13, lambda: cases.random_circuit(13, 150, 4), {"plan": "greedy"}),
    ("rotations", 13, lambda: _rotations(13, 5), {"plan": "greedy"}),
    ("diagonals", 13, lambda: _diag_rich(13, 6), {"plan": "greedy", "low_qubits": 2}),
])
def test_generated_passes_equal_the_oracle(tmp_path, name, n, make, plan):
    circ = _spread(make(), n, 17)
    plan = dict(plan)
    L = plan.pop("L", 12)
    passes = _passes(circ, n, L=L, **plan)
    assert passes and all(t for t, _ in passes)
    if name == "u3cx_lane_split":
        assert any(re.search(r"if \(!?\(?t & [1248]u\)", t) for t, _ in passes)
    e0 = np.zeros(1 << n, complex)
    e0[0] = 1.0
    got = _emulate(tmp_path, passes, n, e0, name, L=L)
    ref = O.run_statevector(circ, fuse=True)
    assert np.max(np.abs(got - ref)) <= 1e-12


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_benchmark_passes_fit_four_waves_per_simd(tmp_path):
    """The 30-qubit benchmark's passes compiled for gfx950: at most 128 VGPRs
    (four waves per SIMD) and no scratch spills."""
    circ = _u3cx(30, 20, 1234)
    comp = lq.CompiledCircuit(circ.to_product(lq), True)
    hdr = open(os.path.join(DEVICE_SRC, "spec_pass.hpp")).read()
    for step in range(len(comp.describe(plan="greedy"))):
        text, words = comp.specialized_pass(step, plan="greedy")
        assert text and len(words) > 0
        src = tmp_path / f"pass{step}.hip"
        src.write_text("#include <hip/hip_runtime.h>\n" + hdr + "\n" + text)
        obj = tmp_path / f"pass{step}.o"
        subprocess.run([HIPCC, "-std=c++17", "-O3", "--offload-arch=gfx950", "--cuda-device-only", "--no-gpu-bundle-output", "-c",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", str(src), "-o", str(obj)],
                       check=True, capture_output=True)
        notes = subprocess.run([READELF, "--notes", str(obj)], check=True, capture_output=True, text=True).stdout
        vgpr = int(notes.split(".vgpr_count:")[1].split()[0])
        spill = int(notes.split(".vgpr_spill_count:")[1].split()[0])
        assert vgpr <= 128 and spill == 0, (step, vgpr, spill)


def test_compile_worker_builds_the_pass_code_object(tmp_path):
    """The compile worker the engine starts (one process per pass, since
    hipRTC serialises the compiles of one process): a generated pass's text on
    stdin gives a gfx950 code object holding lqcs_pass on stdout; text that
    does not compile gives exit status 1 and the compiler's log."""
    worker = os.path.join(ROOT, "litequantumcomputingsimulator_amd", "lib", "lqcs_rtc_worker")
    if not os.path.exists(worker):
        pytest.skip("compile worker not built")
    circ = _u3cx(14, 2, 5)
    text, _ = _passes(circ, 14)[0]
    r = subprocess.run([worker], input=text.encode(), capture_output=True, timeout=600)
    assert r.returncode == 0, r.stderr.decode()[:2000]
    assert r.stdout[:4] == b"\x7fELF"
    co = tmp_path / "pass.co"
    co.write_bytes(r.stdout)
    syms = subprocess.run([READELF, "--syms", str(co)], check=True, capture_output=True, text=True).stdout
    assert "lqcs_pass" in syms
    bad = subprocess.run([worker], input=b"this is not a kernel", capture_output=True, timeout=600)
    assert bad.returncode == 1 and b"error" in bad.stderr

"""CPU check of the register-tile program builder: tests/native/rtile_emulator.hip
builds random tile-op sequences into register programs with the library's own
build_rtile, executes them on the host exactly as the kernel reads them (op
words, per-phase layouts, the LDS slots of both half-tile transpose rounds,
header decoding) and compares with the ops applied directly. No device
needed."""
import os
import shutil
import subprocess

import pytest

from litequantumcomputingsimulator_amd import LIB_DIR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_register_programs_match_direct_application(tmp_path):
    exe = str(tmp_path / "rtile_emulator")
    cmd = [HIPCC, "-std=c++20", "-O2", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "litequantumcomputingsimulator_amd", "csrc", "device"),
           os.path.join(ROOT, "tests", "native", "rtile_emulator.hip"), "-o", exe,
           "-L", LIB_DIR, "-llqcs_device", f"-Wl,-rpath,{LIB_DIR}"]
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([exe, "400"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    line = out.stdout.strip().splitlines()[-1]
    assert line.startswith("ok 400"), line
    fields = dict(kv.split("=") for kv in line.split()[2:])
    assert int(fields["transposes"]) > 1000
    # The layered cases merge singly-controlled X into the neighbouring 2x2
    # ops (select ops, executed and compared like every other op).
    assert int(fields["merged"]) > 500

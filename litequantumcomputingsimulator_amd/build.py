"""In-tree build of the engine.

Products (all under litequantumcomputingsimulator_amd/lib/):
  liblqcs_device.so  HIP kernels for gfx950 + the thin C-ABI (include/lqcs_device.h)
  lqcs_rtc_worker    compile worker the device library starts for specialized passes
  liblqcs.so         C++20 host library: the lqcs:: drop-in API over the C-ABI
  _pylqcs*.so        pybind11 module (the pylqcs Python face), in the package dir

Run `python -m litequantumcomputingsimulator_amd.build` or call build().
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(ROOT, "_build", "obj")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_FLAGS = [
    "-std=c++20", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-Wall",
    # keep hand-written fp32 code off the packed-fp32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]
CXX_FLAGS = ["-std=c++20", "-O3", "-fPIC", "-Wall", "-Wextra", f"-I{INCLUDE}"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def _compile_all(jobs: list[list[str]]) -> None:
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()


GEN = os.path.join(ROOT, "_build", "gen")


def embed_spec_header() -> None:
    """Writes the specialized-pass device templates (spec_pass.hpp) as a raw
    string literal that spec_pass.hip compiles in and hands to hipRTC."""
    os.makedirs(GEN, exist_ok=True)
    with open(os.path.join(CSRC, "device", "spec_pass.hpp")) as f:
        text = f.read()
    if ")LQCSHDR\"" in text:
        raise RuntimeError("spec_pass.hpp contains the raw-string delimiter")
    out = os.path.join(GEN, "spec_pass_hpp.inc")
    body = 'R"LQCSHDR(' + text + ')LQCSHDR"\n'
    if not os.path.exists(out) or open(out).read() != body:
        with open(out, "w") as f:
            f.write(body)


def build_device() -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    embed_spec_header()
    srcs = sorted(glob.glob(os.path.join(CSRC, "device", "*.hip")))
    objs, jobs = [], []
    for s in srcs:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        objs.append(o)
        jobs.append([HIPCC, *HIP_FLAGS, f"-I{GEN}", "-c", s, "-o", o])
    _compile_all(jobs)
    out = os.path.join(LIB, "liblqcs_device.so")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-L/opt/rocm/lib", "-lhiprtc",
          "-Wl,-rpath,/opt/rocm/lib", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_worker() -> str:
    """lqcs_rtc_worker: the specialized passes' compile worker (host code and
    hipRTC only), beside liblqcs_device.so, which starts it."""
    embed_spec_header()
    out = os.path.join(LIB, "lqcs_rtc_worker")
    _run([CXX, "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{GEN}",
          os.path.join(CSRC, "device", "rtc_worker.cpp"), "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib",
          "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_host() -> str:
    srcs = sorted(glob.glob(os.path.join(CSRC, "host", "*.cpp")))
    objs, jobs = [], []
    for s in srcs:
        o = os.path.join(OBJ, "host_" + os.path.basename(s) + ".o")
        objs.append(o)
        jobs.append([CXX, *CXX_FLAGS, "-c", s, "-o", o])
    _compile_all(jobs)
    out = os.path.join(LIB, "liblqcs.so")
    _run([CXX, "-shared", *objs, f"-L{LIB}", "-llqcs_device", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_python() -> str:
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(PKG, "_pylqcs" + suffix)
    src = os.path.join(CSRC, "python", "bindings.cpp")
    py_inc = sysconfig.get_paths()["include"]
    _run([CXX, *CXX_FLAGS, "-shared", f"-I{pybind11.get_include()}", f"-I{py_inc}", src,
          f"-L{LIB}", "-llqcs", "-llqcs_device", "-Wl,-rpath,$ORIGIN/lib", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_oracle() -> None:
    """Checkers: the C restatement and (when the reference sources exist) the
    reference itself. Test infrastructure, never linked into the product."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"])
    if os.path.isdir("/root/reference/src"):
        _run(["sh", os.path.join(ROOT, "oracle", "build_ref.sh")])


def build(oracle: bool = True) -> None:
    build_device()
    build_worker()
    build_host()
    build_python()
    if oracle:
        build_oracle()


if __name__ == "__main__":
    build(oracle="--no-oracle" not in sys.argv)
    print("built", LIB)

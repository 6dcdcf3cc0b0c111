"""MI355X-native statevector engine with the LiteQuantumComputingSimulator API.

The compute path is HIP kernels for gfx950 behind the C-ABI in
include/lqcs_device.h, driven by the C++20 lqcs:: host library; this package
exposes it to Python as ``pylqcs``. Importing fails loudly when the native
libraries are not built: there is no CPU fallback.
"""
from __future__ import annotations

import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_PKG, "lib")


def native_libraries() -> list[str]:
    """Paths of the native libraries this package loads."""
    return [os.path.join(LIB_DIR, "liblqcs_device.so"), os.path.join(LIB_DIR, "liblqcs.so")]


def _load():
    try:
        from . import _pylqcs  # noqa: F401
    except ImportError as e:  # pragma: no cover - exercised only without a build
        raise ImportError(
            "litequantumcomputingsimulator_amd: native engine not built "
            "(run `python -m litequantumcomputingsimulator_amd.build`): " + str(e)
        ) from e
    return _pylqcs

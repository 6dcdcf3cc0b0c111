"""pylqcs: Python face of the engine (the reference's python/pylqcs/__init__.py
re-exports, backed here by the HIP engine), plus a torch autograd bridge."""
from __future__ import annotations

from . import _load

_native = _load()

GateType = _native.GateType
Gate = _native.Gate
QuantumCircuit = _native.QuantumCircuit
Statevector = _native.Statevector
Result = _native.Result
KrausChannel = _native.KrausChannel
NoiseModel = _native.NoiseModel
StatevectorSimulator = _native.StatevectorSimulator
DensityMatrix = _native.DensityMatrix
DensityMatrixSimulator = _native.DensityMatrixSimulator
ParametricCircuit = _native.ParametricCircuit
algorithms = _native.algorithms
io = _native.io
SingleQubitGateFusion = _native.SingleQubitGateFusion
RedundantGateElimination = _native.RedundantGateElimination
device_info = _native.device_info
last_plan_stats = _native.last_plan_stats
MAX_DEVICE_QUBITS = _native.MAX_DEVICE_QUBITS
CompiledCircuit = _native.CompiledCircuit
CircuitGraph = _native.CircuitGraph
statevector = _native.statevector
synchronize = _native.synchronize
profile_enable = _native.profile_enable
profile_read = _native.profile_read
set_tile_kernel = _native.set_tile_kernel
set_dense_mfma_min_targets = _native.set_dense_mfma_min_targets
specialized_kernels_cached = _native.specialized_kernels_cached
profile_read_launches = _native.profile_read_launches
COST_FEATURES = _native.COST_FEATURES
mlp_num_params = _native.mlp_num_params
mlp_forward = _native.mlp_forward
mlp_loss_grad = _native.mlp_loss_grad
mlp_train = _native.mlp_train
costmodel_install = _native.costmodel_install
costmodel_uninstall = _native.costmodel_uninstall
costmodel_load = _native.costmodel_load
last_plan_choice = _native.last_plan_choice

__all__ = [
    "GateType", "Gate", "QuantumCircuit", "Statevector", "Result", "KrausChannel",
    "NoiseModel", "StatevectorSimulator", "DensityMatrix", "DensityMatrixSimulator", "ParametricCircuit", "algorithms", "io", "SingleQubitGateFusion",
    "RedundantGateElimination",
    "device_info", "last_plan_stats", "MAX_DEVICE_QUBITS", "CompiledCircuit",
    "statevector", "synchronize", "profile_enable", "profile_read", "set_tile_kernel", "ExpectationFunction",
    "expectation", "profile_read_launches", "COST_FEATURES", "mlp_num_params", "mlp_forward", "mlp_loss_grad",
    "mlp_train", "costmodel_install", "costmodel_uninstall", "costmodel_load", "last_plan_choice",
    "CircuitGraph", "set_dense_mfma_min_targets", "specialized_kernels_cached",
]

# The packaged cost model (if trained) steers plan="learned".
from . import costmodel as _costmodel  # noqa: E402

_costmodel.install_default(_native)


def _torch():
    import torch

    return torch


class _Expectation:
    """Lazily built torch.autograd.Function: forward = <H>(theta) on the
    engine, backward = adjoint gradient from the same call."""

    _fn = None

    @classmethod
    def get(cls):
        if cls._fn is None:
            torch = _torch()

            class ExpectationFunction(torch.autograd.Function):
                @staticmethod
                def forward(ctx, theta, ansatz, hamiltonian):
                    params = theta.detach().to("cpu", torch.float64).tolist()
                    energy, grad = algorithms.adjoint_gradient(hamiltonian, ansatz, params)
                    ctx.save_for_backward(torch.tensor(grad, dtype=torch.float64, device=theta.device))
                    return torch.tensor(energy, dtype=theta.dtype, device=theta.device)

                @staticmethod
                def backward(ctx, grad_out):
                    (g,) = ctx.saved_tensors
                    return grad_out * g.to(grad_out.dtype), None, None

            cls._fn = ExpectationFunction
        return cls._fn


def ExpectationFunction():
    """The torch.autograd.Function class (built on first use)."""
    return _Expectation.get()


def expectation(theta, ansatz, hamiltonian):
    """Differentiable <psi(theta)|H|psi(theta)> for a ParametricCircuit;
    hamiltonian = [(coeff, pauli), ...].

    When no gradient can flow (theta does not require one, or autograd is
    off), only the forward state is evolved and H read once; otherwise the
    adjoint method returns E and dE/dtheta from one forward and one backward
    sweep, and backward() uses that gradient.

    theta must be a host (CPU) tensor: the engine drives the GPU through the
    system HIP runtime, and the torch wheel carries a HIP runtime of its own;
    whichever of the two opens the device first holds it for the process. A
    theta on another device raises ValueError instead of touching the GPU
    through torch. The result comes back on the CPU in theta's dtype."""
    torch = _torch()
    if theta.device.type != "cpu":
        raise ValueError(
            "expectation: theta must be a CPU tensor (the engine and torch's bundled HIP runtime cannot share "
            "the GPU in one process); pass theta.detach().cpu() from a process where torch has not opened the GPU")
    if not (torch.is_grad_enabled() and theta.requires_grad):
        params = theta.detach().to(torch.float64).tolist()
        sv = StatevectorSimulator().run_statevector(ansatz.bind(params))
        return torch.tensor(algorithms.expectation(sv, hamiltonian), dtype=theta.dtype)
    return _Expectation.get().apply(theta, ansatz, hamiltonian)

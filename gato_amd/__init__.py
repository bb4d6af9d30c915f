"""gato_amd — batched SQP trajectory optimisation (Mo0dy/GATO's hot path) on one AMD Instinct MI355X.

Hand-written CDNA4 HIP kernels behind the C-ABI in include/gato_mi355x.h; PyTorch-ROCm provides
device memory and streams. ``gato_amd.bsqp.bsqpN{N}_{plant}.BSQP_{B}_float`` is the drop-in class.
"""
from .solver import Solver  # noqa: F401

__all__ = ["Solver"]

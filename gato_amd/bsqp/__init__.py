"""Drop-in module surface of the reference's pybind11 extensions ``bsqpN{N}_{plant}``.

The reference compiles one extension module per (knot count, plant), each exposing
``KNOT_POINTS`` and classes ``BSQP_{B}_float`` for B in 1..1024 (python/bindings.cu:212-266), and under USE_DOUBLES
``BSQP_{B}_double`` for B in 1..128 (bindings.cu:244-252); both families are registered here.
Here the modules are created on import of this package and registered under
``gato_amd.bsqp.bsqpN{N}_{plant}``, so ``importlib.import_module(f"gato_amd.bsqp.bsqpN{N}_{plant}")``
and ``getattr(module, f"BSQP_{B}_float")`` work as in python/bsqp/interface.py:44-59.
"""
import sys
import types

from ..solver import Solver, Solver64

KNOTS = (8, 16, 32, 64, 128)
PLANTS = ("iiwa14", "indy7")
BATCHES = (1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024)
BATCHES_DOUBLE = (1, 2, 4, 8, 16, 32, 64, 128)


def _make_class(plant, N, B):
    def __init__(self, *args, **kw):
        Solver.__init__(self, plant, N, B, *args, **kw)

    return type(f"BSQP_{B}_float", (Solver,), {"__init__": __init__, "__module__": f"{__name__}.bsqpN{N}_{plant}",
                                               "__doc__": f"Batched SQP, {plant}, N={N}, batch {B}, float32."})


def _make_class64(plant, N, B):
    def __init__(self, *args, **kw):
        Solver64.__init__(self, plant, N, B, *args, **kw)

    return type(f"BSQP_{B}_double", (Solver64,), {"__init__": __init__, "__module__": f"{__name__}.bsqpN{N}_{plant}",
                                                  "__doc__": f"Batched SQP, {plant}, N={N}, batch {B}, float64."})


def _make_module(plant, N):
    name = f"{__name__}.bsqpN{N}_{plant}"
    mod = types.ModuleType(name, f"bsqpN{N}_{plant}: batched SQP solvers for {plant} with {N} knot points.")
    mod.KNOT_POINTS = N
    for B in BATCHES:
        setattr(mod, f"BSQP_{B}_float", _make_class(plant, N, B))
    for B in BATCHES_DOUBLE:
        setattr(mod, f"BSQP_{B}_double", _make_class64(plant, N, B))
    sys.modules[name] = mod
    return mod


for _p in PLANTS:
    for _n in KNOTS:
        globals()[f"bsqpN{_n}_{_p}"] = _make_module(_p, _n)

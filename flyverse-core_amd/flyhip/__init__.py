"""flyhip: an MI355X (gfx950) HIP backend for the brain step of a connectome-scale fly simulator.

Layers (bottom up): libflyhip.so (hand-written HIP kernels behind a C ABI, include/flyhip.h) -> `hip` (ctypes loader
and validators, the flyverse/cuda.py surface) -> `Brain`, `OpticLobe`, `FlyBrain` (the reference's call surfaces)
and the additions `batch.BatchBrain` (persistent multi-fly stepping) and `adjoint` (optic-lobe adjoint).
"""
from .connectome import Connectome, Retina
from .brain import Brain, LIFParams
from .optic import OpticLobe, OpticParams
from .fly import FlyBrain, StepResult

__all__ = ["Connectome", "Retina", "Brain", "LIFParams", "OpticLobe", "OpticParams", "FlyBrain", "StepResult"]

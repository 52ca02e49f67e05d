"""sgd_amd -- MI355X (gfx950) box-spline footprint CT projectors.

Drop-in for the projector path of hqpwca/SGD: the C ABI of layers/A_*.cu
(include/sgdct.h, libsgdct.so) and the pylops-style operators of
python/operators/operators.py (sgd_amd.operators), plus FP32 device-resident
projectors for torch tensors (sgd_amd.torch_ops).
"""
from ._lib import SgdctError, lib  # noqa: F401

__all__ = ["SgdctError", "lib"]

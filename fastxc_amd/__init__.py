"""FastXC hot path for AMD MI355X (gfx950): sac2spec, xc_fast, ncf_pws, ncf_tfpws.

Device buffers are torch tensors (PyTorch is used for device memory and
streams only); the arithmetic runs in the HIP kernels of libfastxc_amd.so.
"""
from ._lib import FxcError, PRECISIONS, lib  # noqa: F401
from . import ops  # noqa: F401

__all__ = ["FxcError", "PRECISIONS", "lib", "ops"]

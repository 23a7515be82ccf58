"""finchgpu — MI355X (gfx950) execution backend for Finch.jl's compiled sparse kernels.

Layers:
  * ``_lib``   — ctypes binding of libfinchgpu.so (include/finchgpu.h);
  * ``ops``    — device operations on Compressed operands and torch tensors;
  * ``tensors``— Finch's level formats (Dense/SparseList/Element), Tensor,
                 swizzle and device transfer;
  * ``interface`` — @finch/@einsum-style programs routed to the kernels;
  * ``batch``, ``serve`` — the batched persistent form and the serving loop;
  * ``synth``  — seeded synthetic operands.
"""
from . import _lib
from ._lib import (AdmissionError, DimensionMismatch, FinchGPUError, FinchGPUUnavailable,
                   FinchProtocolError, UnsupportedError)

__all__ = ["_lib", "AdmissionError", "DimensionMismatch", "FinchGPUError",
           "FinchGPUUnavailable", "FinchProtocolError", "UnsupportedError"]

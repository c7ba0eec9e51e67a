"""HIP kernels for gfx950 and their PyTorch bindings; the only code that launches kernels.

Same modules and signatures as the reference bindings (dsplat/kernels/*.py of
Asixa/Dirichlet-Splatting): surfel, spectral, search, reduce, response and blas (the
reference's cublas module: the symmetric rank-k update). Every launch goes through
the C ABI of libdsplat_hip.so (include/dsplat_hip.h) on torch's current stream.
Tensors are device float32, complex values complex64, and complex gradients follow
the PyTorch convention dL/dx = Re(conj(upstream) * dy/dx).
"""

import torch


def interleaved(value):
    """Interleaved real/imaginary view [..., 2] of a complex tensor, as the kernels read it
    (kernels/__init__.py:29-31)."""
    return torch.view_as_real(value.to(torch.complex64).resolve_conj().resolve_neg()).contiguous()

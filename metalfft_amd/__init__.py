"""metalfft_amd -- MI355X (gfx950) batched complex FFT behind a C-ABI.

The compute path is libmifft.so (hand-written HIP kernels for gfx950, C-ABI
in include/mifft.h); this package is its Python mirror of MetalFFT's Swift
API.  Importing does not touch the GPU; the first plan creation does.
"""
from ._lib import FFTError, FORWARD, INVERSE  # noqa: F401
from .api import (MetalFFT, MultiChannelFFT, SUPPORTED_SIZES,  # noqa: F401
                  device_name, kernel_variants)

__all__ = ["FFTError", "FORWARD", "INVERSE", "MetalFFT", "MultiChannelFFT",
           "SUPPORTED_SIZES", "device_name", "kernel_variants"]

"""metalfft_amd -- MI355X (gfx950) batched complex FFT behind a C-ABI.

The compute path is libmifft.so (hand-written HIP kernels for gfx950, C-ABI
in include/mifft.h); this package is its Python mirror of MetalFFT's Swift
API.  Importing does not touch the GPU; the first plan creation does.
"""
from ._lib import FFTError, FORWARD, INVERSE  # noqa: F401
from .api import (MetalFFT, MultiChannelFFT, PrecisionFFT, RaggedFFT,  # noqa: F401
                  FOURSTEP_SIZES, RAGGED_DEFAULT, RAGGED_GRAPH, RAGGED_PER_SIZE, RAGGED_QUEUE,
                  RAGGED_SPLIT,
                  SUPPORTED_SIZES,
                  device_name, kernel_variants)
from .dsp import (PSD, STFT, AnalysisResult, Convolver, Correlator,  # noqa: F401
                  CrossSpectral, CrossSpectralResult, FFTAnalyzer, FusedConvolver, Precision,
                  STFTFrame, cross_spectral, to_decibels, window_coefficients)
from .sar import RangeDoppler, RdaParams  # noqa: F401

__all__ = ["FFTError", "FORWARD", "INVERSE", "MetalFFT", "MultiChannelFFT",
           "SUPPORTED_SIZES", "FOURSTEP_SIZES", "device_name", "kernel_variants", "Convolver", "FusedConvolver",
           "Correlator", "Precision", "RaggedFFT", "PrecisionFFT", "RAGGED_GRAPH",
           "RAGGED_PER_SIZE", "RAGGED_DEFAULT", "RAGGED_QUEUE", "RAGGED_SPLIT", "RangeDoppler", "RdaParams", "CrossSpectral",
           "CrossSpectralResult", "cross_spectral", "window_coefficients", "STFT", "STFTFrame",
           "FFTAnalyzer", "AnalysisResult", "PSD", "to_decibels"]

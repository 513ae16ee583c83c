"""sfft_amd -- MI355X-native sparse FFT (gfx950 HIP kernels behind the C ABI of libsfft.so).

Importing loads gpu-sfft_amd/libsfft.so; a missing library raises LibraryNotBuilt (there is
no CPU fallback).  Device arrays are torch tensors on the HIP device.
"""
from ._lib import LIB_PATH, LibraryNotBuilt, lib  # noqa: F401
from .core import (FOLD_ALL_TAPS, FOLD_REFERENCE, NODE_DTYPE, Batch, Plan,  # noqa: F401
                   SfftError, batch_workload, derive_params, dft_any, draw_permutations, fft, fft_plan_passes,
                   filter_taps, filter_window, floor_to_pow2, make_filter, metric,
                   signal_spectrum, sparse_fft, synthesize, to_complex64, truth_spectrum,
                   twiddle, validate)

__version__ = lib.sfft_version().decode()

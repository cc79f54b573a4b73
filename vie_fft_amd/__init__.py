"""MI355X-native (gfx950) VIE-FFT operator and Krylov solvers.

The compute lives in ``libviefft.so`` (hand-written HIP, built by
``vie_fft_amd.build``); :mod:`vie_fft_amd.viefft` is the host-side interface
and :mod:`vie_fft_amd.geometry` builds voxel media.
"""
from . import geometry  # noqa: F401

__all__ = ["geometry", "viefft"]

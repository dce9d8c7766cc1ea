"""obit_amd: MI355X (gfx950) implementation of Obit's visibility-domain hot path.

Two operations, behind Obit's own C entry points (include/ObitCUDA*.h):
the DFT sky-model evaluate/subtract/divide/replace of CLEAN-component
tables (ObitSkyModel / ObitGPUSkyModel) and convolutional uv gridding with
the 7-cell, 100x-oversampled spheroidal kernel (ObitUVGrid / ObitGPUGrid).
The compute is in libobit_hip.so (obit_amd/csrc, hand-written HIP); this
package holds the ctypes binding, the host-side drivers and the synthetic
workload generator.
"""
from . import synth  # noqa: F401  (pure numpy, no native code)

__all__ = ["synth", "load"]


def load():
    """Load the native library (raises if it has not been built)."""
    from . import _lib
    return _lib.load()

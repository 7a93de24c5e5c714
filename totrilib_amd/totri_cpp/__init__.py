"""Drop-in for the reference's compiled module `totri_cpp` (src/totri.cpp:13-54).

Submodules `render`, `reconstruct` and `util` expose the volumetric functions
under the reference's names and positional arguments; each writes into the
output tensor passed by the caller.  They run the HIP kernels of
libtotri_hip.so through its C ABI and never fall back to another
implementation.
"""
from totrilib_amd.totri_cpp import reconstruct, render, util  # noqa: F401

__version__ = "0.1.0"

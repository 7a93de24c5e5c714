"""Steps of the reference's RBF fitting that sit on the backprojection path.

Only the volume-error blob proposal of RbfFitting.add_blobs_volume (reference
totri/fitting/rbf.py:521-565) is provided: the rest of the fitting loop
(losses, optimisation, the mesh renderer) is not part of this package.
"""
from totrilib_amd.fitting.blobs import sample_blob_positions, volume_error_map

__all__ = ["volume_error_map", "sample_blob_positions"]

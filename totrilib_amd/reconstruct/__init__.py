"""totri.reconstruct -- backprojection reconstructions (f-k migration of the
reference is not part of this library)."""
from totrilib_amd.reconstruct.backprojection import (  # noqa: F401
    BackprojectionConfocalVelten,
    BackprojectionExhaustiveArellano,
    BackprojectionExhaustiveVelten,
    second_derivative_filter,
)

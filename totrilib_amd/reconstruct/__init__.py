"""totri.reconstruct -- backprojection reconstructions and f-k migration."""
from totrilib_amd.reconstruct.backprojection import (  # noqa: F401
    BackprojectionConfocalVelten,
    BackprojectionExhaustiveArellano,
    BackprojectionExhaustiveVelten,
    second_derivative_filter,
)
from totrilib_amd.reconstruct.fk_migration import FkMigration, StoltInterpolation  # noqa: F401

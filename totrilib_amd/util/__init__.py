"""totri.util -- the splat and the wall-point helpers of the hot path
(the reference's video, logging and visualisation helpers are not part of
this library)."""
from totrilib_amd.util.format import origin_of_wallpoints  # noqa: F401
from totrilib_amd.util.rbf2volume import Rbf2VolumeGaussian  # noqa: F401

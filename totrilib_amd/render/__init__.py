"""totri.render -- volume transient renderers (the OptiX mesh renderers of the
reference are not part of this library)."""
from totrilib_amd.render.volume import VolumeRenderConfocal  # noqa: F401

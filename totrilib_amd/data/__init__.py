"""totri.data -- only the scan-grid helper lives here; dataset loaders of the
reference are outside this library."""
from totrilib_amd.data.util import make_wall_grid  # noqa: F401

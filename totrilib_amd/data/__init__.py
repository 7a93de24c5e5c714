"""totri.data: the scan-grid helper and the loaders of the .mat and PFM
capture formats.  The HDF5 loaders of the reference (Zaragoza, LindellFk2019)
need h5py, which this environment does not provide."""
from totrilib_amd.data.otooleconfocal2018 import (OTooleConfocal2018,  # noqa: F401
                                                   OTooleConfocal2018Dataset)
from totrilib_amd.data.pfm import PfmTransient  # noqa: F401
from totrilib_amd.data.util import make_wall_grid  # noqa: F401
from totrilib_amd.data.veltennc2012 import VeltenNc2012, VeltenNc2012Dataset  # noqa: F401

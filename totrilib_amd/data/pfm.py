"""PFM transient reader, reference totri/data/pfm.py:4-20."""
from __future__ import annotations

import numpy as np
import torch


def PfmTransient(file_path: str, shape: tuple) -> torch.Tensor:
    """Read a single-channel PFM file ('Pf' header, a size line, scale -1 for
    little-endian) and return its float32 samples viewed as `shape`.  The
    size line is not used; the caller gives the shape."""
    with open(file_path, 'rb') as f:
        if not f.readline().decode().strip().lower().startswith('pf'):
            raise IOError(f'File "{file_path}" is no valid Pfm file.')
        f.readline()  # width height
        if float(f.readline().decode().strip()) != -1:
            raise IOError(f'Line 3 of "{file_path}" must be -1')
        samples = np.fromfile(f, dtype='<f4')
    return torch.from_numpy(samples.astype(np.float32)).view(shape)

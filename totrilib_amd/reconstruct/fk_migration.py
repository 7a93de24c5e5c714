"""f-k migration (Lindell et al. 2019) on the GPU.

Mirrors totri/reconstruct/fk_migration.py of the reference: the same
StoltInterpolation and FkMigration modules, constructor arguments, input
layouts and return shapes.  The reference is pure PyTorch; here the padding,
the Stolt interpolation (with both fftshifts folded into its indexing) and
the magnitude run as HIP kernels of libtotri_hip.so.  For power-of-two
sizes the 3-D FFTs are the library's own LDS kernels; for other sizes they
are torch.fft calls (hipFFT; the forward one real-to-complex).
"""
from __future__ import annotations

import torch

from totrilib_amd import _lib
from totrilib_amd.totri_cpp._common import stream_ptr


def _gpu_device(t: torch.Tensor) -> torch.device:
    return t.device if t.is_cuda else torch.device('cuda', torch.cuda.current_device())


def _stolt(spectrum: torch.Tensor, out: torch.Tensor, depth: float, height: float,
           width: float, layout: int, out_half_zeroed: bool) -> None:
    C, D2, H = (int(x) for x in spectrum.shape[:3])
    W = int(out.shape[-1])
    lib = _lib.load()
    rc = lib.totri_fk_stolt(spectrum.data_ptr(), C, D2, H, W, float(depth), float(height),
                            float(width), layout, int(out_half_zeroed), out.data_ptr(),
                            stream_ptr(spectrum.device))
    _lib.check(rc, 'fk_stolt')


class StoltInterpolation(torch.nn.Module):
    """Stolt interpolation of an fftshift-ed spectrum (reference
    fk_migration.py:5-85).

    depth:  range, number of bins times bin width [m]
    height: height of the scanned wall area [m]
    width:  width of the scanned wall area [m]

    The reference caches the sample grid and weights as [1, 2D, H, W, 3] and
    [1, 2D, H, W] tensors; the kernel computes them per element instead.
    get_buffers returns the same tensors for callers that read them.
    """

    def __init__(self, depth, height, width):
        super().__init__()
        self.depth = depth
        self.height = height
        self.width = width

    def get_buffers(self, D, H, W, device):
        """grid [1, 2D, H, W, 3] and weights [1, 2D, H, W] of the interpolation
        (fk_migration.py:21-54)."""
        x = torch.linspace(-1, 1, W, device=device).view(1, 1, W).expand((2 * D, H, W))
        y = torch.linspace(-1, 1, H, device=device).view(1, H, 1).expand((2 * D, H, W))
        z = torch.linspace(-1, 1, 2 * D, device=device).view(2 * D, 1, 1).expand((2 * D, H, W))
        zz = torch.sqrt(x ** 2 * ((W * self.depth) / (2 * D * self.width)) ** 2
                        + y ** 2 * ((H * self.depth) / (2 * D * self.height)) ** 2 + z ** 2)
        grid = torch.stack((x, y, zz), dim=3).unsqueeze(0)
        weights = ((z > 0) * torch.abs(z) / torch.clamp(zz, min=1e-6)).unsqueeze(0)
        return [grid, weights]

    def forward(self, fourier: torch.Tensor) -> torch.Tensor:
        """fourier: complex [N, C, 2D, H, W] in fftshift order -> the
        interpolated, weighted spectrum of the same shape and order."""
        if not torch.cuda.is_available():
            raise RuntimeError('StoltInterpolation needs a GPU, but no devices are available.')
        if fourier.dim() != 5 or not fourier.is_complex():
            raise ValueError(f'expected a complex [N, C, 2D, H, W] tensor, got '
                             f'{fourier.dtype} {tuple(fourier.shape)}')
        N, C, D2, H, W = fourier.shape
        if D2 % 2 != 0:
            raise ValueError(f'the depth size {D2} must be even')
        dev = _gpu_device(fourier)
        f = fourier.to(dev, torch.complex64).contiguous()
        out = torch.empty_like(f)
        with torch.cuda.device(dev):
            _stolt(f.view(N * C, D2, H, W), out, self.depth, self.height, self.width,
                   _lib.TOTRI_FK_SHIFTED, False)
        return out.to(fourier.device)


class FkMigration(torch.nn.Module):
    """f-k migration of a confocal transient (reference fk_migration.py:87-120).

    forward(transient [N, D, H, W]) -> volume [1, D, H, W] of the first batch
    entry, as the reference returns `volume[0]` of its [N, 1, D, H, W] result;
    the other entries do not reach the output and are not computed.

    Power-of-two sizes (H * W <= 16384, 2D <= 4096) run the library's own FFTs
    (totri_fk_migration: five kernels, no zero padding read or written).
    Other sizes run the pre-pass, Stolt and magnitude kernels around
    torch.fft (hipFFT): the padded input is real, so the forward FFT is
    one-sided and the Stolt kernel reads the other half through conjugate
    symmetry.  Scratch memory is kept between calls of the same shape on the
    same stream (the reference keeps its grid and weights in the same way).
    """

    def __init__(self, depth, height, width):
        super().__init__()
        self.stolt_interpolation = StoltInterpolation(depth, height, width)
        self._fk_key = None
        self._fk_ws = None

    def _workspace(self, key, make):
        if self._fk_key != key:
            self._fk_ws = None
            self._fk_ws = make()
            self._fk_key = key
        return self._fk_ws

    def forward(self, transient: torch.Tensor) -> torch.Tensor:
        if not torch.cuda.is_available():
            raise RuntimeError('FkMigration needs a GPU, but no devices are available.')
        if transient.dim() != 4:
            raise ValueError(f'expected a [N, D, H, W] transient, got {tuple(transient.shape)}')
        N, D, H, W = (int(s) for s in transient.shape)
        if N == 0:
            raise IndexError('FkMigration returns the first batch entry; the batch is empty')
        if D * H * W == 0:
            raise ValueError(f'FkMigration needs a non-empty transient, got {tuple(transient.shape)}')
        dev = _gpu_device(transient)
        t = transient[0].to(dev, torch.float32).contiguous()
        lib = _lib.load()
        with torch.cuda.device(dev):
            st = stream_ptr(dev)
            out = torch.empty((1, D, H, W), dtype=torch.float32, device=dev)
            nbytes = int(lib.totri_fk_migration_workspace(D, H, W))
            if nbytes:
                self._native(lib, t, out, nbytes, dev, st)
            else:
                self._torch_fft(lib, t, out, dev, st)
        return out.to(transient.device)

    def _native(self, lib, t, out, nbytes, dev, st):
        D, H, W = (int(x) for x in t.shape)
        s = self.stolt_interpolation
        ws = self._workspace(('native', dev, st, D, H, W),
                             lambda: torch.empty(nbytes, dtype=torch.uint8, device=dev))
        _lib.check(lib.totri_fk_migration(t.data_ptr(), D, H, W, float(s.depth),
                                          float(s.height), float(s.width), out.data_ptr(),
                                          ws.data_ptr(), nbytes, st), 'fk_migration')

    def _torch_fft(self, lib, t, out, dev, st):
        D, H, W = (int(x) for x in t.shape)
        s = self.stolt_interpolation
        padded, stolt_out = self._workspace(
            ('torch_fft', dev, st, D, H, W),
            lambda: (torch.zeros((2 * D, H, W), dtype=torch.float32, device=dev),
                     torch.zeros((2 * D, H, W), dtype=torch.complex64, device=dev)))
        # both buffers keep their zero halves: only rows d < D of the padded
        # input, and the z > 0 half of the Stolt output, are written
        _lib.check(lib.totri_fk_prepare(t.data_ptr(), D, H, W, padded.data_ptr(), 1, st),
                   'fk_prepare')
        spectrum = torch.fft.rfftn(padded).contiguous()          # [2D, H, W/2+1]
        _stolt(spectrum.view(1, 2 * D, H, W // 2 + 1), stolt_out, s.depth, s.height,
               s.width, _lib.TOTRI_FK_UNSHIFTED_ONESIDED, True)
        del spectrum
        volume = torch.fft.ifftn(stolt_out).contiguous()
        _lib.check(lib.totri_fk_magnitude(volume.data_ptr(), 1, D, H, W, out.data_ptr(), st),
                   'fk_magnitude')

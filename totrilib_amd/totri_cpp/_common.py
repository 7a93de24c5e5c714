"""Shared argument handling for the totri_cpp-compatible binding."""
from __future__ import annotations

import torch

from totrilib_amd import _lib

_DTYPE_CODE = {
    torch.float32: _lib.TOTRI_F32,
    torch.bfloat16: _lib.TOTRI_BF16,
    torch.float16: _lib.TOTRI_F16,
}


def require_gpu(*tensors: torch.Tensor, names=None) -> torch.device:
    """All tensors must live on one HIP device; returns it."""
    dev = None
    for k, t in enumerate(tensors):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"argument {k} must be a torch.Tensor, got {type(t)}")
        if t.device.type != "cuda":
            name = names[k] if names else f"argument {k}"
            raise NotImplementedError(
                f"{name} is on {t.device}; the totri HIP kernels need GPU tensors")
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise ValueError(f"tensors on different devices: {dev} and {t.device}")
    return dev


def require_f32(t: torch.Tensor, name: str) -> None:
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")


def require_f64(t: torch.Tensor, name: str) -> None:
    if t.dtype != torch.float64:
        raise TypeError(f"{name} must be float64 like the volume, got {t.dtype}")


def storage_code(t: torch.Tensor, name: str) -> int:
    try:
        return _DTYPE_CODE[t.dtype]
    except KeyError:
        raise TypeError(
            f"{name} must be float32, bfloat16 or float16, got {t.dtype}") from None


def stream_ptr(dev: torch.device) -> int:
    return torch.cuda.current_stream(dev).cuda_stream


def workspace(nbytes: int, dev: torch.device) -> tuple[torch.Tensor | None, int]:
    if nbytes <= 0:
        return None, 0
    buf = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
    return buf, buf.data_ptr()


class OutputSlot:
    """A contiguous view of an output tensor; copies back on finish() if the
    caller's tensor is not contiguous (the reference writes through strides)."""

    def __init__(self, t: torch.Tensor):
        self.t = t
        self.buf = t if t.is_contiguous() else torch.empty_like(t, memory_format=torch.contiguous_format)

    def ptr(self) -> int:
        return self.buf.data_ptr()

    def finish(self) -> None:
        if self.buf is not self.t:
            self.t.copy_(self.buf)

"""Fixed-budget memory arenas for the serving engine.

`Arena` is the bookkeeping: a first-fit allocator over [0, capacity) with
aligned block sizes and coalescing frees.  It touches no memory, so the
admission policy can be tested without a GPU.  `TensorArena` backs an `Arena`
with one torch allocation (device memory, or pinned host memory) made once,
so that serving never calls the device allocator while requests run.
"""
from __future__ import annotations

import bisect
import math
from typing import Dict, List, Optional, Sequence, Tuple

import torch


class Arena:
    """First-fit allocator over [0, capacity) in `align`-byte granules."""

    def __init__(self, capacity: int, align: int = 256):
        if capacity < 0 or align <= 0 or align & (align - 1):
            raise ValueError("capacity must be >= 0 and align a power of two")
        self.capacity = int(capacity) // align * align
        self.align = align
        self._free_off: List[int] = [0] if self.capacity else []
        self._free_len: List[int] = [self.capacity] if self.capacity else []
        self._used: Dict[int, int] = {}

    def _round(self, nbytes: int) -> int:
        return max(self.align, -(-int(nbytes) // self.align) * self.align)

    def alloc(self, nbytes: int) -> Optional[int]:
        """Offset of a new block of at least nbytes, or None if none fits."""
        if nbytes < 0:
            raise ValueError("negative size")
        size = self._round(nbytes)
        for k, length in enumerate(self._free_len):
            if length >= size:
                off = self._free_off[k]
                if length == size:
                    del self._free_off[k]
                    del self._free_len[k]
                else:
                    self._free_off[k] = off + size
                    self._free_len[k] = length - size
                self._used[off] = size
                return off
        return None

    def alloc_many(self, sizes: Sequence[int]) -> Optional[List[int]]:
        """All blocks or none."""
        got: List[int] = []
        for s in sizes:
            off = self.alloc(s)
            if off is None:
                for o in got:
                    self.free(o)
                return None
            got.append(off)
        return got

    def free(self, off: int) -> None:
        size = self._used.pop(off)
        k = bisect.bisect_left(self._free_off, off)
        self._free_off.insert(k, off)
        self._free_len.insert(k, size)
        # merge with the next block, then with the previous one
        if k + 1 < len(self._free_off) and off + size == self._free_off[k + 1]:
            self._free_len[k] += self._free_len[k + 1]
            del self._free_off[k + 1]
            del self._free_len[k + 1]
        if k > 0 and self._free_off[k - 1] + self._free_len[k - 1] == off:
            self._free_len[k - 1] += self._free_len[k]
            del self._free_off[k]
            del self._free_len[k]

    def fits(self, sizes: Sequence[int]) -> bool:
        """Whether the blocks could be allocated in an empty arena."""
        return sum(self._round(s) for s in sizes) <= self.capacity

    @property
    def used(self) -> int:
        return sum(self._used.values())

    @property
    def largest_free(self) -> int:
        return max(self._free_len, default=0)

    def blocks(self) -> List[Tuple[int, int]]:
        return sorted(self._used.items())


class TensorArena:
    """An `Arena` over one uint8 tensor; blocks are handed out as typed views."""

    def __init__(self, capacity: int, device: torch.device, pinned: bool = False,
                 align: int = 256):
        self.arena = Arena(capacity, align)
        if pinned:
            self.buf = torch.empty(self.arena.capacity, dtype=torch.uint8, pin_memory=True)
        else:
            self.buf = torch.empty(self.arena.capacity, dtype=torch.uint8, device=device)

    def view(self, off: int, shape: Sequence[int], dtype=torch.float32) -> torch.Tensor:
        n = math.prod(shape) * torch.empty((), dtype=dtype).element_size()
        return self.buf[off:off + n].view(dtype).view(tuple(shape))

    def ptr(self, off: int) -> int:
        return self.buf.data_ptr() + off

"""GPU checks of the library's device math helpers against the compiler's
correctly rounded functions."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_short_sqrt_equals_sqrtf_on_its_whole_range():
    """The short correctly rounded square root that the exhaustive
    backprojection, render and adjoint take (where every radicand lies in
    [2^-64, 2^64]) returns the same bits as sqrtf for every float of that
    range: all 1.07e9 bit patterns from 2^-64 to 2^64."""
    from totrilib_amd import _lib
    lib = _lib.load()
    dev = torch.device("cuda:0")
    out = torch.empty(1, dtype=torch.int64, device=dev)
    lo, hi = 0x1F800000, 0x5F800000  # 2^-64 and 2^64
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.totri_debug_sqrt_mid_mismatches(lo, hi, out.data_ptr(), st),
               "debug_sqrt_mid_mismatches")
    torch.cuda.synchronize()
    assert int(out.item()) == 0

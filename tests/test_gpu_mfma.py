"""Lane-layout check of the gfx950 MFMA builtins the MLP kernels are built on."""
import pytest
import torch

from srf_hip import _lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('prec,k,dtype', [(0, 4, None), (1, 32, torch.bfloat16), (2, 32, torch.float16)])
def test_mfma_tile_layout(cuda, prec, k, dtype):
    g = torch.Generator().manual_seed(prec)
    a = torch.randn(16, k, generator=g)
    b = torch.randn(k, 16, generator=g)
    if dtype is not None:
        a = a.to(dtype).float()
        b = b.to(dtype).float()
    d = torch.empty(16, 16, device=cuda)
    ad, bd = a.to(cuda), b.to(cuda)
    _lib.call('srf_selftest_mfma', prec, ad.data_ptr(), bd.data_ptr(), d.data_ptr(), _lib.stream_ptr(cuda))
    torch.cuda.synchronize()
    ref = a.double() @ b.double()
    assert torch.allclose(d.cpu().double(), ref, atol=1e-4, rtol=1e-5), (d.cpu() - ref).abs().max()


def test_lds_dma_fragment_copy(cuda):
    src = torch.randint(0, 256, (8 * 1024,), dtype=torch.uint8, generator=torch.Generator().manual_seed(0))
    dst = torch.zeros(8 * 1024, dtype=torch.uint8, device=cuda)
    s = src.to(cuda)
    _lib.call('srf_selftest_lds_dma', s.data_ptr(), dst.data_ptr(), _lib.stream_ptr(cuda))
    torch.cuda.synchronize()
    assert torch.equal(dst.cpu().view(8, 1024), src.view(8, 1024).flip(0))

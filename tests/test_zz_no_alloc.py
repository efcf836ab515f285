"""The allocation-free compaction entry point (srf_threshold_compact_noalloc, tile counts in the
library's fixed device array) against srf_threshold_compact (caller's workspace) through the C-ABI:
bit-exact indices and counts on consecutive calls (the second smaller, so stale counts of the first
call sit in the array), the empty and single-entry cases, the largest input the array takes, and the
host-side refusal of larger inputs."""
import pytest
import torch

from srf_hip import _lib, tensorf

pytestmark = pytest.mark.gpu

SRF_ETOOLARGE = 4
NOALLOC_MAX = 1 << 28


def _noalloc(v: torch.Tensor, thr: float, idx: torch.Tensor, count: torch.Tensor, n: int | None = None) -> int:
    n = v.numel() if n is None else n
    return _lib.lib().srf_threshold_compact_noalloc(n, _lib.ptr(v), float(thr), _lib.ptr(idx), _lib.ptr(count),
                                                    _lib.stream_ptr(v.device))


def _check_against_workspace_entry(v: torch.Tensor, thr: float):
    n = v.numel()
    print(f'noalloc compaction: n={n} threshold={thr}', flush=True)
    idx_ref, cnt_ref = tensorf.threshold_compact(v, thr)
    idx = torch.full((max(n, 1),), -1, dtype=torch.int32, device=v.device)
    cnt = torch.full((1,), -1, dtype=torch.int32, device=v.device)
    assert _noalloc(v, thr, idx, cnt) == 0
    torch.cuda.synchronize()
    c = int(cnt.item())
    assert c == int(cnt_ref.item())
    assert torch.equal(idx[:c], idx_ref[:c])
    return idx[:c], c


def test_noalloc_matches_workspace_entry_on_consecutive_calls(cuda):
    g = torch.Generator().manual_seed(5)
    for n, thr in ((5_000_003, 0.7), (70_001, 0.2), (1024 * 37, 0.5)):
        v = torch.rand(n, generator=g).to(cuda)
        idx, c = _check_against_workspace_entry(v, thr)
        # the reference's boolean indexing of weights > threshold
        assert torch.equal(idx.long().cpu(), torch.nonzero(v.cpu() > thr).flatten())


def test_noalloc_empty_single_and_none_selected(cuda):
    v = torch.empty(0, device=cuda)
    _check_against_workspace_entry(v, 0.5)
    _check_against_workspace_entry(torch.tensor([0.9], device=cuda), 0.5)
    _, c = _check_against_workspace_entry(torch.zeros(3000, device=cuda), 0.5)
    assert c == 0


def test_noalloc_largest_input(cuda):
    """n = 2^28 fills every tile count of the fixed array."""
    v = torch.rand(NOALLOC_MAX, device=cuda, generator=torch.Generator(device=cuda).manual_seed(3))
    _, c = _check_against_workspace_entry(v, 0.999)
    assert 0 < c < NOALLOC_MAX


def test_noalloc_refuses_inputs_beyond_its_scratch(cuda):
    """Refused on the host before any launch."""
    v = torch.zeros(4, device=cuda)
    idx = torch.zeros(4, dtype=torch.int32, device=cuda)
    cnt = torch.zeros(1, dtype=torch.int32, device=cuda)
    assert _noalloc(v, 0.5, idx, cnt, n=NOALLOC_MAX + 1) == SRF_ETOOLARGE
    assert 'allocation-free' in _lib.lib().srf_status_string(SRF_ETOOLARGE).decode()
    with pytest.raises(_lib.SrfError, match='allocation-free'):
        _lib.call('srf_threshold_compact_noalloc', NOALLOC_MAX + 1, _lib.ptr(v), 0.5, _lib.ptr(idx), _lib.ptr(cnt),
                  _lib.stream_ptr(v.device))

"""GPU parity of the CP-decomposition kernels (density and appearance gathers, forward and
backward) against the oracle (torch grid_sample on the CPU in FP32 and FP64)."""
import pytest
import torch

import oracle.tensorf as ot
from srf_hip import _lib, tensorf

pytestmark = pytest.mark.gpu

RES = [40, 36, 30]
BBOX = torch.tensor([[-1.5, -1.0, -0.5], [1.5, 1.2, 1.0]])


def cp_lines(rank, res, seed, scale=0.3):
    """Lines [1,R,res[v_i],1], v = (2, 1, 0), as created at SimpleTensoRF09.py:989-996 (the scale
    is raised from the reference's 0.1 so the triple products are not tiny)."""
    g = torch.Generator().manual_seed(seed)
    return [scale * torch.randn(1, rank, res[ot.VECTOR_AXES[i]], 1, generator=g) for i in range(3)]


def points(n, seed, outside=0.2):
    g = torch.Generator().manual_seed(seed)
    lo, hi = BBOX[0], BBOX[1]
    pad = (hi - lo) * outside
    return lo - pad + (hi - lo + 2 * pad) * torch.rand(n, 3, generator=g)


def to_dev(ts, dev, texel_major):
    if texel_major:
        return [tensorf.texel_major_param(t.shape[1], t.shape[2], t.shape[3], t, device=dev) for t in ts]
    return [torch.nn.Parameter(t.to(dev)) for t in ts]


def oracle_density(lines, pts, alpha=None, kind='ReLU', offset=-10.0, dtype=torch.float64):
    L = [l.to(dtype) for l in lines]
    pts = pts.to(dtype)
    bb = BBOX.to(dtype)
    valid = ot.bbox_validity(pts, bb)
    if alpha is not None:
        a = ot.alpha_mask_sample(alpha[0].to(dtype), alpha[1].to(dtype), pts[valid])
        valid[valid.clone()] &= a > 0
    feat = torch.zeros(pts.shape[0], dtype=dtype)
    feat[valid] = ot.cp_density_features(L, ot.normalize_points(pts, bb)[valid])
    sigma = torch.zeros_like(feat)
    sigma[valid] = ot.density_predictor(feat[valid], kind, offset)
    return sigma, valid


@pytest.mark.parametrize('rank,texel_major', [(4, True), (16, False), (48, True), (96, True), (256, True)])
def test_cp_density_forward(cuda, rank, texel_major):
    lines = cp_lines(rank, RES, 0)
    pts = points(20000, 1)
    s64, v64 = oracle_density(lines, pts)
    s32, _ = oracle_density(lines, pts, dtype=torch.float32)
    L = to_dev(lines, cuda, texel_major)
    sig, valid = tensorf.cp_density(pts.to(cuda), L, BBOX.flatten().tolist())
    torch.cuda.synchronize()
    assert torch.equal(valid.cpu().bool(), v64)
    # FP32 parity bar 1e-5 abs; the tolerance scales with the feature magnitude for the largest ranks
    tol = 1e-5 * max(1.0, s64.abs().max().item())
    assert (sig.cpu().double() - s64).abs().max() < tol
    assert (sig.cpu() - s32).abs().max() < tol


def test_cp_density_alpha_mask_and_softplus(cuda):
    lines = cp_lines(16, RES, 2)
    g = torch.Generator().manual_seed(3)
    vol = (torch.rand(16, 16, 16, generator=g) < 0.4).float()
    abox = torch.tensor([[-1.4, -0.9, -0.4], [1.4, 1.1, 0.9]])
    pts = points(30000, 4)
    s64, v64 = oracle_density(lines, pts, alpha=(vol, abox), kind='SoftPlus', offset=-2.0)
    mask = tensorf.AlphaMask(vol.to(cuda), abox.flatten().tolist())
    sig, valid = tensorf.cp_density(pts.to(cuda), to_dev(lines, cuda, True), BBOX.flatten().tolist(), alpha=mask,
                                    predictor=tensorf.SOFTPLUS, offset=-2.0)
    torch.cuda.synchronize()
    assert (valid.cpu().bool() != v64).sum().item() == 0
    assert (sig.cpu().double() - s64).abs().max() < 1e-5


@pytest.mark.parametrize('rank,kind', [(16, 'ReLU'), (8, 'SoftPlus'), (64, 'ReLU')])
def test_cp_density_backward_matches_autograd(cuda, rank, kind):
    lines = cp_lines(rank, RES, 5)
    pts = points(8000, 6)
    g = torch.Generator().manual_seed(7)
    gs = torch.randn(8000, generator=g)
    L64 = [l.double().requires_grad_(True) for l in lines]
    s64, _ = oracle_density(L64, pts, kind=kind, offset=-1.0)
    (s64 * gs.double()).sum().backward()
    L = to_dev(lines, cuda, True)
    sig, _ = tensorf.cp_density(pts.to(cuda), L, BBOX.flatten().tolist(), predictor=tensorf.predictor_code(kind),
                                offset=-1.0)
    (sig * gs.to(cuda)).sum().backward()
    torch.cuda.synchronize()
    for got, ref in zip(L, L64):
        assert got.grad.shape == ref.grad.shape
        err = (got.grad.cpu().double() - ref.grad).abs().max() / ref.grad.abs().max()
        assert err < 1e-4, err


@pytest.mark.parametrize('rank,texel_major', [(48, True), (96, False)])
def test_cp_appearance_forward_backward(cuda, rank, texel_major):
    lines = cp_lines(rank, RES, 8)
    pts = points(6000, 9, outside=0.0)
    g = torch.Generator().manual_seed(10)
    idx = torch.nonzero(torch.rand(6000, generator=g) < 0.3).flatten().to(torch.int32)
    L64 = [l.double().requires_grad_(True) for l in lines]
    ref = ot.cp_appearance_features(L64, ot.normalize_points(pts.double(), BBOX.double())[idx.long()])
    gout = torch.randn(ref.shape, generator=g)
    (ref * gout.double()).sum().backward()
    L = to_dev(lines, cuda, texel_major)
    got = tensorf.cp_appearance(pts.to(cuda), L, BBOX.flatten().tolist(), idx=idx.to(cuda))
    (got * gout.to(cuda)).sum().backward()
    torch.cuda.synchronize()
    assert got.shape == ref.shape
    assert (got.detach().cpu().double() - ref.detach()).abs().max() < 1e-6
    for a, b in zip(L, L64):
        err = (a.grad.cpu().double() - b.grad).abs().max() / b.grad.abs().max()
        assert err < 1e-4, err


def test_cp_all_samples_without_index_and_empty_input(cuda):
    lines = cp_lines(16, RES, 11)
    pts = points(3000, 12, outside=0.0)
    L = to_dev(lines, cuda, True)
    got = tensorf.cp_appearance(pts.to(cuda), L, BBOX.flatten().tolist())
    ref = ot.cp_appearance_features([l.double() for l in lines], ot.normalize_points(pts.double(), BBOX.double()))
    assert (got.detach().cpu().double() - ref).abs().max() < 1e-6
    empty = tensorf.cp_appearance(torch.zeros(0, 3, device=cuda), L, BBOX.flatten().tolist())
    assert empty.shape == (0, 16)
    sig, valid = tensorf.cp_density(torch.zeros(0, 3, device=cuda), L, BBOX.flatten().tolist())
    assert sig.shape == (0,) and valid.shape == (0,)


def test_cp_rejects_unsupported_ranks(cuda):
    pts = points(10, 13).to(cuda)
    with pytest.raises(_lib.SrfError):
        tensorf.cp_density(pts, to_dev(cp_lines(6, RES, 14), cuda, True), BBOX.flatten().tolist())
    with pytest.raises(_lib.SrfError):
        tensorf.cp_density(pts, to_dev(cp_lines(260, RES, 15), cuda, True), BBOX.flatten().tolist())
    mixed = to_dev(cp_lines(8, RES, 16), cuda, True)[:2] + to_dev(cp_lines(12, RES, 17), cuda, True)[2:]
    with pytest.raises(_lib.SrfError):
        tensorf.cp_appearance(pts, mixed, BBOX.flatten().tolist())

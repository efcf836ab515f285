"""srf_rays (rays, view directions and the NDC warp in one kernel) against the oracle's get_rays /
get_view_dirs / get_ndc_rays in FP64 and against the PyTorch restatement in FP32 (the path a model
that learns its camera differentiates), on random cameras with rotation, translation, skew and an
off-centre principal point."""
import pytest
import torch

import oracle.nerf as on
from srf_hip import geometry

pytestmark = pytest.mark.gpu

W, H, NEAR = 504, 378, 1.0


def row_rel(ours, theirs):
    """Largest per-ray error relative to that ray's largest component (FP32 rounding of components
    near zero is judged against the vector's magnitude, not its own)."""
    ours, theirs = ours.double().cpu(), theirs.double().cpu()
    scale = theirs.abs().max(dim=1, keepdim=True).values.clamp_min(1e-12)
    return ((ours - theirs).abs() / scale).max().item()


def cameras(n_frames, seed):
    g = torch.Generator().manual_seed(seed)
    K = torch.zeros(n_frames, 3, 3)
    E = torch.zeros(n_frames, 4, 4)
    for f in range(n_frames):
        fx = W * (0.6 + 0.4 * torch.rand(1, generator=g).item())
        fy = fx * (0.9 + 0.2 * torch.rand(1, generator=g).item())
        s = 0.01 * torch.randn(1, generator=g).item()
        cx = W / 2 + 10 * torch.randn(1, generator=g).item()
        cy = H / 2 + 10 * torch.randn(1, generator=g).item()
        K[f] = torch.tensor([[fx, s, cx], [0., fy, cy], [0., 0., 1.]])
        # a rotation within ~10 degrees of identity (forward-facing) and a small translation
        a = 0.1 * torch.randn(3, generator=g)
        A = torch.tensor([[0., -a[2], a[1]], [a[2], 0., -a[0]], [-a[1], a[0], 0.]])
        E[f, :3, :3] = torch.linalg.matrix_exp(A)
        E[f, :3, 3] = 0.2 * torch.randn(3, generator=g)
        E[f, 3, 3] = 1.
    return K, E


def pixel_ids(n, n_frames, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.stack([torch.randint(0, n_frames, (n,), generator=g), torch.randint(0, W, (n,), generator=g),
                        torch.randint(0, H, (n,), generator=g)], 1).int()


@pytest.mark.parametrize('mip', [False, True])
@pytest.mark.parametrize('ndc', [False, True])
def test_rays_match_oracle_and_torch(cuda, mip, ndc):
    n, F = 6000, 4
    K, E = cameras(F, seed=3)
    pid = pixel_ids(n, F, seed=4)
    ro, rd, vd, on_o, on_d = geometry.rays_fused(pid.to(cuda), K.to(cuda), E.to(cuda), mip_nerf_used=mip, ndc=ndc,
                                                 resolution=(H, W), near=NEAR)
    fr = pid[:, 0].long()
    # FP64 oracle
    r_o, r_d = on.get_rays(pid, K[fr].double(), E[fr].double(), mip_nerf_used=mip)
    r_v = on.get_view_dirs(r_d)
    ref = [(ro, r_o), (rd, r_d), (vd, r_v)]
    if ndc:
        n_o, n_d = on.get_ndc_rays(r_o, r_d, (H, W), K[fr].double(), NEAR)
        ref += [(on_o, n_o), (on_d, n_d)]
    else:
        assert on_o is None and on_d is None
    for ours, theirs in ref:
        err = row_rel(ours, theirs)
        assert err < 2e-6, err
    # FP32 PyTorch restatement on the same GPU (LU inverse of the intrinsics): ulp-level agreement
    t_o, t_d = geometry.get_rays_tr(pid.to(cuda), K[fr].to(cuda), E[fr].to(cuda), mip_nerf_used=mip)
    pairs = [(ro, t_o), (rd, t_d), (vd, geometry.get_view_dirs_tr(t_d))]
    if ndc:
        pairs += list(zip((on_o, on_d), geometry.get_ndc_rays_tr(t_o, t_d, (H, W), K[fr].to(cuda), near=NEAR)))
    for ours, theirs in pairs:
        err = row_rel(ours, theirs)
        assert err < 1e-6, err


def test_rays_invalid_frame_is_nan(cuda):
    K, E = cameras(2, seed=5)
    pid = torch.tensor([[0, 10, 20], [2, 10, 20], [-1, 0, 0], [1, 5, 5]], dtype=torch.int32)
    ro, rd, vd, no, nd = geometry.rays_fused(pid.to(cuda), K.to(cuda), E.to(cuda), mip_nerf_used=False, ndc=True,
                                             resolution=(H, W), near=NEAR)
    bad = torch.tensor([False, True, True, False])
    for t in (ro, rd, vd, no, nd):
        t = t.cpu()
        assert torch.isnan(t[bad]).all() and torch.isfinite(t[~bad]).all()

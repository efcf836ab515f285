"""Seeded synthetic workloads for the benchmarks and tests.

Camera: forward-facing pinhole, W x H = 504 x 378 (focal 0.8 W, principal point at the image
centre), pose = R_yaw(theta) R_pitch(phi) with theta, phi ~ U(-5 deg, 5 deg) per frame and
zero translation; NDC with near = 1.0 (the reference's `near`) and samples over
[near_ndc, far_ndc] = [0, 1] (src/data_preprocessors/DataPreprocessor10.py:162-167).
All randomness comes from one torch.Generator so runs are reproducible.
"""
from __future__ import annotations

import math

import torch

from .geometry import get_ndc_rays_tr, get_rays_tr, get_view_dirs_tr


def camera(width=504, height=378, n_frames=3, seed=0):
    g = torch.Generator().manual_seed(seed)
    f = 0.8 * width
    K = torch.tensor([[f, 0., width / 2.], [0., f, height / 2.], [0., 0., 1.]], dtype=torch.float32)
    poses = []
    for _ in range(n_frames):
        th, ph = ((torch.rand(2, generator=g) * 10. - 5.) * math.pi / 180.).tolist()
        ry = torch.tensor([[math.cos(th), 0., math.sin(th)], [0., 1., 0.], [-math.sin(th), 0., math.cos(th)]])
        rx = torch.tensor([[1., 0., 0.], [0., math.cos(ph), -math.sin(ph)], [0., math.sin(ph), math.cos(ph)]])
        P = torch.eye(4)
        P[:3, :3] = ry @ rx
        poses.append(P)
    return dict(K=K, poses=torch.stack(poses).float(), resolution=(height, width), n_frames=n_frames, g=g)


def rays(n_rays: int, seed=0, width=504, height=378, n_frames=3, near=1.0, device='cpu', pixel_id=None):
    """Random pixel rays: returns dict(pixel_id, rays_o, rays_d, rays_o_ndc, rays_d_ndc, view_dirs)."""
    cam = camera(width, height, n_frames, seed)
    g = cam['g']
    if pixel_id is None:
        img = torch.randint(0, n_frames, (n_rays,), generator=g)
        x = torch.randint(0, width, (n_rays,), generator=g)
        y = torch.randint(0, height, (n_rays,), generator=g)
        pixel_id = torch.stack([img, x, y], 1).int()
    K = cam['K'][None].expand(n_rays, 3, 3)
    E = cam['poses'][pixel_id[:, 0].long()]
    rays_o, rays_d = get_rays_tr(pixel_id, K, E, mip_nerf_used=False)
    rays_o_ndc, rays_d_ndc = get_ndc_rays_tr(rays_o, rays_d, cam['resolution'], K, near=near)
    out = dict(pixel_id=pixel_id, rays_o=rays_o, rays_d=rays_d, rays_o_ndc=rays_o_ndc, rays_d_ndc=rays_d_ndc,
               view_dirs=get_view_dirs_tr(rays_d), K=K, E=E)
    return {k: v.to(device) for k, v in out.items()}


def full_image_pixel_ids(width=1008, height=756, image_id=0):
    """All pixels of one frame, row-major (x fastest), as the reference's test grid."""
    y, x = torch.meshgrid(torch.arange(height), torch.arange(width), indexing='ij')
    return torch.stack([torch.full_like(x, image_id), x, y], -1).reshape(-1, 3).int()


def train_batch(n_rays: int, seed=0, n_frames=3, width=504, height=378, sparse_fraction=0.1, device='cpu'):
    """One synthetic training batch: pixel ids of the seeded camera, targets rgb ~ U(0,1), and
    sparse depth on `sparse_fraction` of the rays (NDC depth ~ U(0.2, 0.9), supervised as the
    metric depth it maps to, as SparseDepthMSE14 supervises metric depth)."""
    r = rays(n_rays, seed=seed, width=width, height=height, n_frames=n_frames)
    g = torch.Generator().manual_seed(seed + 1)
    target = torch.rand(n_rays, 3, generator=g)
    sd_mask = torch.rand(n_rays, generator=g) < sparse_fraction
    zn = 0.2 + 0.7 * torch.rand(n_rays, 1, generator=g)
    oz, dz = r['rays_o'][:, 2:3], r['rays_d'][:, 2:3]
    tn = -(1.0 + oz) / dz
    sd_depth = ((oz + tn * dz) / dz * (1 / (1 - zn) - 1) + tn)[:, 0]
    b = dict(pixel_id=r['pixel_id'], target_rgb=target, sd_mask=sd_mask, sd_depth=sd_depth)
    return {k: v.to(device) for k, v in b.items()}


def train_loss(out: dict, batch: dict):
    """Synthetic objective of the benchmark: photometric MSE on the main coarse/fine and both
    augmented coarse renders (MSE14), 0.1 x sparse-depth MSE on metric depth (SparseDepthMSE14),
    and 0.1 x a plain MSE between augmented and main coarse depths standing in for the
    reprojection-gated AugmentationsDepthLoss11.  Each family of terms is evaluated as one stacked
    expression (the sum of the per-term means), a handful of launches instead of several per term."""
    t = batch['target_rgb']
    rgbs = [out[k] for k in ('rgb_coarse', 'rgb_fine', 'points_augmentation_rgb_coarse',
                             'views_augmentation_rgb_coarse') if k in out]
    loss = ((torch.stack(rgbs) - t) ** 2).mean(dim=(1, 2)).sum()
    m = batch['sd_mask']
    depths = [out[k] for k in ('depth_coarse', 'points_augmentation_depth_coarse',
                               'views_augmentation_depth_coarse') if k in out]
    if depths:
        scale = 0.1 * (m.numel() / m.sum().clamp_min(1))
        loss = loss + scale * (((torch.stack(depths) - batch['sd_depth']) ** 2) * m).mean(dim=1).sum()
    aug = [out[k] for k in ('points_augmentation_depth_coarse', 'views_augmentation_depth_coarse') if k in out]
    if aug:
        loss = loss + 0.1 * ((torch.stack(aug) - out['depth_coarse'].detach()) ** 2).mean(dim=1).sum()
    return loss

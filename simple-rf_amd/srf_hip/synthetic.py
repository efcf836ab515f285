"""Seeded synthetic workloads for the benchmarks and tests.

Camera: forward-facing pinhole, W x H = 504 x 378 (focal 0.8 W, principal point at the image
centre), pose = R_yaw(theta) R_pitch(phi) with theta, phi ~ U(-5 deg, 5 deg) per frame and the
centre of frame i at (0.04 i, 0.015 (i mod 2), 0.01 i) (distinct centres, so each frame has one
nearest neighbour for the reprojection losses); NDC with near = 1.0 (the reference's `near`) and
samples over [near_ndc, far_ndc] = [0, 1] (src/data_preprocessors/DataPreprocessor10.py:162-167).
Ground-truth images: every frame sees one textured plane at z = -4, so the frames agree where
they overlap.  All randomness comes from one torch.Generator so runs are reproducible.
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
        i = len(poses)
        P[:3, 3] = torch.tensor([0.04 * i, 0.015 * (i % 2), 0.01 * i])
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


PLANE_Z = -4.0


def _texture(p: torch.Tensor) -> torch.Tensor:
    return torch.stack([0.5 + 0.4 * torch.sin(2.3 * p[..., 0] + 1.1 * p[..., 1]),
                        0.5 + 0.4 * torch.cos(1.7 * p[..., 1] - 0.6 * p[..., 0]),
                        0.5 + 0.3 * torch.sin(1.3 * (p[..., 0] + p[..., 1]))], -1)


def images(width=504, height=378, n_frames=3, seed=0) -> torch.Tensor:
    """Ground-truth frames [n_frames, height, width, 3]: the plane z = PLANE_Z with a smooth colour
    texture, seen through each frame's pixel rays (the reference's ray convention)."""
    cam = camera(width, height, n_frames, seed)
    y, x = torch.meshgrid(torch.arange(height), torch.arange(width), indexing='ij')
    out = []
    for i in range(n_frames):
        pid = torch.stack([torch.full_like(x, i), x, y], -1).reshape(-1, 3).int()
        K = cam['K'][None].expand(pid.shape[0], 3, 3)
        E = cam['poses'][i][None].expand(pid.shape[0], 4, 4)
        o, d = get_rays_tr(pid, K, E, mip_nerf_used=False)
        t = (PLANE_Z - o[:, 2]) / d[:, 2]
        out.append(_texture(o + d * t[:, None]).reshape(height, width, 3))
    return torch.stack(out).float()


def full_image_pixel_ids(width=1008, height=756, image_id=0):
    """All pixels of one frame, row-major (x fastest), as the reference's test grid."""
    y, x = torch.meshgrid(torch.arange(height), torch.arange(width), indexing='ij')
    return torch.stack([torch.full_like(x, image_id), x, y], -1).reshape(-1, 3).int()


def train_batch(n_rays: int, seed=0, n_frames=3, width=504, height=378, sparse_fraction=0.1, device='cpu',
                scene_seed=0):
    """One synthetic training batch as the reference's data loader splits it: a fraction
    `sparse_fraction` of the rays carry a sparse depth (indices_mask_sparse_depth; NDC depth
    ~ U(0.2, 0.9), supervised as the metric depth it maps to, as SparseDepthMSE14 supervises metric
    depth), the others are NeRF rays (indices_mask_nerf) whose target is the ground-truth frame's
    colour at their pixel.  The frames themselves (images) and their poses come with the batch, as
    the reference's common_data carries them for the reprojection losses; they are those of the
    scene camera (camera(seed=scene_seed), the one configs.model_configs describes), while `seed`
    draws the pixels."""
    r = rays(n_rays, seed=seed, width=width, height=height, n_frames=n_frames)
    g = torch.Generator().manual_seed(seed + 1)
    imgs = images(width, height, n_frames, scene_seed)
    pid = r['pixel_id'].long()
    target = imgs[pid[:, 0], pid[:, 2], pid[:, 1]]
    sd_mask = torch.rand(n_rays, generator=g) < sparse_fraction
    zn = 0.2 + 0.7 * torch.rand(n_rays, 1, generator=g)
    oz, dz = r['rays_o'][:, 2:3], r['rays_d'][:, 2:3]
    tn = -(1.0 + oz) / dz
    sd_depth = ((oz + tn * dz) / dz * (1 / (1 - zn) - 1) + tn)[:, 0]
    b = dict(pixel_id=r['pixel_id'], target_rgb=target, indices_mask_nerf=~sd_mask, indices_mask_sparse_depth=sd_mask,
             sd_depth=sd_depth, images=imgs, poses=camera(width, height, n_frames, scene_seed)['poses'])
    return {k: v.to(device) for k, v in b.items()}


# Loss weights of the reference's Simple-NeRF training configuration (src/NerfLlffTrainerTester10.py:381-417;
# the two depth-consistency losses at their weight from iteration 10000 on)
LOSS_WEIGHTS = {'MSE14': 1.0, 'SparseDepthMSE14': 0.1, 'AugmentationsDepthLoss11': 0.1,
                'CoarseFineConsistencyLoss34': 0.1}
RGB_KEYS = ('rgb_coarse', 'rgb_fine', 'points_augmentation_rgb_coarse', 'views_augmentation_rgb_coarse')
SD_KEYS = ('depth_coarse', 'points_augmentation_depth_coarse', 'views_augmentation_depth_coarse')
AUG_NAMES = ('points_augmentation', 'views_augmentation')


def train_loss(out: dict, batch: dict):
    """The reference's training objective on the fused loss kernels: MSE14 on the NeRF rays of the
    main coarse / fine and both augmented coarse renders, 0.1 x SparseDepthMSE14 on the metric depths
    of the coarse renders over the sparse-depth rays, 0.1 x AugmentationsDepthLoss11 (main coarse
    against each augmented coarse depth, reprojection-gated by the ground-truth frames) and 0.1 x
    CoarseFineConsistencyLoss34 (coarse against fine depth, gated the same way, plus coarse towards
    the detached fine depth on the sparse-depth rays)."""
    from . import losses
    w = LOSS_WEIGHTS
    nerf, sd = batch['indices_mask_nerf'], batch['indices_mask_sparse_depth']
    loss = w['MSE14'] * losses.mse_terms([out[k] for k in RGB_KEYS if k in out], batch['target_rgb'], nerf)
    depths = [out[k] for k in SD_KEYS if k in out]
    if depths:
        loss = loss + w['SparseDepthMSE14'] * losses.mse_terms(depths, batch['sd_depth'], sd)
    frames = batch['images']
    res = tuple(frames.shape[1:3])

    def gated(d1, d2, rule):
        return losses.depth_consistency(d1, d2, nerf, out['rays_o'], out['rays_d'], out['extrinsics_all'], frames,
                                        batch['pixel_id'], out['intrinsics'], res, both_invalid_rule=rule)[0]
    aug = [gated(out['depth_coarse'], out[f'{a}_depth_coarse'], True) for a in AUG_NAMES if f'{a}_depth_coarse' in out]
    if aug:
        loss = loss + w['AugmentationsDepthLoss11'] * sum(aug)
    if 'depth_fine' in out:
        cf = gated(out['depth_coarse'], out['depth_fine'], False) + \
            losses.mse_terms([out['depth_coarse']], out['depth_fine'].detach(), sd)
        loss = loss + w['CoarseFineConsistencyLoss34'] * cf
    return loss

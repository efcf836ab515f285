"""Oracle: the Simple-NeRF training render (coarse + fine + augmented coarse models) composed
from the oracle functions, in plain PyTorch on any device / dtype (TEST INFRASTRUCTURE ONLY;
also the timed plain-PyTorch baseline of bench.py).

Follows SimpleNeRF.render_rays (src/models/SimpleNeRF17.py:159-328) for the NDC, training,
augmentation branch, with the random draws (u for the coarse/fine depths, sigma noise per
network) passed in explicitly.
"""
from __future__ import annotations

import torch

from . import nerf as on

AUG_CFG = {
    'coarse': {},
    'fine': {},
    'points_augmentation': {'sigma_pe_degree': 3},
    'views_augmentation': {'use_view_dirs': False, 'view_dependent_rgb': False},
}


def render_train(params: dict, pixel_id, K, E, resolution, *, u_coarse=None, u_fine=None, noise=None,
                 n_coarse=64, n_fine=128, near=1.0, white_bkgd=False, augment=True, z_fine=None):
    """params: {'coarse'|'fine'|'points_augmentation'|'views_augmentation': reference-named tensors}.
    K [N,3,3], E [N,4,4] per ray.  noise: dict name -> [N*S] scaled sigma noise (or None).
    z_fine: the merged [N, n_coarse + n_fine] fine depths to render instead of drawing them from
    the coarse weights (a parity check conditioned on identical fine samples)."""
    rays_o, rays_d = on.get_rays(pixel_id, K, E)
    o_ndc, d_ndc = on.get_ndc_rays(rays_o, rays_d, resolution, K, near)
    view_dirs = on.get_view_dirs(rays_d)
    n = pixel_id.shape[0]
    dt = K.dtype
    out = {}

    def net_render(name, z):
        s = z.shape[1]
        pts = on.sample_points(o_ndc, d_ndc, z).reshape(-1, 3)
        dirs = view_dirs[:, None].expand(n, s, 3).reshape(-1, 3)
        nz = None if noise is None or noise.get(name) is None else noise[name].reshape(-1, 1).to(dt)
        net = on.mlp_forward(params[name], pts, dirs, noise=nz, **AUG_CFG[name])
        return on.volume_rendering(net['sigma'].reshape(n, s), net['rgb'].reshape(n, s, 3), z_vals_ndc=z,
                                   rays_o=rays_o, rays_d=rays_d, rays_d_ndc=d_ndc, ndc=True, white_bkgd=white_bkgd)

    out['rays_o'], out['rays_d'] = rays_o, rays_d
    z_c = on.stratified_z_vals(n, n_coarse, 0., 1., u=u_coarse, dtype=dt, device=K.device)
    vr = net_render('coarse', z_c)
    for k, v in vr.items():
        out[f'{k}_coarse'] = v
    if augment:
        for name in ('points_augmentation', 'views_augmentation'):
            if name in params:
                for k, v in net_render(name, z_c).items():
                    out[f'{name}_{k}_coarse'] = v
    if 'fine' in params:
        z_f = on.fine_z_vals(z_c, vr['weights'], n_fine, u=u_fine)[0] if z_fine is None else z_fine.to(dt)
        for k, v in net_render('fine', z_f).items():
            out[f'{k}_fine'] = v
    return out


LOSS_WEIGHTS = {'MSE14': 1.0, 'SparseDepthMSE14': 0.1, 'AugmentationsDepthLoss11': 0.1,
                'CoarseFineConsistencyLoss34': 0.1}


def reference_objective(out: dict, batch: dict, K):
    """The Simple-NeRF training objective of src/NerfLlffTrainerTester10.py:381-417 (LossComputer03
    weighting the losses, the depth-consistency ones at their weight from iteration 10000 on), on the
    oracle's losses: MSE14 on the NeRF rays, SparseDepthMSE14 on the coarse renders' metric depths over
    the sparse-depth rays, AugmentationsDepthLoss11 (main against each augmented coarse depth) and
    CoarseFineConsistencyLoss34 (coarse against fine, plus its sparse-depth term).  out: render_train's
    output; batch: pixel_id, target_rgb, indices_mask_nerf, indices_mask_sparse_depth, sd_depth, images,
    poses (per frame); K: per-ray intrinsics."""
    from . import losses as ol
    w = LOSS_WEIGHTS
    nerf, sd = batch['indices_mask_nerf'], batch['indices_mask_sparse_depth']
    loss = w['MSE14'] * sum(ol.mse(out[k], batch['target_rgb'], nerf) for k in
                            ('rgb_coarse', 'rgb_fine', 'points_augmentation_rgb_coarse', 'views_augmentation_rgb_coarse')
                            if k in out)
    loss = loss + w['SparseDepthMSE14'] * sum(ol.sparse_depth_mse(out[k], batch['sd_depth'], sd) for k in
                                              ('depth_coarse', 'points_augmentation_depth_coarse',
                                               'views_augmentation_depth_coarse') if k in out)
    images = batch['images'].to(out['rgb_coarse'].dtype)
    res = tuple(images.shape[1:3])
    poses = batch['poses'].to(out['rgb_coarse'].dtype)

    def gated(d1, d2, rule):
        return ol.depth_consistency_nerf(d1, d2, nerf, out['rays_o'], out['rays_d'], poses, images, batch['pixel_id'],
                                         K, res, both_invalid_rule=rule)[0]
    for a in ('points_augmentation', 'views_augmentation'):
        if f'{a}_depth_coarse' in out:
            loss = loss + w['AugmentationsDepthLoss11'] * gated(out['depth_coarse'], out[f'{a}_depth_coarse'], True)
    if 'depth_fine' in out:
        loss = loss + w['CoarseFineConsistencyLoss34'] * (
            gated(out['depth_coarse'], out['depth_fine'], False) +
            ol.coarse_fine_sparse_depth(out['depth_coarse'], out['depth_fine'], sd))
    return loss

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

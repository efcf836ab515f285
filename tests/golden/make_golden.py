"""Writes tests/golden/mlp_golden.npz and tests/golden/loss_golden.npz: input/output vectors of the
reference's own NeRF MLP, TensoRF colour network and depth-consistency losses, for the oracle functions
that no closed form pins (the MLPs' concatenation orders; the losses' reprojection, patch gating and
in-place masking).

The reference classes are imported from the reference tree in THIS container only, instantiated with
seeded weights (narrower than the shipped 256/128 widths, so the fixture stays small; the wiring under
test does not depend on the width), put in eval() (no sigma noise) and evaluated on 64 seeded points
in float32 and float64.  The fixture holds only data: every variant's state_dict, the inputs and the
outputs, as plain arrays (numpy.load with allow_pickle=False reads it).

Variants (reference classes and configs):
  nerf_main        MLP, src/models/SimpleNeRF17.py:616-785, points PE 10, views PE 4, view-dependent rgb
  nerf_points_aug  same class with points_sigma_positional_encoding_degree 3 (the 21/42 split,
                   src/NerfLlffTrainerTester10.py:321-333)
  nerf_views_aug   same class with use_view_dirs False, view_dependent_rgb False (:347-358)
  color_shipped    MlpFeaturesColorPredictor, src/models/SimpleTensoRF09.py:1370-1421, features 27,
                   features PE 0, views PE 0 (src/NerfLlffTrainerTester10.py:958-961)
  color_pe         same class with features PE 2 and views PE 2 (both encoders and their order)
  aug_depth        AugmentationsDepthLoss.compute_loss_nerf, src/loss_functions/AugmentationsDepthLoss11.py:108-181
  coarse_fine      CoarseFineConsistencyLoss.compute_loss_nerf, src/loss_functions/CoarseFineConsistencyLoss34.py:85-164
                   (loss, both loss maps and the gradients of the loss w.r.t. both depths, on a seeded scene
                   of four forward-facing views with smooth images so that the patch gate takes both
                   decisions)
and tests/golden/tensorf_mods_golden.npz: the TensoRF model modifications of VmDecomposedTensor
(src/models/SimpleTensoRF09.py:820-944, 1274-1321) on a 20^3-voxel grid whose density is occupied in a
sub-box only: the alpha-mask update (dense evaluation, 3x3x3 max pool, threshold, new box), the shrink,
the factor upsampling and a second alpha-mask update through the existing mask; and the TV loss
(src/loss_functions/TotalVariationLoss04.py:97-116) and the mass-concentration entropy
(src/loss_functions/MassConcentrationLoss06.py:63-83) with their gradients.

    python tests/golden/make_golden.py [/path/to/reference]
"""
import pathlib
import sys

import numpy as np
import torch

HERE = pathlib.Path(__file__).resolve().parent
N_POINTS = 64
NERF_WIDTH, NERF_VIEWS_WIDTH = 32, 16
COLOR_UNITS = 32


def nerf_mlp_configs(variant: str) -> dict:
    cfg = {'points_net_depth': 8, 'views_net_depth': 1, 'points_net_width': NERF_WIDTH,
           'views_net_width': NERF_VIEWS_WIDTH, 'points_positional_encoding_degree': 10,
           'views_positional_encoding_degree': 4, 'use_view_dirs': True, 'view_dependent_rgb': True,
           'predict_visibility': False}
    if variant == 'points_aug':
        cfg['points_sigma_positional_encoding_degree'] = 3
    elif variant == 'views_aug':
        cfg['use_view_dirs'] = False
        cfg['view_dependent_rgb'] = False
    return cfg


def color_tensor_configs(features_pe: int, views_pe: int) -> dict:
    return {'features_positional_encoding_degree': features_pe, 'views_positional_encoding_degree': views_pe,
            'use_view_dirs': True, 'view_dependent_color': True}


def seeded_(module: torch.nn.Module, seed: int) -> None:
    """Draws every parameter from one seeded generator: weights uniform with the variance 2/fan_in
    that keeps activations from shrinking through the ReLU trunk, biases uniform in [-0.1, 0.1], so
    that the biases are not zero and every input channel moves the outputs."""
    g = torch.Generator().manual_seed(seed)
    with torch.no_grad():
        for p in module.parameters():
            scale = (6.0 / p.shape[-1]) ** 0.5 if p.ndim == 2 else 0.1
            p.copy_((torch.rand(p.shape, generator=g, dtype=torch.float64) * 2 - 1) * scale)


def main(ref_root: str = '/root/reference') -> None:
    sys.path.insert(0, str(pathlib.Path(ref_root) / 'src'))
    from models.SimpleNeRF17 import MLP
    from models.SimpleTensoRF09 import MlpFeaturesColorPredictor

    g = torch.Generator().manual_seed(20240620)
    pts = (torch.rand(N_POINTS, 3, generator=g, dtype=torch.float64) * 2 - 1) * 1.2
    dirs = torch.randn(N_POINTS, 3, generator=g, dtype=torch.float64)
    dirs = dirs / dirs.norm(dim=-1, keepdim=True)
    features = torch.randn(N_POINTS, 27, generator=g, dtype=torch.float64) * 0.7
    arrays = {'pts': pts.numpy(), 'dirs': dirs.numpy(), 'features': features.numpy()}

    configs = {'model': {'raw_noise_std': 1.0}}
    for seed, variant in enumerate(('main', 'points_aug', 'views_aug')):
        mlp = MLP('coarse_model', configs, nerf_mlp_configs(variant), {})
        seeded_(mlp, 110 + seed)
        mlp.eval()
        for p in mlp.state_dict().items():
            arrays[f'nerf_{variant}/param/{p[0]}'] = p[1].numpy().astype(np.float32)
        for dtype, tag in ((torch.float32, 'f32'), (torch.float64, 'f64')):
            m = mlp.to(dtype)
            with torch.no_grad():
                out = m({'pts': pts.to(dtype), 'view_dirs': dirs.to(dtype)})
            arrays[f'nerf_{variant}/{tag}/sigma'] = out['sigma'].numpy()
            arrays[f'nerf_{variant}/{tag}/rgb'] = out['rgb'].numpy()
        mlp.float()

    for seed, (name, fpe, vpe) in enumerate((('shipped', 0, 0), ('pe', 2, 2))):
        cp = MlpFeaturesColorPredictor({}, color_tensor_configs(fpe, vpe), {}, 27, fpe, vpe, COLOR_UNITS)
        seeded_(cp, 200 + seed)
        cp.eval()
        for k, v in cp.state_dict().items():
            arrays[f'color_{name}/param/{k}'] = v.numpy().astype(np.float32)
        for dtype, tag in ((torch.float32, 'f32'), (torch.float64, 'f64')):
            m = cp.to(dtype)
            with torch.no_grad():
                out = m({'features': features.to(dtype), 'view_dirs': dirs.to(dtype)})
            arrays[f'color_{name}/{tag}/color'] = out['color'].numpy()
        cp.float()

    out_path = HERE / 'mlp_golden.npz'
    np.savez_compressed(out_path, **arrays)
    print('wrote', out_path, sum(a.nbytes for a in arrays.values()), 'bytes of arrays')
    write_loss_golden()
    write_tensorf_mods_golden()


def loss_scene(seed: int = 7, n_rays: int = 256, n_views: int = 4, h: int = 36, w: int = 44, plane: float = 4.0):
    """Four forward-facing views (distinct centres, so the nearest view is unique) of one textured plane
    z = -plane, each image rendered by intersecting its pixel rays (the reference's convention,
    CommonUtils04.py:73-95) with the plane, so that the views agree where the plane is seen.  Each ray
    carries two depth estimates near the true one (2 % and 6 % noise): their reprojections land inside
    and outside the image, and the 5x5 patch RMSE falls on both sides of the 0.1 threshold."""
    g = torch.Generator().manual_seed(seed)
    f = 0.9 * w
    K = torch.tensor([[f, 0., w / 2.], [0., f, h / 2.], [0., 0., 1.]], dtype=torch.float64)
    poses = torch.zeros(n_views, 4, 4, dtype=torch.float64)
    for v in range(n_views):
        th = (torch.rand((), generator=g, dtype=torch.float64) - 0.5) * 0.1
        c, s_ = torch.cos(th), torch.sin(th)
        poses[v, :3, :3] = torch.tensor([[c, 0., s_], [0., 1., 0.], [-s_, 0., c]], dtype=torch.float64)
        poses[v, :3, 3] = torch.tensor([0.12 * v, 0.05 * (v % 2), 0.02 * v], dtype=torch.float64)
        poses[v, 3, 3] = 1.

    def rays(v, x, y):
        d_cam = torch.stack([(x - w / 2.) / f, -(y - h / 2.) / f, -torch.ones_like(x)], -1)
        return poses[v, :3, 3].expand_as(d_cam), (poses[v, :3, :3] @ d_cam[..., None])[..., 0]

    def texture(p):
        return torch.stack([0.5 + 0.4 * torch.sin(2.3 * p[..., 0] + 1.1 * p[..., 1]),
                            0.5 + 0.4 * torch.cos(1.7 * p[..., 1] - 0.6 * p[..., 0]),
                            0.5 + 0.3 * torch.sin(1.3 * (p[..., 0] + p[..., 1]))], -1)

    yy, xx = torch.meshgrid(torch.arange(h, dtype=torch.float64), torch.arange(w, dtype=torch.float64), indexing='ij')
    imgs = []
    for v in range(n_views):
        o, d = rays(v, xx, yy)
        t = (-plane - o[..., 2]) / d[..., 2]
        imgs.append(texture(o + d * t[..., None]))
    imgs = torch.stack(imgs)
    img = torch.randint(0, n_views, (n_rays,), generator=g)
    x = torch.randint(0, w, (n_rays,), generator=g).double()
    y = torch.randint(0, h, (n_rays,), generator=g).double()
    pixel_id = torch.stack([img, x.long(), y.long()], -1)
    rays_o, rays_d = torch.stack([poses[i, :3, 3] for i in img.tolist()]), None
    rays_d = torch.stack([rays(int(i), x[k], y[k])[1] for k, i in enumerate(img.tolist())])
    t_true = (-plane - rays_o[:, 2]) / rays_d[:, 2]
    depth1 = t_true * (1 + 0.02 * torch.randn(n_rays, generator=g, dtype=torch.float64))
    depth2 = t_true * (1 + 0.06 * torch.randn(n_rays, generator=g, dtype=torch.float64))
    mask = torch.rand(n_rays, generator=g) < 0.85
    return dict(intrinsics=K[None].repeat(n_views, 1, 1), extrinsics=poses, images=imgs, pixel_id=pixel_id,
                rays_o=rays_o, rays_d=rays_d, depth1=depth1, depth2=depth2, indices_mask_nerf=mask,
                resolution=np.array([h, w]))


def write_loss_golden(ref_root: str = '/root/reference') -> None:
    from loss_functions.AugmentationsDepthLoss11 import AugmentationsDepthLoss
    from loss_functions.CoarseFineConsistencyLoss34 import CoarseFineConsistencyLoss
    sc = loss_scene()
    arrays = {k: (v.numpy() if isinstance(v, torch.Tensor) else v) for k, v in sc.items()}
    configs = {'model': {'coarse_model': {}, 'fine_model': {}, 'augmentations': []}, 'data_loader': {}}
    loss_configs = {'rmse_threshold': 0.1, 'patch_size': [5, 5]}
    classes = {'aug_depth': AugmentationsDepthLoss, 'coarse_fine': CoarseFineConsistencyLoss}
    for name, cls in classes.items():
        obj = cls(configs, loss_configs)
        # FP32 only: reproject builds its axis permutation in the default dtype (CommonUtils04.py:247-248)
        for dtype, tag in ((torch.float32, 'f32'),):
            d1 = sc['depth1'].to(dtype).clone().requires_grad_(True)
            d2 = sc['depth2'].to(dtype).clone().requires_grad_(True)
            loss, m1, m2 = obj.compute_loss_nerf(
                d1, d2, sc['indices_mask_nerf'], sc['rays_o'].to(dtype), sc['rays_d'].to(dtype),
                sc['extrinsics'].to(dtype), sc['images'].to(dtype), sc['pixel_id'], sc['intrinsics'].to(dtype),
                tuple(int(v) for v in sc['resolution']))
            loss.backward()
            arrays[f'{name}/{tag}/loss'] = np.array(loss.item())
            arrays[f'{name}/{tag}/map1'] = m1.detach().numpy()
            arrays[f'{name}/{tag}/map2'] = m2.detach().numpy()
            arrays[f'{name}/{tag}/grad1'] = d1.grad.numpy()
            arrays[f'{name}/{tag}/grad2'] = d2.grad.numpy()
    out_path = HERE / 'loss_golden.npz'
    np.savez_compressed(out_path, **arrays)
    print('wrote', out_path, sum(np.asarray(a).nbytes for a in arrays.values()), 'bytes of arrays')



TENSORF_CFG = {'decomposition_type': 'VectorMatrix', 'num_samples_max': 1e6, 'num_components_density': [4, 4, 4],
               'num_components_color': [8, 4, 4], 'bounding_box': [[-1.5, -1.67, -1.0], [1.5, 1.67, 1.0]],
               'num_voxels_initial': 20 ** 3, 'num_voxels_final': 40 ** 3, 'tensor_upsampling_iters': [2000, 3000],
               'num_voxels_per_sample': 0.5, 'alpha_mask_update_iters': [2500], 'alpha_mask_threshold': 0.0001,
               'ray_marching_weight_threshold': 0.0001, 'use_view_dirs': True, 'view_dependent_color': True,
               'views_positional_encoding_degree': 0, 'features_positional_encoding_degree': 0,
               'features_dimension_color': 27, 'density_offset': -10, 'distance_scale': 25, 'density_predictor': 'ReLU',
               'color_predictor': 'MLP_Features', 'num_units_color_predictor': 128, 'predict_visibility': False}


def write_tensorf_mods_golden() -> None:
    from models.SimpleTensoRF09 import VmDecomposedTensor
    from loss_functions.TotalVariationLoss04 import TotalVariationLoss
    from loss_functions.MassConcentrationLoss06 import MassConcentrationLoss
    arrays = {}
    m = VmDecomposedTensor('coarse_model', {'data_loader': {'ndc': True}, 'model': {}}, TENSORF_CFG,
                           {'bounding_box': TENSORF_CFG['bounding_box']})
    m.train()
    g = torch.Generator().manual_seed(31)
    with torch.no_grad():
        # density occupied in a sub-box only: the xy plane of component 0 positive on a rectangle, its z
        # line positive on a range, every other density factor zero; colour factors random
        for p in list(m.matrices_density) + list(m.vectors_density):
            p.zero_()
        pl, ln = m.matrices_density[0], m.vectors_density[0]
        pl[0, 0, 6:15, 4:13] = 0.5 + torch.rand(9, 9, generator=g)
        ln[0, 0, 3:9, 0] = 0.5 + torch.rand(6, generator=g)
        for p in list(m.matrices_color) + list(m.vectors_color):
            p.copy_(0.1 * torch.randn(p.shape, generator=g))

    def snap(tag):
        arrays[f'{tag}/resolution'] = m.resolution.numpy().astype(np.int64)
        arrays[f'{tag}/bounding_box'] = m.bounding_box.numpy()
        arrays[f'{tag}/step_size'] = np.array(float(m.step_size))
        arrays[f'{tag}/num_samples'] = np.array(int(m.num_samples))
        for kind in ('density', 'color'):
            for i in range(3):
                arrays[f'{tag}/matrices_{kind}.{i}'] = getattr(m, f'matrices_{kind}')[i].detach().numpy()
                arrays[f'{tag}/vectors_{kind}.{i}'] = getattr(m, f'vectors_{kind}')[i].detach().numpy()
    snap('initial')
    with torch.no_grad():
        bb = m.update_alpha_mask(2500)
    arrays['alpha1/alpha_volume'] = m.alpha_mask.alpha_volume.numpy().astype(np.uint8)  # 0 / 1
    arrays['alpha1/bounding_box_new'] = bb.numpy()
    with torch.no_grad():
        m.shrink_tensor(bb)
    snap('shrunk')
    with torch.no_grad():
        m.upsample_model_resolution(3000)
    snap('upsampled')
    with torch.no_grad():
        bb2 = m.update_alpha_mask(2500)
    arrays['alpha2/alpha_volume'] = m.alpha_mask.alpha_volume.numpy().astype(np.uint8)
    arrays['alpha2/bounding_box_new'] = bb2.numpy()
    # TV on the upsampled planes (VM) and lines (as CP components), at iteration weight 0.7
    for kind in ('density', 'color'):
        for comps, tag in ((getattr(m, f'matrices_{kind}'), 'planes'), (getattr(m, f'vectors_{kind}'), 'lines')):
            leaves = [p.detach().clone().requires_grad_(True) for p in comps]
            tv = TotalVariationLoss.compute_tv_loss(leaves, 0.7, False)['loss_value']
            tv.backward()
            arrays[f'tv/{kind}_{tag}/loss'] = np.array(tv.item())
            for i, p in enumerate(leaves):
                arrays[f'tv/{kind}_{tag}/grad.{i}'] = p.grad.numpy()
    # mass concentration on seeded weights of 40 rays x 443 samples, 11 bins
    w = torch.rand(40, 443, generator=g) ** 4 / 40
    mask = torch.rand(40, generator=g) < 0.8
    mc = MassConcentrationLoss({'model': {}}, {'num_bins': 11})
    wl = w.clone().requires_grad_(True)
    ent = mc.compute_entropy_loss(wl, mask, False)['loss_value']
    ent.backward()
    arrays['entropy/weights'] = w.numpy()
    arrays['entropy/mask'] = mask.numpy()
    arrays['entropy/num_bins'] = np.array(11)
    arrays['entropy/loss'] = np.array(ent.item())
    arrays['entropy/grad'] = wl.grad.numpy()
    out_path = HERE / 'tensorf_mods_golden.npz'
    np.savez_compressed(out_path, **arrays)
    print('wrote', out_path, sum(np.asarray(a).nbytes for a in arrays.values()), 'bytes of arrays')


if __name__ == '__main__':
    main(*sys.argv[1:])

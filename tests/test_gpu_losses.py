"""GPU parity of the fused loss kernels (srf_mse_terms_*, srf_depth_consistency_*) and of the loss
plugins built on them, against the oracle (oracle/losses.py) and the vectors the reference's own
classes wrote (tests/golden/loss_golden.npz)."""
import pathlib

import numpy as np
import pytest
import torch

import oracle.losses as ol
from srf_hip import losses

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / 'golden' / 'loss_golden.npz'


@pytest.fixture(scope='module')
def scene():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: torch.from_numpy(z[k]) if z[k].ndim else z[k] for k in z.files}


def _gpu_pair(scene, cuda, name, d1=None, d2=None, patch=(5, 5)):
    d1 = (scene['depth1'] if d1 is None else d1).float().to(cuda).requires_grad_(True)
    d2 = (scene['depth2'] if d2 is None else d2).float().to(cuda).requires_grad_(True)
    out = losses.depth_consistency(d1, d2, scene['indices_mask_nerf'].to(cuda), scene['rays_o'].float().to(cuda),
                                   scene['rays_d'].float().to(cuda), scene['extrinsics'].float().to(cuda),
                                   scene['images'].float().to(cuda), scene['pixel_id'].to(cuda),
                                   scene['intrinsics'].float().to(cuda), tuple(int(v) for v in scene['resolution']),
                                   patch_size=patch, both_invalid_rule=(name == 'aug_depth'))
    out[0].backward()
    return out, d1.grad, d2.grad


@pytest.mark.parametrize('name', ['aug_depth', 'coarse_fine'])
def test_depth_consistency_matches_reference_vectors(cuda, scene, name):
    """Loss, loss maps and both depth gradients against the vectors written by the reference's
    AugmentationsDepthLoss11 / CoarseFineConsistencyLoss34: the gate decisions (masks) exactly, the
    values to FP32 rounding."""
    (loss, m1, m2, flags), g1, g2 = _gpu_pair(scene, cuda, name)
    sel = scene['indices_mask_nerf']
    ref_loss = float(scene[f'{name}/f32/loss'])
    assert abs(loss.item() - ref_loss) <= 1e-6 * abs(ref_loss)
    assert torch.allclose(m1.cpu()[sel], scene[f'{name}/f32/map1'], rtol=1e-6, atol=1e-9)
    assert torch.equal(m2.cpu()[sel], scene[f'{name}/f32/map2'])
    assert torch.allclose(g1.cpu(), scene[f'{name}/f32/grad1'], rtol=1e-5, atol=1e-9)
    assert torch.equal(g2.cpu(), scene[f'{name}/f32/grad2'])
    # the gate decisions agree with the oracle's exactly (the oracle is bit-exact to the vectors)
    d1 = scene['depth1'].float().requires_grad_(True)
    d2 = scene['depth2'].float().requires_grad_(True)
    _, _, _, k1, k2 = ol.depth_consistency_nerf(d1, d2, sel, scene['rays_o'].float(), scene['rays_d'].float(),
                                                scene['extrinsics'].float(), scene['images'].float(),
                                                scene['pixel_id'], scene['intrinsics'].float(),
                                                tuple(int(v) for v in scene['resolution']),
                                                both_invalid_rule=(name == 'aug_depth'))
    f = flags.cpu()[sel]
    assert torch.equal((f & 1).bool(), k1) and torch.equal((f & 2).bool(), k2)
    assert not (flags.cpu()[~sel] & 4).any()


def _random_scene(n, seed, h=120, w=160, views=5):
    """A larger seeded scene for the oracle comparison: smooth images, cameras on a short baseline,
    depths spread so that every branch of the gate is taken."""
    g = torch.Generator().manual_seed(seed)
    f = 0.9 * w
    K = torch.tensor([[f, 0., w / 2.], [0., f, h / 2.], [0., 0., 1.]])
    E = torch.eye(4).repeat(views, 1, 1)
    E[:, :3, 3] = torch.rand(views, 3, generator=g) * torch.tensor([0.6, 0.3, 0.1])
    yy, xx = torch.meshgrid(torch.arange(h).float(), torch.arange(w).float(), indexing='ij')
    imgs = torch.stack([torch.stack([0.5 + 0.45 * torch.sin(0.07 * xx + 0.05 * yy + 0.3 * v),
                                     0.5 + 0.45 * torch.cos(0.06 * yy - 0.02 * xx),
                                     0.5 + 0.3 * torch.sin(0.04 * (xx + yy) + 0.1 * v)], -1) for v in range(views)])
    img = torch.randint(0, views, (n,), generator=g)
    x = torch.randint(0, w, (n,), generator=g)
    y = torch.randint(0, h, (n,), generator=g)
    d_cam = torch.stack([(x.float() - w / 2.) / f, -(y.float() - h / 2.) / f, -torch.ones(n)], -1)
    rays_d = (E[img, :3, :3] @ d_cam[..., None])[..., 0]
    rays_o = E[img, :3, 3]
    d1 = 1.0 + 6.0 * torch.rand(n, generator=g)
    d2 = d1 * (1 + 0.1 * torch.randn(n, generator=g))
    sel = torch.rand(n, generator=g) < 0.8
    return dict(depth1=d1, depth2=d2, indices_mask_nerf=sel, rays_o=rays_o, rays_d=rays_d, extrinsics=E,
                images=imgs, pixel_id=torch.stack([img, x, y], -1), intrinsics=K[None].repeat(views, 1, 1),
                resolution=np.array([h, w]))


@pytest.mark.parametrize('rule,patch', [(True, (5, 5)), (False, (5, 5)), (True, (7, 7))])
def test_depth_consistency_matches_oracle_at_batch_size(cuda, rule, patch):
    """4096 rays: the gate decisions agree with the FP32 oracle except on rays whose reprojection sits
    within FP32 rounding of a half-pixel (the rounding to a pixel) or whose RMSEs tie to rounding; the
    loss and the gradient agree to FP32 rounding once those rays are set aside.  The 5 x 5 patches run
    the kernel's compile-time patch path, 7 x 7 the general one."""
    sc = _random_scene(4096, 11)
    name = 'aug_depth' if rule else 'coarse_fine'
    (loss, m1, m2, flags), g1, g2 = _gpu_pair(sc, cuda, name, patch=patch)
    d1 = sc['depth1'].clone().requires_grad_(True)
    d2 = sc['depth2'].clone().requires_grad_(True)
    r_loss, r1, r2, k1, k2 = ol.depth_consistency_nerf(d1, d2, sc['indices_mask_nerf'], sc['rays_o'], sc['rays_d'],
                                                       sc['extrinsics'], sc['images'], sc['pixel_id'],
                                                       sc['intrinsics'], tuple(int(v) for v in sc['resolution']),
                                                       patch_size=patch, both_invalid_rule=rule)
    r_loss.backward()
    sel = sc['indices_mask_nerf']
    f = flags.cpu()[sel]
    differ = ((f & 1).bool() != k1) | (((f & 2) >> 1).bool() != k2)
    assert int(differ.sum()) <= 4, int(differ.sum())
    assert int(k1.sum()) > 100 and int(k2.sum()) > 100
    agree = torch.zeros_like(sel)
    agree[sel] = ~differ
    assert torch.allclose(g1.cpu()[agree], d1.grad[agree], rtol=1e-5, atol=1e-9)
    if int(differ.sum()) == 0:
        assert abs(loss.item() - r_loss.item()) <= 1e-5 * abs(r_loss.item())


def test_mse_terms_match_oracle_and_autograd(cuda):
    """Several rgb terms with weights, a boolean and an index selection: loss, maps and gradients
    against the oracle (MSE14.compute_mse per term) in FP64."""
    g = torch.Generator().manual_seed(3)
    n = 3000
    target = torch.rand(n, 3, generator=g)
    preds = [torch.rand(n, 3, generator=g) for _ in range(4)]
    weights = [1.0, 0.5, 2.0, 0.25]
    mask = torch.rand(n, generator=g) < 0.7
    for sel in (mask, torch.nonzero(mask)[:, 0]):
        pg = [p.to(cuda).requires_grad_(True) for p in preds]
        loss, maps = losses.mse_terms(pg, target.to(cuda), sel.to(cuda), weights, return_maps=True)
        loss.backward()
        pd = [p.double().requires_grad_(True) for p in preds]
        ref = sum(w * ol.mse(p, target.double(), mask) for w, p in zip(weights, pd))
        ref.backward()
        assert abs(loss.item() - ref.item()) <= 2e-6 * ref.item()
        for i in range(4):
            want = torch.where(mask, ((preds[i] - target) ** 2).mean(1), torch.zeros(n))
            assert torch.allclose(maps[i].cpu(), want, rtol=1e-6, atol=1e-8)
            assert torch.allclose(pg[i].grad.cpu().double(), pd[i].grad, rtol=1e-5, atol=1e-12)


def test_sparse_depth_terms_and_empty_selection(cuda):
    g = torch.Generator().manual_seed(4)
    n = 1000
    gt = torch.rand(n, generator=g) * 5
    ds = [torch.rand(n, generator=g) * 5 for _ in range(3)]
    mask = torch.rand(n, generator=g) < 0.1
    dg = [d.to(cuda).requires_grad_(True) for d in ds]
    loss = losses.mse_terms(dg, gt.to(cuda), mask.to(cuda), [1.0, 1.0, 1.0])
    loss.backward()
    ref = sum(ol.sparse_depth_mse(d.double(), gt.double(), mask) for d in ds)
    assert abs(loss.item() - ref.item()) <= 2e-6 * ref.item()
    for d in dg:
        assert not d.grad.cpu()[~mask].any()
    empty = losses.mse_terms([ds[0].to(cuda)], gt.to(cuda), torch.zeros(n, dtype=torch.bool, device=cuda))
    assert empty.item() == 0.0


def test_loss_plugins_match_reference_formulas(cuda, scene):
    """The drop-in plugins (loaded by config name like the reference's LossComputer) on one batch:
    MSEHip14 and SparseDepthMSEHip14 against the oracle sums, AugmentationsDepthLossHip11 and
    CoarseFineConsistencyLossHip34 against the reference vectors' loss."""
    import importlib
    n = scene['depth1'].shape[0]
    g = torch.Generator().manual_seed(5)
    configs = {'model': {'coarse_model': {}, 'fine_model': {},
                         'augmentations': [{'name': 'points_augmentation', 'coarse_model': {}}]},
               'data_loader': {'ndc': True}}
    regs = {'main_coarse': 1, 'main_fine': 1, 'points_augmentation_coarse': 1}
    out = {'rgb_coarse': torch.rand(n, 3, generator=g), 'rgb_fine': torch.rand(n, 3, generator=g),
           'points_augmentation_rgb_coarse': torch.rand(n, 3, generator=g),
           'depth_coarse': scene['depth1'].float(), 'depth_fine': scene['depth2'].float(),
           'points_augmentation_depth_coarse': scene['depth2'].float(),
           'rays_o': scene['rays_o'].float(), 'rays_d': scene['rays_d'].float(),
           'extrinsics_all': scene['extrinsics'].float(), 'intrinsics': scene['intrinsics'].float()}
    sd = torch.rand(n, generator=g) < 0.2
    inp = {'target_rgb': torch.rand(n, 3, generator=g), 'indices_mask_nerf': scene['indices_mask_nerf'],
           'indices_mask_sparse_depth': sd, 'sparse_depth_values': torch.rand(n, 1, generator=g) * 4,
           'pixel_id': scene['pixel_id'],
           'common_data': {'images': scene['images'].float(), 'resolution': tuple(int(v) for v in scene['resolution'])}}
    to = lambda d: {k: (to(v) if isinstance(v, dict) else v.to(cuda) if isinstance(v, torch.Tensor) else v)
                    for k, v in d.items()}
    gi, go = to(inp), to(out)
    mod = lambda name: getattr(importlib.import_module(f'loss_functions.{name}'), name[:-2])
    mse = mod('MSEHip14')(configs, {'models_to_regularize': regs}).compute_loss(gi, go, None)['loss_value']
    ref = sum(ol.mse(out[k], inp['target_rgb'], inp['indices_mask_nerf'])
              for k in ('rgb_coarse', 'rgb_fine', 'points_augmentation_rgb_coarse'))
    assert abs(mse.item() - ref.item()) <= 2e-6 * ref.item()
    sdl = mod('SparseDepthMSEHip14')(configs, {'models_to_regularize': regs}).compute_loss(gi, go, None)['loss_value']
    ref = sum(ol.sparse_depth_mse(out[k], inp['sparse_depth_values'][:, 0], sd)
              for k in ('depth_coarse', 'depth_fine', 'points_augmentation_depth_coarse'))
    assert abs(sdl.item() - ref.item()) <= 2e-6 * ref.item()
    lc = {'patch_size': [5, 5], 'rmse_threshold': 0.1}
    aug = mod('AugmentationsDepthLossHip11')(configs, lc).compute_loss(gi, go, None, return_loss_maps=True)
    assert abs(aug['loss_value'].item() - float(scene['aug_depth/f32/loss'])) <= 1e-6 * float(scene['aug_depth/f32/loss'])
    assert set(aug['loss_maps']) == {'AugmentationsDepthLoss11_points_augmentation_coarse_main',
                                     'AugmentationsDepthLoss11_points_augmentation_coarse_augmented'}
    cf = mod('CoarseFineConsistencyLossHip34')(configs, lc).compute_loss(gi, go, None)['loss_value']
    assert abs(cf.item() - float(scene['coarse_fine/f32/loss'])) <= 1e-6 * float(scene['coarse_fine/f32/loss'])


@pytest.mark.parametrize('ndc', [True, False])
def test_coarse_fine_consistency_33_plugin(cuda, ndc):
    """CoarseFineConsistencyLossHip33 against CoarseFineConsistencyLoss33.py:35-57 in FP64: the mean of
    (coarse - fine)^2 over all rays on the NDC (or metric) depths, both depths differentiated, and its map."""
    import importlib
    g = torch.Generator().manual_seed(8)
    n = 3001
    key = 'depth_ndc' if ndc else 'depth'
    c0, f0 = torch.rand(n, generator=g), torch.rand(n, generator=g)
    c, f = c0.to(cuda).requires_grad_(True), f0.to(cuda).requires_grad_(True)
    configs = {'model': {'coarse_model': {}, 'fine_model': {}}, 'data_loader': {'ndc': ndc}}
    cls = importlib.import_module('loss_functions.CoarseFineConsistencyLossHip33').CoarseFineConsistencyLossHip
    res = cls(configs, {}).compute_loss({'target_rgb': torch.zeros(n, 3, device=cuda)},
                                        {f'{key}_coarse': c, f'{key}_fine': f}, None, return_loss_maps=True)
    res['loss_value'].backward()
    c64, f64 = c0.double().requires_grad_(True), f0.double().requires_grad_(True)
    m64 = torch.square(c64 - f64)
    ref = m64.mean()
    ref.backward()
    assert abs(res['loss_value'].item() - ref.item()) <= 1e-6 * ref.item()
    assert torch.allclose(c.grad.cpu().double(), c64.grad, rtol=1e-5, atol=1e-12)
    assert torch.allclose(f.grad.cpu().double(), f64.grad, rtol=1e-5, atol=1e-12)
    assert torch.allclose(res['loss_maps']['CoarseFineConsistencyLoss33'].detach().cpu().double(), m64.detach(),
                          rtol=1e-6, atol=1e-12)

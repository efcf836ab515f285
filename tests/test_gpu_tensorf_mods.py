"""GPU parity of the TensoRF model modifications and regularisers (csrc/tensorf_mods.hip) against the
vectors the reference's own VmDecomposedTensor, TotalVariationLoss and MassConcentrationLoss wrote
(tests/golden/tensorf_mods_golden.npz): the drop-in's VmDecomposedTensor is loaded with the reference's
initial factors and taken through the same sequence (alpha-mask update, shrink, upsampling, a second
alpha-mask update through the first mask)."""
import pathlib

import numpy as np
import pytest
import torch

from models.SimpleTensoRFHip09 import VmDecomposedTensor
from srf_hip import tensorf_mods

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / 'golden' / 'tensorf_mods_golden.npz'
CFG = {'decomposition_type': 'VectorMatrix', 'num_samples_max': 1e6, 'num_components_density': [4, 4, 4],
       'num_components_color': [8, 4, 4], 'bounding_box': [[-1.5, -1.67, -1.0], [1.5, 1.67, 1.0]],
       'num_voxels_initial': 20 ** 3, 'num_voxels_final': 40 ** 3, 'tensor_upsampling_iters': [2000, 3000],
       'num_voxels_per_sample': 0.5, 'alpha_mask_update_iters': [2500], 'alpha_mask_threshold': 0.0001,
       'ray_marching_weight_threshold': 0.0001, 'use_view_dirs': True, 'view_dependent_color': True,
       'views_positional_encoding_degree': 0, 'features_positional_encoding_degree': 0,
       'features_dimension_color': 27, 'density_offset': -10, 'distance_scale': 25, 'density_predictor': 'ReLU',
       'color_predictor': 'MLP_Features', 'num_units_color_predictor': 128, 'predict_visibility': False}


@pytest.fixture(scope='module')
def g():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: torch.from_numpy(z[k]) if z[k].ndim else z[k] for k in z.files}


def _model(cuda, g):
    configs = {'data_loader': {'ndc': True}, 'model': {'white_bkgd': False}}
    m = VmDecomposedTensor('coarse_model', configs, CFG, {'bounding_box': CFG['bounding_box']}).to(cuda)
    m.train()
    with torch.no_grad():
        for kind in ('density', 'color'):
            for i in range(3):
                getattr(m, f'matrices_{kind}')[i].copy_(g[f'initial/matrices_{kind}.{i}'])
                getattr(m, f'vectors_{kind}')[i].copy_(g[f'initial/vectors_{kind}.{i}'])
    assert m.resolution.tolist() == g['initial/resolution'].tolist()
    return m


def _check_factors(m, g, tag, atol):
    for kind in ('density', 'color'):
        for i in range(3):
            for key in ('matrices', 'vectors'):
                got = getattr(m, f'{key}_{kind}')[i].detach().cpu()
                want = g[f'{tag}/{key}_{kind}.{i}']
                assert got.shape == want.shape, (tag, key, kind, i)
                assert (got - want).abs().max().item() <= atol, (tag, key, kind, i)
                # the new factors stay texel-major (the layout the gathers read)
                assert getattr(m, f'{key}_{kind}')[i].permute(0, 2, 3, 1).is_contiguous()


def test_modification_sequence_matches_reference(cuda, g):
    m = _model(cuda, g)
    bb = m.update_alpha_mask(2500)
    vol = m.alpha_mask.alpha_volume.detach().cpu()[0, 0]
    assert torch.equal(vol, g['alpha1/alpha_volume'][0, 0].float())
    assert torch.equal(bb.cpu(), g['alpha1/bounding_box_new'])
    m.shrink_tensor(bb)
    assert m.resolution.tolist() == g['shrunk/resolution'].tolist()
    assert torch.equal(m.bounding_box.cpu(), g['shrunk/bounding_box'])
    _check_factors(m, g, 'shrunk', 0.0)
    m.upsample_model_resolution(3000)
    assert m.resolution.tolist() == g['upsampled/resolution'].tolist()
    assert int(m.num_samples) == int(g['upsampled/num_samples'])
    assert abs(float(m.step_size) - float(g['upsampled/step_size'])) <= 1e-7
    _check_factors(m, g, 'upsampled', 1e-6)
    bb2 = m.update_alpha_mask(2500)
    vol2 = m.alpha_mask.alpha_volume.detach().cpu()[0, 0]
    assert torch.equal(vol2, g['alpha2/alpha_volume'][0, 0].float())
    assert torch.equal(bb2.cpu(), g['alpha2/bounding_box_new'])


def test_run_model_modifications_reconfigures_the_optimiser(cuda, g):
    """run_model_modifications at an upsampling iteration swaps this model's parameter groups in the
    main optimiser for the new parameters and drops the old parameters' Adam state
    (src/models/SimpleTensoRF09.py:917-944)."""
    m = _model(cuda, g)
    opt_cfg = {'name': 'optimizer_main', 'lr_initial_tensor': 0.02, 'lr_initial_network': 1e-3}
    m.configs['optimizers'] = [opt_cfg]
    opt = torch.optim.Adam(m.get_trainable_parameters(opt_cfg), betas=(0.9, 0.99))
    for p in m.parameters():
        p.grad = torch.ones_like(p)
    opt.step()
    m.optimizers = {'optimizer_nerf': opt}
    old = {id(p) for p in m.parameters()}
    m.run_model_modifications(3000)
    assert m.resolution.tolist() != g['initial/resolution'].tolist()
    names = [pg['name'] for pg in opt.param_groups]
    assert sorted(names) == ['coarse_model_network_params', 'coarse_model_tensor_params']
    in_opt = {id(p) for pg in opt.param_groups for p in pg['params']}
    assert in_opt == {id(p) for p in m.parameters() if p.requires_grad}
    tensor_group = next(pg for pg in opt.param_groups if pg['name'] == 'coarse_model_tensor_params')
    assert not any(id(p) in old for p in tensor_group['params'])
    assert all(p not in opt.state for p in tensor_group['params'])


@pytest.mark.parametrize('kind', ['density', 'color'])
@pytest.mark.parametrize('tag,key', [('planes', 'matrices'), ('lines', 'vectors')])
def test_tv_loss_matches_reference(cuda, g, kind, tag, key):
    from srf_hip.tensorf import texel_major_param
    comps = []
    for i in range(3):
        t = g[f'upsampled/{key}_{kind}.{i}']
        comps.append(texel_major_param(t.shape[1], t.shape[2], t.shape[3], t, device=cuda))
    tv = tensorf_mods.tv_loss(comps, 0.7)
    tv.backward()
    want = float(g[f'tv/{kind}_{tag}/loss'])
    assert abs(tv.item() - want) <= 2e-6 * want
    for i in range(3):
        assert torch.allclose(comps[i].grad.cpu(), g[f'tv/{kind}_{tag}/grad.{i}'], rtol=1e-4, atol=1e-9)


def test_entropy_loss_matches_reference(cuda, g):
    w = g['entropy/weights'].to(cuda).requires_grad_(True)
    loss, per_ray = tensorf_mods.entropy_loss(w, g['entropy/mask'].to(cuda), int(g['entropy/num_bins']),
                                              return_maps=True)
    loss.backward()
    want = float(g['entropy/loss'])
    assert abs(loss.item() - want) <= 2e-6 * abs(want)
    assert torch.allclose(w.grad.cpu(), g['entropy/grad'], rtol=1e-5, atol=1e-9)
    assert not per_ray.cpu()[~g['entropy/mask']].any()


def test_upsample_reproduces_linear_ramps_and_identity(cuda):
    """Closed forms: bilinear upsampling with align_corners reproduces a linear function exactly (to
    rounding) and a same-size resample is the identity."""
    from srf_hip.tensorf import texel_major_param
    h, w, c = 7, 9, 5
    y = torch.arange(h).float()[:, None, None]
    x = torch.arange(w).float()[None, :, None]
    ch = torch.arange(c).float()[None, None, :]
    ramp = (0.5 * y - 0.25 * x + ch).permute(2, 0, 1)[None]
    p = texel_major_param(c, h, w, ramp, device=cuda)
    up = tensorf_mods.upsample_factor(p, 13, 17).detach().cpu()
    yy = torch.linspace(0, h - 1, 13)[:, None, None]
    xx = torch.linspace(0, w - 1, 17)[None, :, None]
    want = (0.5 * yy - 0.25 * xx + ch).permute(2, 0, 1)[None]
    assert (up - want).abs().max().item() < 1e-5
    same = tensorf_mods.upsample_factor(p, h, w).detach().cpu()
    assert torch.equal(same, ramp)


def test_regulariser_plugins_match_oracle(cuda):
    """The drop-in TotalVariationLossHip04 and MassConcentrationLossHip06 plugins on a training forward
    of the Simple-TensoRF demo (one points-augmented VM model) against the oracle's restatement of the
    two losses in FP64: loss values, the augmented factors' TV gradients, the weights' entropy gradient
    and the per-ray map (selected rays only, as the reference's map)."""
    import oracle.tensorf_mods as om
    from loss_functions.MassConcentrationLossHip06 import MassConcentrationLossHip
    from loss_functions.TotalVariationLossHip04 import TotalVariationLossHip
    from models.ModelFactory02 import get_model
    from srf_hip import configs as cfgs, synthetic
    torch.manual_seed(0)
    c = cfgs.simple_tensorf_configs(perturb=False, coarse=dict(num_voxels_initial=32 ** 3),
                                    augmented=dict(num_voxels_initial=16 ** 3))
    c['optimizers'] = [{'name': 'optimizer_main', 'lr_decay_ratio': 0.1, 'lr_decay_iters': None}]
    c['num_iterations'] = 25000
    model = get_model(c, model_configs=cfgs.model_configs()).to(cuda)
    model.train()
    batch = synthetic.train_batch(512, seed=3, device=cuda)
    out = model({'pixel_id': batch['pixel_id'], 'num_frames': 3, 'iter_num': 10}, retraw=True)
    inp = {'target_rgb': batch['target_rgb'], 'indices_mask_nerf': batch['indices_mask_nerf'], 'iter_num': 10}
    tv = TotalVariationLossHip(c, {'name': 'TotalVariationLossHip04', 'weight': 1e-2, 'weight_density': 1.0,
                                   'weight_color': 0.5})
    mc = MassConcentrationLossHip(c, {'name': 'MassConcentrationLossHip06', 'num_bins': 5})
    res_tv = tv.compute_loss(inp, out, model, return_loss_maps=True)
    res_mc = mc.compute_loss(inp, out, model, return_loss_maps=True)
    aug = model.augmented_models[0]['coarse_model']
    comps = list(aug.matrices_density) + list(aug.matrices_color)
    g_tv = torch.autograd.grad(res_tv['loss_value'], comps, retain_graph=True)
    w = out['points_augmentation_weights_coarse']
    g_w, = torch.autograd.grad(res_mc['loss_value'], w)

    iw = 0.1 ** (11 / 25000)
    c64 = [p.detach().cpu().double().requires_grad_(True) for p in comps]
    ref_tv = om.tv_loss(c64[:3], iw) * 1.0 + om.tv_loss(c64[3:], iw) * 0.5
    ref_tv.backward()
    assert abs(res_tv['loss_value'].item() - ref_tv.item()) <= 1e-5 * abs(ref_tv.item())
    for a, b in zip(g_tv, c64):
        assert (a.cpu().double() - b.grad).abs().max().item() <= 1e-4 * b.grad.abs().max().item()

    sel = batch['indices_mask_nerf'].cpu()
    w64 = w.detach().cpu().double().requires_grad_(True)
    ref_mc = om.entropy_loss(w64, sel, 5)
    ref_mc.backward()
    assert abs(res_mc['loss_value'].item() - ref_mc.item()) <= 1e-5 * abs(ref_mc.item())
    assert (g_w.cpu().double() - w64.grad).abs().max().item() <= 1e-5 * w64.grad.abs().max().item()
    m = res_mc['loss_maps']['MassConcentrationLoss06_points_augmentation_coarse']
    assert m.shape[0] == int(sel.sum())
    sizes = om.entropy_bins(w64.shape[1], 5)
    binned = torch.stack([b.sum(1) for b in torch.split(w64.detach()[sel], sizes, 1)], 1)
    assert torch.allclose(m.cpu().double(), torch.special.entr(binned + 1e-3).sum(1), rtol=1e-5, atol=1e-7)

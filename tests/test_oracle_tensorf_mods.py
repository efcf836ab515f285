"""The oracle's TensoRF model modifications and regularisers (oracle/tensorf_mods.py) against the
vectors the reference's own VmDecomposedTensor, TotalVariationLoss and MassConcentrationLoss wrote
(tests/golden/tensorf_mods_golden.npz, tests/golden/make_golden.py), plus closed forms."""
import pathlib

import numpy as np
import pytest
import torch

import oracle.tensorf_mods as om

GOLDEN = pathlib.Path(__file__).resolve().parent / 'golden' / 'tensorf_mods_golden.npz'
CFG = dict(threshold=1e-4, upsampling_iters=[2000, 3000], n0=20 ** 3, n1=40 ** 3, per_sample=0.5)


@pytest.fixture(scope='module')
def g():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: torch.from_numpy(z[k]) if z[k].ndim else z[k] for k in z.files}


def _factors(g, tag, kind='density'):
    return ([g[f'{tag}/matrices_{kind}.{i}'] for i in range(3)], [g[f'{tag}/vectors_{kind}.{i}'] for i in range(3)])


def test_alpha_mask_update_matches_reference(g):
    planes, lines = _factors(g, 'initial')
    alpha, bb = om.update_alpha_mask(planes, lines, g['initial/bounding_box'], g['initial/resolution'],
                                     float(g['initial/step_size']), CFG['threshold'])
    assert torch.equal(alpha, g['alpha1/alpha_volume'][0, 0].float())
    assert torch.equal(bb, g['alpha1/bounding_box_new'])
    # the mask is the occupied sub-box dilated by the 3x3x3 pool: a proper box inside the grid
    assert 0 < alpha.sum() < alpha.numel()


def test_shrink_and_upsample_match_reference(g):
    res, bbox = g['initial/resolution'], g['initial/bounding_box']
    t_l, b_r, res_new, bb_new = om.shrink_bounds(g['alpha1/bounding_box_new'], bbox, res, res)
    assert torch.equal(res_new, g['shrunk/resolution'])
    assert torch.equal(bb_new, g['shrunk/bounding_box'])
    for kind in ('density', 'color'):
        planes, lines = _factors(g, 'initial', kind)
        sp, sl = om.shrink_vm_factors(planes, lines, t_l, b_r)
        for i in range(3):
            assert torch.equal(sp[i], g[f'shrunk/matrices_{kind}.{i}'])
            assert torch.equal(sl[i], g[f'shrunk/vectors_{kind}.{i}'])
    n_vox = om.new_num_voxels(3000, CFG['upsampling_iters'], CFG['n0'], CFG['n1'])
    import oracle.tensorf as ot
    res_up = ot.compute_resolution_in_voxels(n_vox, g['shrunk/bounding_box'])
    assert torch.equal(res_up, g['upsampled/resolution'])
    assert ot.compute_num_samples(res_up, CFG['per_sample'], 1e6) == int(g['upsampled/num_samples'])
    for kind in ('density', 'color'):
        up, ul = om.upsample_vm_factors(*_factors(g, 'shrunk', kind), res_up)
        for i in range(3):
            assert torch.allclose(up[i], g[f'upsampled/matrices_{kind}.{i}'], rtol=0, atol=1e-7)
            assert torch.allclose(ul[i], g[f'upsampled/vectors_{kind}.{i}'], rtol=0, atol=1e-7)


def test_second_alpha_update_through_existing_mask(g):
    """The first mask keeps the bounding box it was built in (the grid's before the shrink)."""
    planes, lines = _factors(g, 'upsampled')
    alpha, bb = om.update_alpha_mask(planes, lines, g['upsampled/bounding_box'], g['upsampled/resolution'],
                                     float(g['upsampled/step_size']), CFG['threshold'],
                                     alpha_volume=g['alpha1/alpha_volume'][0, 0].float(),
                                     alpha_bbox=g['initial/bounding_box'])
    assert torch.equal(alpha, g['alpha2/alpha_volume'][0, 0].float())
    assert torch.equal(bb, g['alpha2/bounding_box_new'])


def test_max_pool_dilates_a_single_voxel_to_a_cube():
    """Closed form: one occupied voxel becomes the 3x3x3 cube around it (clipped at the border)."""
    bbox = torch.tensor([[0., 0., 0.], [1., 1., 1.]])
    res = torch.tensor([9, 8, 7])
    import oracle.tensorf as ot
    z = torch.zeros(1, 1, 7, 1)  # lines along z (vector axis 2) for plane 0 (x, y)
    z[0, 0, 3, 0] = 1.0
    planes = [torch.zeros(1, 1, 8, 9), torch.zeros(1, 1, 7, 9), torch.zeros(1, 1, 7, 8)]
    planes[0][0, 0, 2, 5] = 1.0  # y = 2, x = 5
    lines = [z, torch.zeros(1, 1, 8, 1), torch.zeros(1, 1, 9, 1)]
    alpha, bb = om.update_alpha_mask(planes, lines, bbox, res, 1.0, 1e-4)
    want = torch.zeros(7, 8, 9)
    want[2:5, 1:4, 4:7] = 1
    assert torch.equal(alpha, want)
    xs = torch.linspace(0, 1, 9)
    assert torch.allclose(bb[0, 0], xs[4]) and torch.allclose(bb[1, 0], xs[6])


def test_tv_loss_matches_reference_and_closed_form(g):
    for kind in ('density', 'color'):
        for tag, key in (('planes', 'matrices'), ('lines', 'vectors')):
            comps = [g[f'upsampled/{key}_{kind}.{i}'].clone().requires_grad_(True) for i in range(3)]
            tv = om.tv_loss(comps, 0.7)
            tv.backward()
            assert abs(tv.item() - float(g[f'tv/{kind}_{tag}/loss'])) <= 1e-6 * float(g[f'tv/{kind}_{tag}/loss'])
            for i in range(3):
                assert torch.allclose(comps[i].grad, g[f'tv/{kind}_{tag}/grad.{i}'], rtol=1e-5, atol=1e-10)
    # closed form on a ramp: every difference along H is 1, along W is 2
    h, w = 4, 5
    ramp = (torch.arange(h)[:, None] + 2 * torch.arange(w)[None, :]).float()[None, None]
    assert abs(om.tv_loss([ramp], 1.0).item() - 2 * (1.0 + 4.0)) < 1e-6


def test_entropy_loss_matches_reference(g):
    w = g['entropy/weights'].clone().requires_grad_(True)
    loss = om.entropy_loss(w, g['entropy/mask'], int(g['entropy/num_bins']))
    loss.backward()
    assert abs(loss.item() - float(g['entropy/loss'])) <= 1e-6 * abs(float(g['entropy/loss']))
    assert torch.allclose(w.grad, g['entropy/grad'], rtol=1e-5, atol=1e-9)
    sizes = om.entropy_bins(443, 11)
    assert sum(sizes) == 443 and sorted(set(sizes)) == [40, 41]
    # closed form: all weight in one bin of one ray -> entr(1 + 1e-3) + (bins - 1) entr(1e-3)
    one = torch.zeros(1, 22)
    one[0, 0] = 1.0
    e = lambda x: -x * np.log(x)
    assert abs(om.entropy_loss(one, torch.tensor([True]), 11).item() - (e(1.001) + 10 * e(1e-3))) < 1e-5

"""The oracle's losses (oracle/losses.py) pinned on the CPU.

* The depth-consistency losses against vectors written by the reference's own classes
  (tests/golden/loss_golden.npz, tests/golden/make_golden.py from AugmentationsDepthLoss11.py:108-181
  and CoarseFineConsistencyLoss34.py:85-164): loss, both loss maps and both depth gradients, bit for
  bit in FP32.  The vectors show the reference's in-place masking: the second term and the gradient
  of the second depth are zero on every ray.
* The MSE terms and the reprojection by closed forms.
"""
import pathlib

import numpy as np
import pytest
import torch

import oracle.losses as ol

GOLDEN = pathlib.Path(__file__).resolve().parent / 'golden' / 'loss_golden.npz'


@pytest.fixture(scope='module')
def scene():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: torch.from_numpy(z[k]) if z[k].ndim else z[k] for k in z.files}


def _run(scene, name, dtype=torch.float32):
    d1 = scene['depth1'].detach().to(dtype).clone().requires_grad_(True)
    d2 = scene['depth2'].detach().to(dtype).clone().requires_grad_(True)
    out = ol.depth_consistency_nerf(d1, d2, scene['indices_mask_nerf'], scene['rays_o'].to(dtype),
                                    scene['rays_d'].to(dtype), scene['extrinsics'].to(dtype),
                                    scene['images'].to(dtype), scene['pixel_id'], scene['intrinsics'].to(dtype),
                                    tuple(int(v) for v in scene['resolution']),
                                    both_invalid_rule=(name == 'aug_depth'))
    out[0].backward()
    return out, d1.grad, d2.grad


@pytest.mark.parametrize('name', ['aug_depth', 'coarse_fine'])
def test_depth_consistency_matches_reference_vectors(scene, name):
    (loss, m1, m2, k1, k2), g1, g2 = _run(scene, name)
    assert loss.item() == float(scene[f'{name}/f32/loss'])
    assert torch.equal(m1.detach(), scene[f'{name}/f32/map1'])
    assert torch.equal(m2.detach(), scene[f'{name}/f32/map2'])
    assert torch.equal(g1, scene[f'{name}/f32/grad1'])
    assert torch.equal(g2, scene[f'{name}/f32/grad2'])
    # the vectors exercise both gate decisions and the in-place masking of the second term
    assert k1.sum() >= 5 and k2.sum() >= 5
    assert (scene[f'{name}/f32/map1'] > 0).sum() >= 5
    assert not scene[f'{name}/f32/map2'].any() and not scene[f'{name}/f32/grad2'].any()


def test_depth_consistency_float64_close_to_float32_vectors(scene):
    """The restatement is dtype-generic; in FP64 it lands within FP32 rounding of the vectors (the gate
    decisions of this scene are not within rounding of a boundary)."""
    (loss, m1, _, _, _), g1, _ = _run(scene, 'aug_depth', torch.float64)
    assert abs(loss.item() - float(scene['aug_depth/f32/loss'])) < 1e-6 * max(1.0, loss.item())
    assert (g1 - scene['aug_depth/f32/grad1'].double()).abs().max().item() < 1e-6


def test_vectors_reject_symmetric_masking(scene):
    """A restatement without the in-place aliasing (the second term over mask1 alone) misses the vectors:
    the reference's second depth gets no gradient."""
    (_, _, _, k1, _), _, _ = _run(scene, 'aug_depth')
    d1, d2 = scene['depth1'].float()[scene['indices_mask_nerf']], scene['depth2'].float()[scene['indices_mask_nerf']]
    symmetric_map2 = torch.where(k1, (d2 - d1) ** 2, torch.zeros_like(d1))
    assert (symmetric_map2 - scene['aug_depth/f32/map2']).abs().max() > 1e-4


def test_mse_closed_form():
    g = torch.Generator().manual_seed(0)
    p, t = torch.rand(50, 3, generator=g), torch.rand(50, 3, generator=g)
    m = torch.rand(50, generator=g) < 0.6
    want = sum(((p[i] - t[i]) ** 2).sum().item() / 3 for i in range(50) if m[i]) / int(m.sum())
    assert abs(ol.mse(p, t, m).item() - want) < 1e-6
    d, gt = torch.rand(50, generator=g), torch.rand(50, generator=g)
    want = sum((d[i] - gt[i]).item() ** 2 for i in range(50) if m[i]) / int(m.sum())
    assert abs(ol.sparse_depth_mse(d, gt, m).item() - want) < 1e-6
    assert ol.mse(p, t, torch.zeros(50, dtype=torch.bool)).item() == 0.0
    want = sum((d[i] - gt[i]).item() ** 2 for i in range(50) if m[i]) / int(m.sum())
    assert abs(ol.coarse_fine_sparse_depth(d, gt, m).item() - want) < 1e-6


def test_reproject_pinhole_closed_form():
    """A point at depth z in front of a camera (looking down -z, y up) lands at
    (cx + f x / z, cy - f y / z); a point on the camera's optical axis lands on the principal point."""
    f, cx, cy = 50.0, 20.0, 15.0
    K = torch.tensor([[[f, 0, cx], [0, f, cy], [0, 0, 1.]]], dtype=torch.float64)
    th = 0.3
    R = torch.tensor([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]], dtype=torch.float64)
    pose = torch.eye(4, dtype=torch.float64)
    pose[:3, :3] = R
    pose[:3, 3] = torch.tensor([0.5, -0.2, 1.0])
    cam_pts = torch.tensor([[0.3, 0.1, -2.0], [0., 0., -3.0], [-0.4, 0.25, -1.5]], dtype=torch.float64)
    world = (R @ cam_pts[..., None])[..., 0] + pose[:3, 3]
    pos = ol.reproject(world, pose[None].repeat(3, 1, 1), K)
    z = -cam_pts[:, 2]
    want = torch.stack([cx + f * cam_pts[:, 0] / z, cy - f * cam_pts[:, 1] / z], -1)
    assert (pos - want).abs().max().item() < 1e-9
    assert (pos[1] - torch.tensor([cx, cy], dtype=torch.float64)).abs().max().item() < 1e-12


def test_closest_view_is_the_nearest_other_camera():
    E = torch.eye(4).repeat(4, 1, 1)
    E[:, :3, 3] = torch.tensor([[0., 0, 0], [1., 0, 0], [0.3, 0, 0], [5., 0, 0]])
    ids = ol.closest_view_ids(E, torch.tensor([0, 1, 2, 3]))
    assert ids.tolist() == [2, 2, 0, 1]

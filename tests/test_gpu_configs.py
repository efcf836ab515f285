"""Parity at the workload shapes themselves (GPU):

* the plain coarse-only Simple-NeRF FP32 step on 1024 rays (504 x 378 camera, 64 samples, white background);
* the 300^2 VM TensoRF model (density 3 x 16, appearance 3 x 48, features PE L = 2, 64^3 alpha mask at 40 %,
  443 samples) on 8192 rays: a 64-ray subset of the full batch against the FP64 oracle, and the factor and
  network gradients of a 256-ray batch on the full-size grids;
* the uneven-batch TensoRF model (512^2 planes, 24 / 64 components, ragged log-normal sample counts) on
  32,768 rays: a 64-ray subset against the oracle rendering each ray on its own, gradients of a 256-ray
  ragged batch, the gradients of all 32,768 rays against the FP64 oracle's autograd (the oracle run in
  FP64 on the GPU), and the kernels' gradients of a 64-ray batch against FP64 central finite differences.
"""
import pytest
import torch
import torch.nn.functional as F

import oracle.tensorf as ot
import oracle.train_step as ots
from models.ModelFactory02 import get_model
from srf_hip import configs as cfgs, synthetic, tensorf

pytestmark = pytest.mark.gpu

THR = 1e-4


def _rel(a, b):
    return ((a - b).abs().max() / b.abs().max().clamp_min(1e-30)).item()


def _assert_grads(pairs, what, tol=1e-3):
    for name, a, b in pairs:
        e = _rel(a.cpu().double(), b.cpu())
        print(f'{what} {name}: grad rel err {e:.2e}')
        assert e < tol, (name, e)


# FP32 and FP64 may take a discrete decision differently where a sample sits within rounding of its
# threshold: the surface test (weight > 1e-4; FP32 weights carry ~2e-5 relative error), the
# occupancy test (trilinear alpha > 0, exactly 0 on one side of a voxel plane) and the ReLU kink of
# the density.  Such a sample sends its whole gradient in one precision and none in the other to
# the few texels it touches, so the gradient checks leave out the rays that hold one.
def _unstable_rays(weights, ray_of, n_rays, feat=None, valid_mismatch=None):
    bad = torch.zeros(n_rays, dtype=torch.bool)
    near_thr = (weights - THR).abs() < 1e-3 * THR
    bad[ray_of[near_thr]] = True
    if feat is not None:
        bad[ray_of[feat]] = True
    if valid_mismatch is not None:
        bad[ray_of[valid_mismatch]] = True
    return bad


def test_coarse_only_fp32_step_1024_rays(cuda):
    torch.manual_seed(0)
    c = cfgs.simple_nerf_configs(precision='fp32', augmentations=False, fine=False, perturb=False,
                                 raw_noise_std=0.0, white_bkgd=True)
    mc = cfgs.model_configs()
    model = get_model(c, model_configs=mc).to(cuda).train()
    batch = synthetic.train_batch(1024, seed=11, device=cuda)
    out = model({'pixel_id': batch['pixel_id'], 'num_frames': 3, 'iter_num': 0, 'sub_batch_index': 0})
    loss = ((out['rgb_coarse'] - batch['target_rgb']) ** 2).mean()
    loss.backward()
    pid = batch['pixel_id'].cpu()
    K = torch.from_numpy(mc['intrinsics'])[pid[:, 0].long()].double()
    E = torch.from_numpy(mc['extrinsics'])[pid[:, 0].long()].double()
    def oracle(dtype):
        prm = {'coarse': {k: v.detach().cpu().to(dtype).requires_grad_(True)
                          for k, v in model.coarse_model.state_dict().items()}}
        ref = ots.render_train(prm, pid, K.to(dtype), E.to(dtype), mc['resolution'], augment=False, white_bkgd=True)
        ((ref['rgb_coarse'] - batch['target_rgb'].cpu().to(dtype)) ** 2).mean().backward()
        return ref, prm

    ref, p64 = oracle(torch.float64)
    _, p32 = oracle(torch.float32)
    assert (out['rgb_coarse'].detach().cpu().double() - ref['rgb_coarse']).abs().max().item() < 1e-5
    a = out['depth_ndc_coarse'].detach().cpu().double() * (out['acc_coarse'].detach().cpu().double() + 1e-6)
    b = (ref['depth_ndc_coarse'] * (ref['acc_coarse'] + 1e-6)).detach()
    assert (a - b).abs().max().item() < 1e-5
    # gradients of a random-init network are ~1e-8 here: relative error in norm, within 1e-3 or within 3x the
    # FP32 oracle's own error against FP64 (the same bar as the full-model test)
    for n, p in model.coarse_model.named_parameters():
        r = p64['coarse'][n].grad
        ek = ((p.grad.cpu().double() - r).norm() / r.norm().clamp_min(1e-30)).item()
        eo = ((p32['coarse'][n].grad.double() - r).norm() / r.norm().clamp_min(1e-30)).item()
        assert ek < max(1e-3, 3 * eo), (n, ek, eo)


def _vm300(cuda):
    torch.manual_seed(0)
    box = ((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    c = cfgs.simple_tensorf_configs(
        perturb=False, white_bkgd=True, augmentations=False,
        coarse=dict(num_components_density=(16, 16, 16), num_components_color=(48, 48, 48), bounding_box=box,
                    num_voxels_initial=300 ** 3 + 4, num_samples_max=443, features_positional_encoding_degree=2))
    model = get_model(c, model_configs=cfgs.model_configs()).to(cuda).eval()
    from models.SimpleTensoRFHip09 import AlphaGridMask
    g = torch.Generator().manual_seed(7)
    vol = (torch.rand(64, 64, 64, generator=g) < 0.4).float()
    model.coarse_model.alpha_mask = AlphaGridMask(vol, model.coarse_model.bounding_box.cpu()).to(cuda)
    return model, vol


def _vm_params(tm, grad, device='cpu'):
    f64 = lambda t: t.detach().to(device).double().contiguous().requires_grad_(grad)
    return {'bbox': tm.bounding_box.detach().to(device).double(),
            'planes_density': [f64(p) for p in tm.matrices_density], 'lines_density': [f64(p) for p in tm.vectors_density],
            'planes_color': [f64(p) for p in tm.matrices_color], 'lines_color': [f64(p) for p in tm.vectors_color],
            'basis': f64(tm.basis_matrix_color.weight),
            'color_mlp': {k: f64(v) for k, v in tm.color_predictor.state_dict().items()}}


def _vm_oracle(out, prm, vol, rows):
    d = lambda k: out[k].detach().cpu().double()[rows]
    o, dn, z = d('rays_o_ndc'), d('rays_d_ndc'), d('z_vals_coarse')
    return ot.tensorf_render(prm, o[:, None] + dn[:, None] * z[..., None], z, d('rays_o'), d('rays_d'), dn,
                             d('view_dirs'), alpha_volume=vol.double(), distance_scale=25.0, weight_threshold=THR,
                             white_bkgd=True, features_pe_degree=2)


def test_vm300_full_batch_subset_and_gradients(cuda):
    model, vol = _vm300(cuda)
    tm = model.coarse_model
    assert tm.resolution.tolist() == [300, 300, 300] and int(tm.num_samples) == 443
    batch = synthetic.train_batch(8192, seed=12, device=cuda)
    with torch.no_grad():
        out = model({'pixel_id': batch['pixel_id'], 'num_frames': 3}, retraw=True)
    rows = torch.randperm(8192, generator=torch.Generator().manual_seed(13))[:64]
    ref = _vm_oracle(out, _vm_params(tm, False), vol, rows)
    keep = ~((ref['weights'] - THR).abs() < 1e-9).any(-1)
    assert keep.sum() >= 62 and (ref['weights'] > THR).sum() > 2000
    for k in ('rgb', 'acc', 'depth_ndc', 'weights'):
        e = (out[f'{k}_coarse'].cpu().double()[rows][keep] - ref[k][keep]).abs().max().item()
        assert e < 1e-5, (k, e)
    # gradients of a 256-ray batch on the full-size grids
    b2 = synthetic.train_batch(256, seed=14, device=cuda)
    out2 = model({'pixel_id': b2['pixel_id'], 'num_frames': 3}, retraw=True)
    prm = _vm_params(tm, True)
    ref2 = _vm_oracle(out2, prm, vol, torch.arange(256))
    d = lambda k: out2[k].detach().cpu().double()
    pts64 = (d('rays_o_ndc')[:, None] + d('rays_d_ndc')[:, None] * d('z_vals_coarse')[..., None]).reshape(-1, 3)
    pts32 = (out2['rays_o_ndc'][:, None] + out2['rays_d_ndc'][:, None] * out2['z_vals_coarse'][..., None]).reshape(-1, 3)
    in64 = ot.bbox_validity(pts64, prm['bbox'])
    occ64 = torch.zeros_like(in64)
    occ64[in64] = ot.alpha_mask_sample(vol.double(), prm['bbox'], pts64[in64]) > 0
    occ32 = (tensorf.alpha_mask_sample(tm.alpha_mask.kernel_mask(), pts32.detach()) > 0).cpu() & in64
    feat = torch.zeros(pts64.shape[0], dtype=torch.float64)
    feat[occ64] = ot.vm_density_features([p.detach() for p in prm['planes_density']],
                                         [p.detach() for p in prm['lines_density']],
                                         ot.normalize_points(pts64[occ64], prm['bbox']))
    ray_of = torch.arange(256).repeat_interleave(443)
    bad = _unstable_rays(ref2['weights'].detach().reshape(-1), ray_of, 256, feat=occ64 & (feat.abs() < 1e-6),
                         valid_mismatch=occ64 != occ32)
    print('vm300: rays left out of the gradient check:', int(bad.sum()))
    assert bad.sum() <= 8
    gr = torch.randn(256, 3, generator=torch.Generator().manual_seed(15))
    gr[bad] = 0
    (out2['rgb_coarse'] * gr.to(cuda)).sum().backward()
    (ref2['rgb'] * gr.double()).sum().backward()
    pairs = []
    for nm, key in (('matrices_density', 'planes_density'), ('vectors_density', 'lines_density'),
                    ('matrices_color', 'planes_color'), ('vectors_color', 'lines_color')):
        pairs += [(f'{nm}.{i}', x.grad, y.grad) for i, (x, y) in enumerate(zip(getattr(tm, nm), prm[key]))]
    pairs.append(('basis', tm.basis_matrix_color.weight.grad, prm['basis'].grad))
    pairs += [(k, dict(tm.color_predictor.named_parameters())[k].grad, v.grad) for k, v in prm['color_mlp'].items()]
    _assert_grads(pairs, 'vm300')


def _vm512(cuda, color_precision='fp32'):
    torch.manual_seed(0)
    c = cfgs.simple_tensorf_configs(
        augmentations=False, ndc=False, perturb=False, white_bkgd=True,
        coarse=dict(num_components_density=(24, 24, 24), num_components_color=(64, 64, 64),
                    bounding_box=((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)), num_voxels_initial=512 ** 3,
                    features_positional_encoding_degree=2))
    c['model']['color_precision'] = color_precision
    return get_model(c, model_configs=cfgs.model_configs()).coarse_model.to(cuda).eval()


def _ragged(n, step, seed):
    g = torch.Generator().manual_seed(seed)
    counts = (96 * torch.exp(0.5 * torch.randn(n, generator=g))).round().clamp(1, 512).long()
    off = torch.zeros(n + 1, dtype=torch.int64)
    off[1:] = torch.cumsum(counts, 0)
    o = -0.9 + 1.8 * torch.rand(n, 3, generator=g)
    d = F.normalize(torch.randn(n, 3, generator=g), dim=-1)
    ray = torch.repeat_interleave(torch.arange(n), counts)
    k = torch.arange(int(off[-1])) - off[:-1][ray]
    z = step * (k.float() + torch.rand(n, generator=g)[ray])
    return {'pts': o[ray] + d[ray] * z[:, None], 'z_vals': z, 'ray_offsets': off.int(), 'rays_o': o, 'rays_d': d,
            'view_dirs': d, 'max_samples': int(counts.max())}


def _subset(b, rows):
    off = b['ray_offsets'].long()
    segs = [torch.arange(int(off[r]), int(off[r + 1])) for r in rows]
    idx = torch.cat(segs)
    counts = torch.tensor([len(s) for s in segs])
    o2 = torch.zeros(len(rows) + 1, dtype=torch.int64)
    o2[1:] = torch.cumsum(counts, 0)
    return {'pts': b['pts'][idx].double(), 'z_vals': b['z_vals'][idx].double(), 'ray_offsets': o2,
            'rays_o': b['rays_o'][rows].double(), 'rays_d': b['rays_d'][rows].double(),
            'view_dirs': b['view_dirs'][rows].double()}, idx


def test_vm512_ragged_full_batch_subset_and_gradients(cuda):
    tm = _vm512(cuda)
    assert tm.resolution.tolist() == [512, 512, 512]
    step = float(tm.step_size)
    b = _ragged(32768, step, 21)
    inp = {k: (v.to(cuda) if torch.is_tensor(v) else v) for k, v in b.items()}
    with torch.no_grad():
        out = tm.forward_ragged(inp, retraw=True)
    rows = torch.randperm(32768, generator=torch.Generator().manual_seed(22))[:64]
    sub, idx = _subset(b, rows)
    prm = _vm_params(tm, False)
    ref = ot.tensorf_render_ragged(prm, sub['pts'], sub['z_vals'], sub['ray_offsets'], sub['rays_o'], sub['rays_d'],
                                   sub['view_dirs'], distance_scale=25.0, weight_threshold=THR, white_bkgd=True,
                                   features_pe_degree=2)
    assert (ref['weights'] > THR).sum() > 500
    counts = sub['ray_offsets'][1:] - sub['ray_offsets'][:-1]
    keep = ~_unstable_rays(ref['weights'], torch.arange(64).repeat_interleave(counts), 64)
    assert keep.sum() >= 62
    for k in ('rgb', 'acc'):
        assert (out[k].cpu().double()[rows][keep] - ref[k][keep]).abs().max().item() < 1e-5, k
    assert (out['weights'].cpu().double()[idx] - ref['weights']).abs().max().item() < 1e-5
    # gradients of a 256-ray ragged batch on the full-size grids
    b2 = _ragged(256, step, 23)
    inp2 = {k: (v.to(cuda) if torch.is_tensor(v) else v) for k, v in b2.items()}
    out2 = tm.forward_ragged(inp2)
    prm2 = _vm_params(tm, True)
    sub2, _ = _subset(b2, torch.arange(256))
    ref2 = ot.tensorf_render_ragged(prm2, sub2['pts'], sub2['z_vals'], sub2['ray_offsets'], sub2['rays_o'],
                                    sub2['rays_d'], sub2['view_dirs'], distance_scale=25.0, weight_threshold=THR,
                                    white_bkgd=True, features_pe_degree=2)
    counts = (sub2['ray_offsets'][1:] - sub2['ray_offsets'][:-1])
    ray_of = torch.arange(256).repeat_interleave(counts)
    in64 = ot.bbox_validity(sub2['pts'], prm2['bbox'])
    feat = torch.zeros(sub2['pts'].shape[0], dtype=torch.float64)
    feat[in64] = ot.vm_density_features([p.detach() for p in prm2['planes_density']],
                                        [p.detach() for p in prm2['lines_density']],
                                        ot.normalize_points(sub2['pts'][in64], prm2['bbox']))
    bad = _unstable_rays(ref2['weights'].detach(), ray_of, 256, feat=in64 & (feat.abs() < 1e-6))
    print('vm512 ragged: rays left out of the gradient check:', int(bad.sum()))
    assert bad.sum() <= 8
    gr = torch.randn(256, 3, generator=torch.Generator().manual_seed(24))
    gr[bad] = 0
    (out2['rgb'] * gr.to(cuda)).sum().backward()
    (ref2['rgb'] * gr.double()).sum().backward()
    pairs = []
    for nm, key in (('matrices_density', 'planes_density'), ('vectors_density', 'lines_density'),
                    ('matrices_color', 'planes_color'), ('vectors_color', 'lines_color')):
        pairs += [(f'{nm}.{i}', x.grad, y.grad) for i, (x, y) in enumerate(zip(getattr(tm, nm), prm2[key]))]
    pairs += [(k, dict(tm.color_predictor.named_parameters())[k].grad, v.grad) for k, v in prm2['color_mlp'].items()]
    _assert_grads(pairs, 'vm512 ragged')


def _ragged_dev(b, device, dtype):
    """The batch on `device`, floating-point entries in `dtype` (the offsets stay int32)."""
    return {k: ((v.to(device=device, dtype=dtype) if v.is_floating_point() else v.to(device))
                if torch.is_tensor(v) else v) for k, v in b.items()}


def _grad_pairs(tm, prm):
    pairs = []
    for nm, key in (('matrices_density', 'planes_density'), ('vectors_density', 'lines_density'),
                    ('matrices_color', 'planes_color'), ('vectors_color', 'lines_color')):
        pairs += [(f'{nm}.{i}', x.grad, y.grad) for i, (x, y) in enumerate(zip(getattr(tm, nm), prm[key]))]
    pairs.append(('basis', tm.basis_matrix_color.weight.grad, prm['basis'].grad))
    pairs += [(k, dict(tm.color_predictor.named_parameters())[k].grad, v.grad) for k, v in prm['color_mlp'].items()]
    return pairs


def _density_kink(prm, pts):
    """Samples whose density feature sits at the ReLU kink (|feature| < 1e-6), as a mask over pts."""
    valid = ot.bbox_validity(pts, prm['bbox'])
    feat = torch.zeros(pts.shape[0], dtype=pts.dtype, device=pts.device)
    feat[valid] = ot.vm_density_features([p.detach() for p in prm['planes_density']],
                                         [p.detach() for p in prm['lines_density']],
                                         ot.normalize_points(pts[valid], prm['bbox']))
    return valid & (feat.abs() < 1e-6)


@pytest.mark.parametrize('color_precision', ['fp32', 'bf16_split'])
def test_vm512_ragged_all_rays_against_fp64_autograd(cuda, color_precision):
    """The 32,768-ray uneven batch, forward and hand-written backward on the kernels, against the
    oracle's autograd in FP64 (run on the GPU for speed): rgb of every stable ray within 1e-5 and the
    gradients of all factors and networks within 1e-3 relative; with the colour network on FP32 and on
    split-BF16 operands."""
    tm = _vm512(cuda, color_precision)
    b = _ragged(32768, float(tm.step_size), 21)
    out = tm.forward_ragged(_ragged_dev(b, cuda, torch.float32), retraw=True)
    prm = _vm_params(tm, True, device=cuda)
    sub = _ragged_dev(b, cuda, torch.float64)
    ref = ot.tensorf_render_ragged(prm, sub['pts'], sub['z_vals'], b['ray_offsets'], sub['rays_o'], sub['rays_d'],
                                   sub['view_dirs'], distance_scale=25.0, weight_threshold=THR, white_bkgd=True,
                                   features_pe_degree=2)
    counts = (b['ray_offsets'][1:] - b['ray_offsets'][:-1]).long()
    ray_of = torch.arange(32768).repeat_interleave(counts)
    # Besides the samples near the thresholds, leave out the rays where the kernels' FP32 arithmetic and
    # FP64 take a decision differently: the surface test (where alpha = 1 - exp(-sigma delta) is tiny its
    # FP32 rounding is relative to sigma delta, far wider than the band) and the density ReLU (a density
    # feature is a sum of 3 x 24 products with cancellation; its FP32 value can be a few 1e-6 off, beyond
    # the 1e-6 kink band).  The reference's own FP32 arithmetic flips the same samples.
    w64 = ref['weights'].detach().cpu()
    flips = (out['weights'].detach().cpu() > THR) != (w64 > THR)
    relu_flips = (out['raw_sigma'].detach().cpu().flatten() > 0) != (ref['raw_sigma'].detach().cpu().flatten() > 0)
    bad = _unstable_rays(w64, ray_of, 32768, feat=_density_kink(prm, sub['pts']).cpu(),
                         valid_mismatch=flips | relu_flips)
    print('vm512 all rays: rays left out:', int(bad.sum()), 'with a flipped surface decision:',
          int(torch.unique(ray_of[flips]).numel()), 'with a flipped density ReLU:',
          int(torch.unique(ray_of[relu_flips]).numel()))
    assert bad.sum() <= 32768 // 50
    keep = ~bad
    # rgb against FP64, ray by ray: within 1e-5, or as close as the reference's own FP32 arithmetic
    # (the oracle in FP32) gets on that ray; the transmittance of a 512-sample ray accumulates ~512
    # FP32 roundings
    with torch.no_grad():
        p32 = {k: ([t.float() for t in v] if isinstance(v, list) else
                   ({n: t.float() for n, t in v.items()} if isinstance(v, dict) else v.float()))
               for k, v in prm.items()}
        s32 = _ragged_dev(b, cuda, torch.float32)
        ref32 = ot.tensorf_render_ragged(p32, s32['pts'], s32['z_vals'], b['ray_offsets'], s32['rays_o'],
                                         s32['rays_d'], s32['view_dirs'], distance_scale=25.0, weight_threshold=THR,
                                         white_bkgd=True, features_pe_degree=2)
    r64 = ref['rgb'].detach().cpu()
    ek = (out['rgb'].detach().cpu().double() - r64).abs().amax(-1)[keep]
    eo = (ref32['rgb'].cpu().double() - r64).abs().amax(-1)[keep]
    worst = ek.argmax()
    print(f'vm512 all rays: rgb max abs err {ek.max().item():.2e} (FP32 oracle {eo.max().item():.2e}); worst ray '
          f'{ek[worst].item():.2e} against the FP32 oracle\'s {eo[worst].item():.2e} there; '
          f'{int((ek > 1e-5).sum())} rays above 1e-5')
    assert bool((ek <= torch.clamp(2 * eo, min=1e-5)).all())
    gr = torch.randn(32768, 3, generator=torch.Generator().manual_seed(26))
    gr[bad] = 0
    (out['rgb'] * gr.to(cuda)).sum().backward()
    (ref['rgb'] * gr.to(cuda).double()).sum().backward()
    _assert_grads(_grad_pairs(tm, prm), 'vm512 all rays')


def test_vm512_ragged_bf16_colour_rgb_against_fp64(cuda):
    """The uneven batch as the benchmark runs it, the colour network on BF16 operands: rgb of every
    stable ray of the 32,768 within 2e-3 of the FP64 oracle (the bar for BF16 renders)."""
    tm = _vm512(cuda, 'bf16')
    b = _ragged(32768, float(tm.step_size), 21)
    with torch.no_grad():
        out = tm.forward_ragged(_ragged_dev(b, cuda, torch.float32), retraw=True)
        prm = _vm_params(tm, False, device=cuda)
        sub = _ragged_dev(b, cuda, torch.float64)
        ref = ot.tensorf_render_ragged(prm, sub['pts'], sub['z_vals'], b['ray_offsets'], sub['rays_o'],
                                       sub['rays_d'], sub['view_dirs'], distance_scale=25.0, weight_threshold=THR,
                                       white_bkgd=True, features_pe_degree=2)
    counts = (b['ray_offsets'][1:] - b['ray_offsets'][:-1]).long()
    ray_of = torch.arange(32768).repeat_interleave(counts)
    w64 = ref['weights'].cpu()
    flips = (out['weights'].cpu() > THR) != (w64 > THR)
    bad = _unstable_rays(w64, ray_of, 32768, valid_mismatch=flips)
    assert bad.sum() <= 32768 // 50
    err = (out['rgb'].cpu().double() - ref['rgb'].cpu()).abs().amax(-1)[~bad]
    print(f'vm512 BF16 colour: rgb max abs err {err.max().item():.2e} over {int((~bad).sum())} rays, '
          f'median {err.median().item():.2e}')
    assert err.max().item() < 2e-3


@pytest.mark.parametrize('color_precision', ['fp32', 'bf16_split'])
def test_vm512_ragged_gradients_against_fp64_finite_differences(cuda, color_precision):
    """The kernels' gradients on a 64-ray uneven batch against FP64 central differences of the
    oracle's loss: for each of seven parameter tensors (density and appearance planes and lines, the
    basis, the colour network's first weight and last bias) the four entries with the largest gradient;
    the colour network on FP32 and on split-BF16 operands."""
    tm = _vm512(cuda, color_precision)
    b = _ragged(64, float(tm.step_size), 25)
    prm = _vm_params(tm, False)
    sub = _ragged_dev(b, 'cpu', torch.float64)
    render = lambda p: ot.tensorf_render_ragged(p, sub['pts'], sub['z_vals'], b['ray_offsets'], sub['rays_o'],
                                                sub['rays_d'], sub['view_dirs'], distance_scale=25.0,
                                                weight_threshold=THR, white_bkgd=True, features_pe_degree=2)
    ref = render(prm)
    counts = (b['ray_offsets'][1:] - b['ray_offsets'][:-1]).long()
    bad = _unstable_rays(ref['weights'], torch.arange(64).repeat_interleave(counts), 64,
                         feat=_density_kink(prm, sub['pts']))
    assert bad.sum() <= 2
    gr = torch.randn(64, 3, generator=torch.Generator().manual_seed(27), dtype=torch.float64)
    gr[bad] = 0
    out = tm.forward_ragged(_ragged_dev(b, cuda, torch.float32))
    (out['rgb'] * gr.to(cuda).float()).sum().backward()
    loss = lambda p: (render(p)['rgb'] * gr).sum().item()
    cases = [('planes_density', 0, tm.matrices_density[0]), ('lines_density', 1, tm.vectors_density[1]),
             ('planes_color', 2, tm.matrices_color[2]), ('lines_color', 0, tm.vectors_color[0]),
             ('basis', None, tm.basis_matrix_color.weight), ('color_mlp', 'mlp.0.weight', None),
             ('color_mlp', 'mlp.4.bias', None)]
    mlp = dict(tm.color_predictor.named_parameters())
    for key, i, param in cases:
        if key == 'color_mlp':
            param = mlp[i]
            target = prm['color_mlp'][i]
        elif i is None:
            target = prm[key]
        else:
            target = prm[key][i]
        g = param.grad.detach().cpu().double().contiguous()
        top = torch.topk(g.abs().flatten(), min(4, g.numel())).indices
        scale = g.abs().max().item()
        assert scale > 0, key
        for flat in top.tolist():
            pos = target.view(-1)
            v0 = pos[flat].item()
            h = 1e-6 * max(1.0, abs(v0))
            pos[flat] = v0 + h
            lp = loss(prm)
            pos[flat] = v0 - h
            lm = loss(prm)
            pos[flat] = v0
            fd = (lp - lm) / (2 * h)
            e = abs(g.view(-1)[flat].item() - fd) / scale
            print(f'fd {key}[{i}] entry {flat}: kernel {g.view(-1)[flat].item():.6e} fd {fd:.6e} rel {e:.1e}')
            assert e < 1e-3, (key, i, flat, e)


def test_vm150_low_rank_augmented_model_subset_and_gradients(cuda):
    """The low-rank points-augmented model of the 300^2 configuration (4 density and 48 / 12 / 12 appearance
    components per plane at 150^2, no alpha mask, the main model's 443 samples): its renders of a 64-ray subset
    of 8192 rays within 1e-5 of the FP64 oracle, and the gradients of all its factors and networks on a 256-ray
    batch within 1e-3."""
    torch.manual_seed(0)
    box = ((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    c = cfgs.simple_tensorf_configs(
        perturb=False, white_bkgd=True,
        coarse=dict(num_components_density=(16, 16, 16), num_components_color=(48, 48, 48), bounding_box=box,
                    num_voxels_initial=300 ** 3 + 4, num_samples_max=443, features_positional_encoding_degree=2),
        augmented=dict(num_components_density=(4, 4, 4), bounding_box=box, num_voxels_initial=150 ** 3 + 4,
                       num_samples_max=443))
    # training mode: the augmented models render in training only (white background and deterministic
    # sampling, so the forward is a function of the inputs)
    model = get_model(c, model_configs=cfgs.model_configs()).to(cuda).train()
    am = model.augmented_models[0]['coarse_model']
    assert am.resolution.tolist() == [150, 150, 150]
    assert [p.shape[1] for p in am.matrices_density] == [4, 4, 4]

    def oracle(out, prm, rows):
        d = lambda k: out[k].detach().cpu().double()[rows]
        o, dn, z = d('rays_o_ndc'), d('rays_d_ndc'), d('z_vals_coarse')
        return ot.tensorf_render(prm, o[:, None] + dn[:, None] * z[..., None], z, d('rays_o'), d('rays_d'), dn,
                                 d('view_dirs'), distance_scale=25.0, weight_threshold=THR, white_bkgd=True)

    batch = synthetic.train_batch(8192, seed=16, device=cuda)
    with torch.no_grad():
        out = model({'pixel_id': batch['pixel_id'], 'num_frames': 3}, retraw=True)
    rows = torch.randperm(8192, generator=torch.Generator().manual_seed(17))[:64]
    ref = oracle(out, _vm_params(am, False), rows)
    keep = ~((ref['weights'] - THR).abs() < 1e-9).any(-1)
    assert keep.sum() >= 62 and (ref['weights'] > THR).sum() > 1000
    for k in ('rgb', 'acc', 'weights'):
        e = (out[f'points_augmentation_{k}_coarse'].cpu().double()[rows][keep] - ref[k][keep]).abs().max().item()
        assert e < 1e-5, (k, e)
    # gradients of a 256-ray batch on the full-size grids
    b2 = synthetic.train_batch(256, seed=18, device=cuda)
    out2 = model({'pixel_id': b2['pixel_id'], 'num_frames': 3}, retraw=True)
    prm = _vm_params(am, True)
    ref2 = oracle(out2, prm, torch.arange(256))
    w32 = out2['points_augmentation_weights_coarse'].detach().cpu().reshape(-1)
    w64 = ref2['weights'].detach().reshape(-1)
    d = lambda k: out2[k].detach().cpu().double()
    pts64 = (d('rays_o_ndc')[:, None] + d('rays_d_ndc')[:, None] * d('z_vals_coarse')[..., None]).reshape(-1, 3)
    ray_of = torch.arange(256).repeat_interleave(out2['z_vals_coarse'].shape[1])
    bad = _unstable_rays(w64, ray_of, 256, feat=_density_kink(prm, pts64),
                         valid_mismatch=(w32 > THR) != (w64 > THR))
    print('vm150 augmented: rays left out of the gradient check:', int(bad.sum()))
    # 4 density components give small features, so more samples sit near the ReLU kink than at 300^2
    assert bad.sum() <= 12
    gr = torch.randn(256, 3, generator=torch.Generator().manual_seed(19))
    gr[bad] = 0
    (out2['points_augmentation_rgb_coarse'] * gr.to(cuda)).sum().backward()
    (ref2['rgb'] * gr.double()).sum().backward()
    _assert_grads(_grad_pairs(am, prm), 'vm150 augmented')

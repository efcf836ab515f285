"""End-to-end parity of the drop-in SimpleNeRFHip model with the oracle render (GPU)."""
import pytest
import torch

import oracle.train_step as ots
from models.ModelFactory02 import get_model
from srf_hip import configs as cfgs, synthetic

pytestmark = pytest.mark.gpu


def _setup(cuda, precision, n_rays=512):
    torch.manual_seed(0)
    c = cfgs.simple_nerf_configs(precision=precision, perturb=False, raw_noise_std=0.0)
    mc = cfgs.model_configs()
    model = get_model(c, model_configs=mc).to(cuda)
    model.train()
    batch = synthetic.train_batch(n_rays, seed=2, device=cuda)
    inp = {'pixel_id': batch['pixel_id'], 'num_frames': 3, 'iter_num': 0, 'sub_batch_index': 0}
    nets = {'coarse': model.coarse_model, 'fine': model.fine_model,
            'points_augmentation': model.augmented_models[0]['coarse_model'],
            'views_augmentation': model.augmented_models[1]['coarse_model']}
    pid = batch['pixel_id'].cpu()
    K = torch.from_numpy(mc['intrinsics'])[pid[:, 0].long()]
    E = torch.from_numpy(mc['extrinsics'])[pid[:, 0].long()]
    return model, nets, batch, inp, pid, K, E, mc


KEYS = ['rgb_coarse', 'depth_coarse', 'depth_ndc_coarse', 'acc_coarse', 'rgb_fine', 'depth_fine',
        'points_augmentation_rgb_coarse', 'points_augmentation_depth_coarse',
        'views_augmentation_rgb_coarse', 'views_augmentation_depth_coarse']


def test_train_render_and_grads_match_oracle_fp32(cuda):
    model, nets, batch, inp, pid, K, E, mc = _setup(cuda, 'fp32')
    out = model(inp)
    loss = synthetic.train_loss(out, batch)
    loss.backward()
    params = {k: {n: p.detach().cpu().double().requires_grad_(True) for n, p in m.state_dict().items()}
              for k, m in nets.items()}
    ref = ots.render_train(params, pid, K.double(), E.double(), mc['resolution'])
    for k in KEYS:
        a, b = out[k].detach().cpu().double(), ref[k]
        if 'depth' in k:
            # depth = sum(w z) / (acc + 1e-6) is ill-conditioned where acc ~ 1e-6 (random-init
            # networks): compare the numerator sum(w z), which is what the kernels accumulate
            acc_k = k.replace('depth_ndc', 'acc').replace('depth', 'acc')
            wa, wb = out[acc_k].detach().cpu().double() + 1e-6, ref[acc_k] + 1e-6
            a, b = a * wa, b * wb
            tol = 2e-4 * max(1.0, b.abs().max().item())
        else:
            tol = 1e-4
        assert (a - b).abs().max().item() < tol, (k, (a - b).abs().max().item())
    cpu_batch = {k: v.cpu() for k, v in batch.items()}

    def oracle_grads(dtype):
        ps = {k: {n: p.detach().cpu().to(dtype).requires_grad_(True) for n, p in m.state_dict().items()}
              for k, m in nets.items()}
        o = ots.render_train(ps, pid, K.to(dtype), E.to(dtype), mc['resolution'])
        b = {**cpu_batch, 'target_rgb': cpu_batch['target_rgb'].to(dtype), 'sd_depth': cpu_batch['sd_depth'].to(dtype)}
        lo = synthetic.train_loss(o, b)
        lo.backward()
        return lo, ps

    ref_loss, p64 = oracle_grads(torch.float64)
    _, p32 = oracle_grads(torch.float32)
    assert abs(loss.item() - ref_loss.item()) < 1e-4 * max(1.0, abs(ref_loss.item()))
    # The depth losses differentiate sum(w z)/(acc + 1e-6), which is ill-conditioned where acc is
    # ~1e-6: the kernels' gradients must be as close to FP64 as the FP32 oracle's own (and within
    # 1e-3 wherever the FP32 oracle is).
    for key, m in nets.items():
        ek = eo = 0.0
        for n, p in m.named_parameters():
            r = p64[key][n].grad
            ek = max(ek, ((p.grad.cpu().double() - r).norm() / r.norm().clamp_min(1e-12)).item())
            eo = max(eo, ((p32[key][n].grad.double() - r).norm() / r.norm().clamp_min(1e-12)).item())
        print(f'{key}: kernel grad rel err {ek:.2e}, fp32 oracle {eo:.2e}')
        assert ek < max(1e-3, 3 * eo), (key, ek, eo)


def test_bf16_train_render_close_to_fp64(cuda):
    model, nets, batch, inp, pid, K, E, mc = _setup(cuda, 'bf16', n_rays=1024)
    with torch.no_grad():
        out = model(inp)
    params = {k: {n: p.detach().cpu().double() for n, p in m.state_dict().items()} for k, m in nets.items()}
    ref = ots.render_train(params, pid, K.double(), E.double(), mc['resolution'])
    err = max((out[k].cpu().double() - ref[k]).abs().max().item() for k in
              ('rgb_coarse', 'rgb_fine', 'points_augmentation_rgb_coarse', 'views_augmentation_rgb_coarse'))
    print('bf16 rgb max abs error vs fp64:', err)
    assert err < 2e-3

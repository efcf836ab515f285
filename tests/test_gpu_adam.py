"""The fused HIP Adam (srf_hip.adam.AdamHip, csrc/adam.hip) against torch.optim.Adam, the optimiser
the reference builds (src/optimizers/OptimizerFactory02.py:9-21): several steps over two parameter
groups (one with weight decay), odd sizes, a texel-major (permuted dense) parameter and more than
32 tensors per step; optimizer state dicts interchange with torch's Adam in both directions."""
import pytest
import torch

from srf_hip.adam import AdamHip

pytestmark = pytest.mark.gpu

SHAPES = [(257, 31), (1000,), (3,), (64, 64, 4)] + [(17 + k,) for k in range(36)]


def _params(device, dtype=torch.float32, seed=0):
    g = torch.Generator().manual_seed(seed)
    ps = [torch.randn(s, generator=g, dtype=torch.float64) for s in SHAPES]
    # a texel-major factor: [1, c, H, W] view of [H, W, c] storage
    tm = torch.randn(1, 9, 11, 5, generator=g, dtype=torch.float64).permute(0, 2, 3, 1).contiguous().permute(0, 3, 1, 2)
    ps.append(tm)
    return [torch.nn.Parameter(p.to(device=device, dtype=dtype)) for p in ps]


def _groups(params):
    return [{'params': params[:20], 'lr': 2e-2}, {'params': params[20:], 'lr': 1e-3, 'weight_decay': 1e-2}]


def _grads(params, step, device, dtype):
    """Seeded gradients in each parameter's own layout (as autograd leaves them)."""
    g = torch.Generator().manual_seed(100 + step)
    for p in params:
        p.grad = torch.empty_like(p).copy_(torch.randn(p.shape, generator=g, dtype=torch.float64))


def test_adam_matches_torch(cuda):
    ours = _params(cuda)
    ref32 = _params(cuda)
    ref64 = _params('cpu', torch.float64)
    assert ours[-1].stride() != ours[-1].contiguous().stride()
    opt = AdamHip(_groups(ours), betas=(0.9, 0.99))
    o32 = torch.optim.Adam(_groups(ref32), betas=(0.9, 0.99), foreach=False)
    o64 = torch.optim.Adam(_groups(ref64), betas=(0.9, 0.99), foreach=False)
    for step in range(5):
        for ps, o, dev, dt in ((ours, opt, cuda, torch.float32), (ref32, o32, cuda, torch.float32),
                               (ref64, o64, 'cpu', torch.float64)):
            _grads(ps, step, dev, dt)
            o.step()
    torch.cuda.synchronize()
    worst = 0.0
    for a, b, c in zip(ours, ref32, ref64):
        ea = (a.detach().cpu().double() - c.detach()).abs().max().item()
        eb = (b.detach().cpu().double() - c.detach()).abs().max().item()
        worst = max(worst, ea)
        # as close to FP64 as torch's own FP32 Adam (a few ulp of the parameters)
        assert ea <= 2 * eb + 1e-6, (ea, eb)
    print('AdamHip max |p - p_fp64| after 5 steps:', worst)
    for a, b in zip(ours, ref32):
        assert int(opt.state[a]['step'].item()) == int(o32.state[b]['step'].item()) == 5


def test_state_dicts_interchange_with_torch_adam(cuda):
    a = _params(cuda)
    b = _params(cuda)
    opt = AdamHip(_groups(a), betas=(0.9, 0.99))
    tor = torch.optim.Adam(_groups(b), betas=(0.9, 0.99), foreach=False)
    for step in range(2):
        _grads(a, step, cuda, torch.float32)
        _grads(b, step, cuda, torch.float32)
        opt.step()
        tor.step()
    # swap the optimisers' states: each continues from the other's
    sd_ours, sd_torch = opt.state_dict(), tor.state_dict()
    opt2 = AdamHip(_groups(b), betas=(0.9, 0.99))
    opt2.load_state_dict(sd_torch)
    tor2 = torch.optim.Adam(_groups(a), betas=(0.9, 0.99), foreach=False)
    tor2.load_state_dict(sd_ours)
    _grads(a, 2, cuda, torch.float32)
    _grads(b, 2, cuda, torch.float32)
    opt2.step()
    tor2.step()
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert (x.detach() - y.detach()).abs().max().item() <= 1e-5

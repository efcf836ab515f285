"""GPU parity of the fused TensoRF colour network (basis + PE + MLP, forward and backward) against
the oracle (FP64 autograd of basis matrix + MlpFeaturesColorPredictor)."""
import pytest
import torch
import torch.nn.functional as F

import oracle.tensorf as ot
from srf_hip import color

pytestmark = pytest.mark.gpu

CASES = {
    'shipped': dict(C=72, Lf=0, Lv=0, view=True),       # [48,12,12] components, PE 0 (demo config)
    'vm300': dict(C=144, Lf=2, Lv=0, view=True),        # 3 x 48 components, features PE L=2
    'no_view': dict(C=48, Lf=1, Lv=0, view=False),
    'view_pe': dict(C=16, Lf=0, Lv=2, view=True),
}


def params(c, seed):
    g = torch.Generator().manual_seed(seed)
    d1 = 27 * (2 * c['Lf'] + 1) + (3 * (2 * c['Lv'] + 1) if c['view'] else 0)
    lin = lambda o, i: ((torch.rand(o, i, generator=g) * 2 - 1) / i ** 0.5, (torch.rand(o, generator=g) * 2 - 1) * 0.1)
    basis = (torch.rand(27, c['C'], generator=g) * 2 - 1) / c['C'] ** 0.5
    w1, b1 = lin(128, d1)
    w2, b2 = lin(128, 128)
    w3, b3 = lin(3, 128)
    return [basis, w1, b1, w2, b2, w3, b3]


def oracle(c, prm, app, dirs):
    basis, w1, b1, w2, b2, w3, b3 = prm
    feats = F.linear(app, basis)
    mlp = {'mlp.0.weight': w1, 'mlp.0.bias': b1, 'mlp.2.weight': w2, 'mlp.2.bias': b2, 'mlp.4.weight': w3,
           'mlp.4.bias': b3}
    return ot.color_mlp_forward(mlp, feats, dirs if c['view'] else None, c['Lf'], c['Lv'])


@pytest.mark.parametrize('case', list(CASES))
def test_color_forward_backward(cuda, case):
    c = CASES[case]
    g = torch.Generator().manual_seed(1)
    n_rays, S = 40, 37
    m = 1000                                   # surface rows (not a multiple of the wave width)
    idx = torch.sort(torch.randperm(n_rays * S, generator=g)[:m])[0].to(torch.int32)
    app = torch.randn(m, c['C'], generator=g) * 0.3
    vd = F.normalize(torch.randn(n_rays, 3, generator=g), dim=-1)
    dirs = vd[idx.long() // S]
    prm = params(c, 2)
    gout = torch.randn(m, 3, generator=g)
    p64 = [t.double().requires_grad_(True) for t in prm]
    a64 = app.double().requires_grad_(True)
    ref = oracle(c, p64, a64, dirs.double())
    (ref * gout.double()).sum().backward()
    spec = color.ColorSpec(n_features=c['C'], features_pe=c['Lf'], use_view_dirs=c['view'], views_pe=c['Lv'])
    pd = [t.to(cuda).requires_grad_(True) for t in prm]
    ad = app.to(cuda).requires_grad_(True)
    rgb = color.color_forward(spec, ad, pd[0], pd[1:], view_dirs=vd.to(cuda), idx=idx.to(cuda), n_samples=S)
    (rgb * gout.to(cuda)).sum().backward()
    torch.cuda.synchronize()
    err = (rgb.detach().cpu().double() - ref.detach()).abs().max().item()
    assert err < 1e-5, err
    ea = ((ad.grad.cpu().double() - a64.grad).abs().max() / a64.grad.abs().max()).item()
    assert ea < 1e-4, ea
    for name, a, b in zip(['basis', 'w1', 'b1', 'w2', 'b2', 'w3', 'b3'], pd, p64):
        e = ((a.grad.cpu().double() - b.grad).abs().max() / b.grad.abs().max()).item()
        assert e < 1e-4, (name, e)


def test_color_inference_matches_training(cuda):
    c = CASES['vm300']
    g = torch.Generator().manual_seed(3)
    app = torch.randn(700, c['C'], generator=g).to(cuda)
    vd = F.normalize(torch.randn(700, 3, generator=g), dim=-1).to(cuda)
    prm = [t.to(cuda) for t in params(c, 4)]
    spec = color.ColorSpec(n_features=c['C'], features_pe=c['Lf'])
    with torch.no_grad():
        a = color.color_forward(spec, app, prm[0], prm[1:], view_dirs=vd)
    pr = [t.clone().requires_grad_(True) for t in prm]
    b = color.color_forward(spec, app, pr[0], pr[1:], view_dirs=vd)
    assert torch.equal(a, b.detach())

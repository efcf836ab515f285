"""GPU parity of the fused PE + MLP kernels (forward and backward) against the oracle."""
import pytest
import torch

import oracle.nerf as on
from srf_hip import mlp as hm

pytestmark = pytest.mark.gpu

VARIANTS = {
    'main': dict(spec=dict(), oracle=dict()),
    'points_aug': dict(spec=dict(sigma_pe_degree=3), oracle=dict(sigma_pe_degree=3)),
    'views_aug': dict(spec=dict(view_dependent=False), oracle=dict(use_view_dirs=False, view_dependent_rgb=False)),
}


def make_params(spec, seed, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    ws, bs = [], []
    for s in spec.param_shapes():
        if s is None:
            ws.append(None)
            bs.append(None)
            continue
        fan_in = s[0][1]
        bound = (6.0 / (fan_in + s[0][0])) ** 0.5 * scale       # Xavier-uniform
        ws.append((torch.rand(s[0], generator=g) * 2 - 1) * bound)
        bs.append((torch.rand(s[1], generator=g) * 2 - 1) * 0.05)
    return ws, bs


def as_dict(ws, bs):
    d = {}
    for i, name in enumerate(hm.PARAM_NAMES):
        if ws[i] is not None:
            d[name + '.weight'] = ws[i]
            d[name + '.bias'] = bs[i]
    return d


def points(n, seed):
    g = torch.Generator().manual_seed(seed)
    pts = torch.rand(n, 3, generator=g) * 2 - 1
    pts[:, 2] = torch.rand(n, generator=g)          # NDC depth in [0,1]
    dirs = torch.nn.functional.normalize(torch.randn(n, 3, generator=g), dim=-1)
    return pts, dirs


def run_oracle(variant, params, pts, dirs, noise=None, dtype=torch.float64):
    p = {k: v.to(dtype) for k, v in params.items()}
    # the reference's sigma noise has the shape of the raw sigma, [P, 1] (SimpleNeRF17.py:740)
    return on.mlp_forward(p, pts.to(dtype), dirs.to(dtype), noise=None if noise is None else noise.to(dtype)[:, None],
                          **VARIANTS[variant]['oracle'])


@pytest.mark.parametrize('variant', list(VARIANTS))
@pytest.mark.parametrize('prec,tol_s,tol_c', [('fp32', 2e-5, 1e-5), ('bf16_split', 2e-5, 1e-5),
                                            ('bf16', 3e-2, 1e-2)])
def test_forward(cuda, variant, prec, tol_s, tol_c):
    spec = hm.MLPSpec(prec=prec, **VARIANTS[variant]['spec'])
    ws, bs = make_params(spec, 1)
    n = 1000                                   # not a multiple of 64: exercises the tail wave
    pts, dirs = points(n, 2)
    noise = torch.randn(n, generator=torch.Generator().manual_seed(3)) * 0.1
    sig, rgb = hm.mlp_forward(spec, [None if w is None else w.to(cuda) for w in ws],
                              [None if b is None else b.to(cuda) for b in bs], pts.to(cuda), dirs.to(cuda),
                              noise.to(cuda))
    ref = run_oracle(variant, as_dict(ws, bs), pts, dirs, noise)
    es = (sig.cpu().double() - ref['sigma'][:, 0]).abs().max().item()
    ec = (rgb.cpu().double() - ref['rgb']).abs().max().item()
    assert es < tol_s * max(1.0, ref['sigma'].abs().max().item()), es
    assert ec < tol_c, ec


@pytest.mark.parametrize('variant', list(VARIANTS))
@pytest.mark.parametrize('prec,tol', [('fp32', 1e-4), ('bf16_split', 1e-4), ('bf16', 0.3)])
def test_backward(cuda, variant, prec, tol):
    """Gradients vs the FP64 oracle.  FP32 operands: within 1e-4 (observed ~1e-6).  Split BF16
    (FP32-level forward, two-plane backward): within 1e-4 as well.  BF16 operands:
    the error of BF16 training itself (~4% at the heads growing to ~15% at layer 0 over ten
    chained roundings), reported, bounded loosely here and pinned tightly by
    test_bf16_backward_matches_rounding_model."""
    spec = hm.MLPSpec(prec=prec, **VARIANTS[variant]['spec'])
    ws, bs = make_params(spec, 5)
    n = 777
    pts, dirs = points(n, 6)
    g = torch.Generator().manual_seed(7)
    gs = torch.randn(n, generator=g)
    gc = torch.randn(n, 3, generator=g)
    # reference gradients in float64
    pd = {k: v.double().requires_grad_(True) for k, v in as_dict(ws, bs).items()}
    ref = on.mlp_forward(pd, pts.double(), dirs.double(), **VARIANTS[variant]['oracle'])
    loss = (ref['sigma'][:, 0] * gs.double()).sum() + (ref['rgb'] * gc.double()).sum()
    names = list(pd)
    rg = torch.autograd.grad(loss, [pd[k] for k in names])
    wg = [None if w is None else w.to(cuda).requires_grad_(True) for w in ws]
    bg = [None if b is None else b.to(cuda).requires_grad_(True) for b in bs]
    sig, rgb = hm.mlp_forward(spec, wg, bg, pts.to(cuda), dirs.to(cuda))
    ((sig * gs.to(cuda)).sum() + (rgb * gc.to(cuda)).sum()).backward()
    got = {}
    for i, name in enumerate(hm.PARAM_NAMES):
        if wg[i] is not None:
            got[name + '.weight'] = wg[i].grad
            got[name + '.bias'] = bg[i].grad
    errs = {k: ((got[k].cpu().double() - r).norm() / r.norm().clamp_min(1e-12)).item() for k, r in zip(names, rg)}
    print(prec, variant, ' '.join(f'{k}={v:.2e}' for k, v in errs.items()))
    assert max(errs.values()) < tol, errs


@pytest.mark.parametrize('prec', ['bf16', 'bf16_split'])
def test_inference_matches_training_forward(cuda, prec):
    spec = hm.MLPSpec(prec=prec)
    ws, bs = make_params(spec, 9)
    pts, dirs = points(4096, 10)
    wd = [w.to(cuda) for w in ws]
    bd = [b.to(cuda) for b in bs]
    with torch.no_grad():
        s0, c0 = hm.mlp_forward(spec, wd, bd, pts.to(cuda), dirs.to(cuda))
    wg = [w.requires_grad_(True) for w in wd]
    s1, c1 = hm.mlp_forward(spec, wg, bd, pts.to(cuda), dirs.to(cuda))
    assert torch.equal(s0, s1.detach()) and torch.equal(c0, c1.detach())


def test_fp16_training_is_reported_unsupported(cuda):
    """FP16 operands cover the inference forward only; a training forward says so."""
    spec = hm.MLPSpec(prec='fp16')
    ws, bs = make_params(spec, 1)
    pts, dirs = points(64, 2)
    wg = [None if w is None else w.to(cuda).requires_grad_(True) for w in ws]
    with pytest.raises(hm._lib.SrfError, match='status 3'):
        hm.mlp_forward(spec, wg, [None if b is None else b.to(cuda) for b in bs], pts.to(cuda), dirs.to(cuda))


@pytest.mark.parametrize('variant', ['main', 'points_aug', 'views_aug'])
def test_fp16_inference_close_to_fp64(cuda, variant):
    """FP16 operands (10-bit mantissa, 3 more than BF16) with FP32 accumulation: the rgb error
    against FP64 is below the BF16 path's and inside the 2e-3 bar."""
    errs = {}
    for prec in ('fp16', 'bf16'):
        spec = hm.MLPSpec(prec=prec, **VARIANTS[variant]['spec'])
        ws, bs = make_params(spec, 3)
        pts, dirs = points(8192, 4)
        with torch.no_grad():
            sig, rgb = hm.mlp_forward(spec, [None if w is None else w.to(cuda) for w in ws],
                                      [None if b is None else b.to(cuda) for b in bs], pts.to(cuda), dirs.to(cuda))
        ref = run_oracle(variant, as_dict(ws, bs), pts, dirs)
        errs[prec] = (rgb.cpu().double() - ref['rgb']).abs().max().item()
        es = (sig.cpu().double() - ref['sigma'][:, 0]).abs().max().item()
        assert es < 3e-2 * max(1.0, ref['sigma'].abs().max().item()), (prec, es)
    print(variant, 'rgb max abs err vs fp64:', errs)
    assert errs['fp16'] < 2e-3 and errs['fp16'] <= errs['bf16']


def _bf(x):
    return x.to(torch.bfloat16).to(x.dtype)


def bf16_model(variant, ws, bs, pts, dirs, gs, gc):
    """The kernel's BF16 arithmetic restated in float64: operands (weights, PE features, stored
    activations, pre-activation gradients) are rounded to BF16 exactly where the kernel rounds
    them; products accumulate in high precision.  Returns forward outputs and weight/bias grads."""
    cfg = VARIANTS[variant]['oracle']
    D = 21 if variant == 'points_aug' else 63
    vd = variant != 'views_aug'
    W = [None if w is None else _bf(w.double()) for w in ws]
    B = [None if b is None else b.double() for b in bs]
    pe = _bf(on.positional_encoding(pts.double(), 10))
    x0 = pe[:, :D]
    acts, masks = [], []
    h = x0
    for i in range(8):
        inp = torch.cat([x0, h], -1) if i == 5 else h
        acts.append(inp)
        z = inp @ W[i].T + B[i]
        masks.append(z > 0)
        h = _bf(torch.relu(z))
    acts.append(h)                                      # input of the heads
    if not vd:
        o = h @ W[9].T + B[9]
        sig_raw, rgb_raw = o[:, 0], o[:, 1:4]
    else:
        sig_raw = (h @ W[9].T + B[9])[:, 0]
        feat = _bf(h @ W[8].T + B[8])
        extra = _bf(torch.cat([on.positional_encoding(pts.double(), 10)[:, D:],
                               on.positional_encoding(dirs.double(), 4)], -1))
        vin = torch.cat([feat, extra], -1)
        zv = vin @ W[10].T + B[10]
        mv = zv > 0
        v = _bf(torch.relu(zv))
        rgb_raw = v @ W[11].T + B[11]
    sigma = torch.relu(sig_raw)
    rgb = torch.sigmoid(rgb_raw)
    # backward
    dsr = gs.double() * (sigma > 0)
    drr = gc.double() * rgb * (1 - rgb)
    gW = [None] * 12
    gB = [None] * 12
    if not vd:
        dz8 = _bf(torch.cat([dsr[:, None], drr], -1))
        gW[9], gB[9] = dz8.T @ h, dz8.sum(0)
        dh = dz8 @ W[9]
    else:
        dz10 = _bf(drr)
        gW[11], gB[11] = dz10.T @ v, dz10.sum(0)
        dz9 = _bf((dz10 @ W[11]) * mv)
        gW[10], gB[10] = dz9.T @ vin, dz9.sum(0)
        dF = _bf(dz9 @ W[10][:, :256])
        gW[8], gB[8] = dF.T @ h, dF.sum(0)
        ds8 = _bf(dsr)
        gW[9], gB[9] = ds8[:, None].T @ h, ds8.sum()[None]
        dh = dF @ W[8] + ds8[:, None] @ W[9]
    for i in range(7, -1, -1):
        dz = _bf(dh * masks[i])
        gW[i], gB[i] = dz.T @ acts[i], dz.sum(0)
        if i > 0:
            dh = dz @ (W[i][:, D:] if i == 5 else W[i])
    return sigma, rgb, gW, gB


@pytest.mark.parametrize('variant', list(VARIANTS))
def test_bf16_backward_matches_rounding_model(cuda, variant):
    spec = hm.MLPSpec(prec='bf16', **VARIANTS[variant]['spec'])
    ws, bs = make_params(spec, 5)
    n = 777
    pts, dirs = points(n, 6)
    g = torch.Generator().manual_seed(7)
    gs = torch.randn(n, generator=g)
    gc = torch.randn(n, 3, generator=g)
    sig_m, rgb_m, gW, gB = bf16_model(variant, ws, bs, pts, dirs, gs, gc)
    wg = [None if w is None else w.to(cuda).requires_grad_(True) for w in ws]
    bg = [None if b is None else b.to(cuda).requires_grad_(True) for b in bs]
    sig, rgb = hm.mlp_forward(spec, wg, bg, pts.to(cuda), dirs.to(cuda))
    assert (sig.detach().cpu().double() - sig_m).abs().max() < 2e-3 * max(1.0, sig_m.abs().max().item())
    assert (rgb.detach().cpu().double() - rgb_m).abs().max() < 2e-3
    ((sig * gs.to(cuda)).sum() + (rgb * gc.to(cuda)).sum()).backward()
    errs = {}
    for i, name in enumerate(hm.PARAM_NAMES):
        if wg[i] is None:
            continue
        for kind, got, ref in (('weight', wg[i].grad, gW[i]), ('bias', bg[i].grad, gB[i])):
            errs[f'{name}.{kind}'] = ((got.cpu().double() - ref.reshape(got.shape)).norm()
                                      / ref.norm().clamp_min(1e-12)).item()
    print('bf16-model', variant, ' '.join(f'{k}={v:.1e}' for k, v in errs.items()))
    assert max(errs.values()) < 2e-2, errs


@pytest.mark.parametrize('variant', list(VARIANTS))
@pytest.mark.parametrize('prec', ['fp32', 'bf16'])
def test_ray_form_matches_materialised_points(cuda, variant, prec):
    """srf_mlp_forward_rays generates rays_o + rays_d * z (the reference's product and sum, each
    rounded) and expands the per-ray view directions in the kernel: sigma, rgb and the weight
    gradients equal those of srf_mlp_forward on the materialised points, bit for bit; also with a
    ray count that leaves the last 256-point block partly empty, and with sigma noise."""
    spec = hm.MLPSpec(prec=prec, **VARIANTS[variant]['spec'])
    ws, bs = make_params(spec, 11)
    n_rays, n_s = 333, 64
    g = torch.Generator().manual_seed(12)
    o = (torch.rand(n_rays, 3, generator=g) * 2 - 1).to(cuda)
    d = torch.randn(n_rays, 3, generator=g).to(cuda)
    z = torch.sort(torch.rand(n_rays, n_s, generator=g), dim=-1).values.to(cuda)
    vd = torch.nn.functional.normalize(d, dim=-1)
    noise = torch.randn(n_rays * n_s, generator=g).to(cuda)
    pts = (o[..., None, :] + d[..., None, :] * z[..., :, None]).reshape(-1, 3)
    dirs = vd[:, None].expand(n_rays, n_s, 3).reshape(-1, 3)

    def run(ray_form):
        w = [None if t is None else t.to(cuda).requires_grad_() for t in ws]
        b = [None if t is None else t.to(cuda).requires_grad_() for t in bs]
        if ray_form:
            s, c = hm.mlp_forward_rays(spec, w, b, o, d, z, vd, noise)
        else:
            s, c = hm.mlp_forward(spec, w, b, pts, dirs, noise)
        (s.square().sum() + (c * torch.linspace(-1, 1, 3, device=cuda)).sum()).backward()
        return s.detach(), c.detach(), [t.grad for t in w + b if t is not None]

    s0, c0, g0 = run(False)
    s1, c1, g1 = run(True)
    assert torch.equal(s0, s1) and torch.equal(c0, c1)
    for a, b in zip(g0, g1):
        # the same products, summed by FP32 atomics whose order may differ between the launches
        assert (a - b).abs().max() <= 1e-5 * a.abs().max() + 1e-12

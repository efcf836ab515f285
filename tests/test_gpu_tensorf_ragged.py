"""Ragged (occupancy-pruned) TensoRF batches on the HIP path: LowRankTensor.forward_ragged against
the oracle rendering each ray on its own in FP64 (outputs and factor / network gradients), for
the VM and CP decompositions, on 64 rays with log-normal sample counts; and the ragged train step
replayed from a captured HIP graph against eager steps."""
import copy

import pytest
import torch
import torch.nn.functional as F

import oracle.tensorf as ot
from models.ModelFactory02 import get_model
from srf_hip import configs as cfgs, losses

pytestmark = pytest.mark.gpu

THR = 1e-4
CP = dict(decomposition_type='CandecompParafac', num_components_density=(16,), num_components_color=(48,))


def ragged_batch(n_rays, median, max_count, bbox, seed, step=0.02):
    """Rays from points inside the box in random directions; per-ray counts ~ round(median *
    exp(0.5 N(0,1))) truncated to [1, max_count]; samples t0 + step (k + u) along each ray."""
    g = torch.Generator().manual_seed(seed)
    lo, hi = bbox[0], bbox[1]
    o = lo + (hi - lo) * (0.2 + 0.6 * torch.rand(n_rays, 3, generator=g))
    d = F.normalize(torch.randn(n_rays, 3, generator=g), dim=-1) * (0.5 + torch.rand(n_rays, 1, generator=g))
    counts = (median * torch.exp(0.5 * torch.randn(n_rays, generator=g))).round().clamp(1, max_count).long()
    off = torch.zeros(n_rays + 1, dtype=torch.int64)
    off[1:] = torch.cumsum(counts, 0)
    ray = torch.repeat_interleave(torch.arange(n_rays), counts)
    k = torch.arange(int(off[-1])) - off[:-1][ray]
    u = torch.rand(n_rays, generator=g)
    z = 0.05 + step * (k.float() + u[ray])
    pts = o[ray] + d[ray] * z[:, None]
    return dict(pts=pts, z_vals=z, ray_offsets=off.int(), rays_o=o, rays_d=d, view_dirs=F.normalize(d, dim=-1),
                max_samples=int(counts.max()))


def tensor_model(cuda, cp):
    torch.manual_seed(0)
    coarse = dict(num_voxels_initial=32 ** 3)
    if cp:
        coarse.update(CP)
    c = cfgs.simple_tensorf_configs(augmentations=False, ndc=False, white_bkgd=True, perturb=False, coarse=coarse)
    model = get_model(c, model_configs=cfgs.model_configs()).to(cuda).eval()
    tm = model.coarse_model
    with torch.no_grad():   # densities that put weight on a good share of the samples
        for p in list(tm.vectors_density) + list(getattr(tm, 'matrices_density', [])):
            p.mul_(6.0 if cp else 3.0)
        if cp:
            for p in tm.vectors_color:
                p.mul_(6.0)
    return tm


def oracle_params(tm, cp):
    f64 = lambda t: t.detach().cpu().double().contiguous().requires_grad_(True)
    prm = {'bbox': tm.bounding_box.detach().cpu().double(),
           'lines_density': [f64(p) for p in tm.vectors_density], 'lines_color': [f64(p) for p in tm.vectors_color],
           'basis': f64(tm.basis_matrix_color.weight),
           'color_mlp': {k: f64(v) for k, v in tm.color_predictor.state_dict().items()}}
    if cp:
        prm['decomposition'] = 'CandecompParafac'
    else:
        prm['planes_density'] = [f64(p) for p in tm.matrices_density]
        prm['planes_color'] = [f64(p) for p in tm.matrices_color]
    return prm


@pytest.mark.parametrize('cp', [False, True])
def test_ragged_render_and_grads_match_oracle(cuda, cp):
    tm = tensor_model(cuda, cp)
    b = ragged_batch(64, 24, 96, tm.bounding_box.cpu(), seed=1)
    inp = {k: (v.to(cuda) if torch.is_tensor(v) else v) for k, v in b.items()}
    out = tm.forward_ragged(inp, retraw=True)
    g = torch.Generator().manual_seed(2)
    g_rgb, g_d = torch.randn(64, 3, generator=g), torch.randn(64, generator=g)
    ((out['rgb'] * g_rgb.to(cuda)).sum() + (out['acc'] * g_d.to(cuda)).sum()).backward()
    prm = oracle_params(tm, cp)
    d64 = {k: (v.double() if torch.is_tensor(v) and v.is_floating_point() else v) for k, v in b.items()}
    ref = ot.tensorf_render_ragged(prm, d64['pts'], d64['z_vals'], b['ray_offsets'], d64['rays_o'], d64['rays_d'],
                                   d64['view_dirs'], distance_scale=25.0, weight_threshold=THR, white_bkgd=True)
    w = ref['weights'].detach()
    assert (w > THR).sum() > 300                     # the colour path is exercised
    assert not ((w - THR).abs() < 1e-9).any()
    for k in ('rgb', 'acc', 'weights'):
        e = (out[k].detach().cpu().double() - ref[k].detach()).abs().max().item()
        assert e < 1e-5, (k, e)
    # metric depth sum(w z)/(acc + 1e-6): compare the numerator
    a = out['depth'].detach().cpu().double() * (out['acc'].detach().cpu().double() + 1e-6)
    r = (ref['depth'] * (ref['acc'] + 1e-6)).detach()
    assert (a - r).abs().max().item() < 1e-5 * max(1.0, r.abs().max().item())
    ((ref['rgb'] * g_rgb.double()).sum() + (ref['acc'] * g_d.double()).sum()).backward()
    pairs = [(tm.vectors_density, prm['lines_density']), (tm.vectors_color, prm['lines_color'])]
    if not cp:
        pairs += [(tm.matrices_density, prm['planes_density']), (tm.matrices_color, prm['planes_color'])]
    for ours, theirs in pairs:
        for x, y in zip(ours, theirs):
            err = ((x.grad.cpu().double() - y.grad).abs().max() / y.grad.abs().max().clamp_min(1e-30)).item()
            assert err < 1e-3, err
    nets = [(tm.basis_matrix_color.weight, prm['basis'])] + \
        [(dict(tm.color_predictor.named_parameters())[k], v) for k, v in prm['color_mlp'].items()]
    for x, y in nets:
        err = ((x.grad.cpu().double() - y.grad).abs().max() / y.grad.abs().max().clamp_min(1e-30)).item()
        assert err < 1e-3, err


@pytest.mark.parametrize('color_precision', ['fp32', 'bf16'])
def test_ragged_graph_replays_match_eager_steps(cuda, color_precision):
    """The ragged train step (forward_ragged, the MSE kernel, backward, Adam) captured once in a HIP
    graph: each replay matches an eager step of a copy of the model from the same state (parameters
    and optimiser state copied in before every replay), with eager work and host synchronisation
    between replays."""
    torch.manual_seed(0)
    c = cfgs.simple_tensorf_configs(augmentations=False, ndc=False, white_bkgd=True, perturb=False,
                                    coarse=dict(num_voxels_initial=32 ** 3, features_positional_encoding_degree=2))
    c['model']['color_precision'] = color_precision
    a = get_model(c, model_configs=cfgs.model_configs()).to(cuda).train().coarse_model
    with torch.no_grad():
        for p in list(a.vectors_density) + list(a.matrices_density):
            p.mul_(3.0)
    b = copy.deepcopy(a)
    batch = ragged_batch(256, 24, 96, a.bounding_box.cpu(), seed=5)
    inp = {k: (v.to(cuda) if torch.is_tensor(v) else v) for k, v in batch.items()}
    target = torch.rand(256, 3, generator=torch.Generator().manual_seed(6)).to(cuda)
    pa = [p for p in a.parameters() if p.requires_grad]
    pb = [p for p in b.parameters() if p.requires_grad]
    opt_a = torch.optim.Adam(pa, lr=1e-3, capturable=True)
    opt_b = torch.optim.Adam(pb, lr=1e-3, capturable=True)

    def step(model, opt):
        out = model.forward_ragged(inp)
        loss = losses.mse_terms([out['rgb']], target)
        loss.backward()
        opt.step()
        return loss

    for _ in range(2):
        opt_a.zero_grad(set_to_none=True)
        step(a, opt_a)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            opt_b.zero_grad(set_to_none=True)
            step(b, opt_b)
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    opt_b.zero_grad(set_to_none=True)
    with torch.cuda.graph(graph):
        loss_b = step(b, opt_b)
    for k in range(3):
        with torch.no_grad():
            for x, y in zip(pb, pa):
                x.copy_(y)
            for x, y in zip(pb, pa):
                for key, v in opt_a.state.get(y, {}).items():
                    if isinstance(v, torch.Tensor):
                        opt_b.state[x][key].copy_(v)
        opt_a.zero_grad(set_to_none=True)
        la = step(a, opt_a).item()
        graph.replay()
        torch.cuda.synchronize()
        lb = loss_b.item()
        _ = sum(float(p.detach().abs().sum()) for p in pb)   # eager work between replays
        assert la > 0 and abs(la - lb) <= 1e-5 * max(1.0, abs(la)), (k, la, lb)
        for (na, xa), xb in zip(a.named_parameters(), b.parameters()):
            err = (xa.detach() - xb.detach()).abs().max().item()
            assert err <= 1e-3, (k, na, err)

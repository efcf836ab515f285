"""The Simple-TensoRF train step is device-resident and replays from a captured HIP graph: the
surface count of the ray-marching threshold stays on the GPU (the appearance gather and the colour
network stop at it), so the step has no host synchronisation, and every replay of a captured step
matches an eager run of the same model from the same state, with eager work and host
synchronisation between replays."""
import copy

import pytest
import torch

from models.ModelFactory02 import get_model
from srf_hip import configs as cfgs, synthetic

pytestmark = pytest.mark.gpu

CP = dict(decomposition_type='CandecompParafac', num_components_density=(16,), num_components_color=(48,))


def _model(cuda, cp=False, features_pe=0, color_precision='fp32'):
    torch.manual_seed(0)
    coarse = dict(num_voxels_initial=32 ** 3, features_positional_encoding_degree=features_pe)
    if cp:
        coarse.update(CP)
    # white background and deterministic sampling: both runs see the same inputs at every step
    c = cfgs.simple_tensorf_configs(white_bkgd=True, perturb=False, coarse=coarse,
                                    augmented=dict(num_voxels_initial=16 ** 3))
    c['model']['color_precision'] = color_precision
    m = get_model(c, model_configs=cfgs.model_configs()).to(cuda).train()
    if cp:
        with torch.no_grad():
            for p in list(m.coarse_model.vectors_density) + list(m.coarse_model.vectors_color):
                p.mul_(6.0)
    return m


def _opt(kind, model):
    params = [p for p in model.parameters() if p.requires_grad]
    if kind == 'sgd':
        return torch.optim.SGD(params, lr=1e-2, momentum=0.9)
    return torch.optim.Adam(params, lr=1e-3, capturable=True)


def _inputs(cuda):
    batch = synthetic.train_batch(1024, seed=2, device=cuda)
    return batch, {'pixel_id': batch['pixel_id'], 'num_frames': 3, 'iter_num': 0, 'sub_batch_index': 0}


def test_train_step_has_no_host_synchronisation(cuda):
    """After the first call (which reads the grid geometry to the host once, as after every model
    modification), a training forward and backward make no synchronising call."""
    model = _model(cuda)
    batch, inp = _inputs(cuda)
    synthetic.train_loss(model(inp), batch).backward()
    model.zero_grad(set_to_none=True)
    torch.cuda.synchronize()
    torch.cuda.set_sync_debug_mode('error')
    try:
        out = model(inp)
        loss = synthetic.train_loss(out, batch)
        loss.backward()
    finally:
        torch.cuda.set_sync_debug_mode('default')
    torch.cuda.synchronize()
    assert torch.isfinite(loss).item()
    assert model.coarse_model.matrices_color[0].grad.abs().sum().item() > 0


# Each replay is compared with an eager step from the same state (the captured model's parameters
# and optimiser state are set to the eager model's before every replay), so that the comparison
# sees one step's summation-order differences only: over several steps those would grow wherever a
# density ReLU or the surface threshold flips, and the trajectories would part for that reason alone.
@pytest.mark.parametrize('kind,cp,features_pe,param_tol,color_precision', [
    ('sgd', False, 0, None, 'fp32'), ('adam', False, 0, 1e-3, 'fp32'), ('sgd', True, 0, None, 'fp32'),
    ('sgd', False, 2, None, 'fp32'), ('sgd', False, 2, None, 'bf16'), ('adam', False, 2, 1e-3, 'bf16_split')])
def test_graph_replays_match_eager_steps(cuda, kind, cp, features_pe, param_tol, color_precision):
    a = _model(cuda, cp, features_pe, color_precision)
    b = copy.deepcopy(a)
    batch, inp = _inputs(cuda)
    opt_a, opt_b = _opt(kind, a), _opt(kind, b)
    pa = [p for p in a.parameters() if p.requires_grad]
    pb = [p for p in b.parameters() if p.requires_grad]

    def step(model, opt):
        out = model(inp)
        loss = synthetic.train_loss(out, batch)
        loss.backward()
        opt.step()
        return loss

    for _ in range(2):   # the eager model's optimiser state exists from here on, as the captured one's
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
        assert abs(la - lb) <= 1e-5 * max(1.0, abs(la)), (k, la, lb)
        for (na, xa), xb in zip(a.named_parameters(), b.parameters()):
            err = (xa.detach() - xb.detach()).abs().max().item()
            tol = param_tol if param_tol is not None else 1e-5 * max(1.0, xa.detach().abs().max().item())
            assert err <= tol, (k, na, err)

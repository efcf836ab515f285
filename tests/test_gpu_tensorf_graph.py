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


def _model(cuda, cp=False, features_pe=0):
    torch.manual_seed(0)
    coarse = dict(num_voxels_initial=32 ** 3, features_positional_encoding_degree=features_pe)
    if cp:
        coarse.update(CP)
    # white background and deterministic sampling: both runs see the same inputs at every step
    c = cfgs.simple_tensorf_configs(white_bkgd=True, perturb=False, coarse=coarse,
                                    augmented=dict(num_voxels_initial=16 ** 3))
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


# The FP32 atomics of the factor gathers and of the weight-gradient GEMM make two runs differ at
# rounding level: SGD passes that through; Adam's m / sqrt(v) can turn a rounding-level gradient of
# either sign into a full +-lr step per element and step.
# features PE 2: the colour backward's LDS tile then exceeds the 64 KB default of dynamic LDS
@pytest.mark.parametrize('kind,cp,features_pe,param_tol', [('sgd', False, 0, None), ('adam', False, 0, 1e-3 * 5),
                                                          ('sgd', True, 0, None), ('sgd', False, 2, None)])
def test_graph_replays_match_eager_steps(cuda, kind, cp, features_pe, param_tol):
    a = _model(cuda, cp, features_pe)
    b = copy.deepcopy(a)
    batch, inp = _inputs(cuda)
    opt_a, opt_b = _opt(kind, a), _opt(kind, b)

    def step(model, opt):
        out = model(inp)
        loss = synthetic.train_loss(out, batch)
        loss.backward()
        opt.step()
        return loss

    ref = []
    for _ in range(5):
        opt_a.zero_grad(set_to_none=True)
        ref.append(step(a, opt_a).item())
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        got = []
        for _ in range(2):
            opt_b.zero_grad(set_to_none=True)
            got.append(step(b, opt_b).item())
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    opt_b.zero_grad(set_to_none=True)
    with torch.cuda.graph(graph):
        loss = step(b, opt_b)
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        got.append(loss.item())
        _ = sum(float(p.detach().abs().sum()) for p in b.parameters())   # eager work between replays
    for i, (x, y) in enumerate(zip(ref, got)):
        assert abs(x - y) <= 1e-4 * max(1.0, abs(x)), (i, x, y)
    for (na, pa), (nb, pb) in zip(a.named_parameters(), b.named_parameters()):
        err = (pa.detach() - pb.detach()).abs().max().item()
        tol = param_tol if param_tol is not None else 1e-5 * max(1.0, pa.detach().abs().max().item())
        assert err <= tol, (na, err)

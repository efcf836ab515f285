"""The Simple-NeRF train step captured in a HIP graph replays the eager step: every replay is
checked against an eager run of the same model from the same state, with eager work and host
synchronisation between replays."""
import copy

import pytest
import torch

from models.ModelFactory02 import get_model
from srf_hip import configs as cfgs, synthetic

pytestmark = pytest.mark.gpu


def _model(cuda):
    torch.manual_seed(0)
    # deterministic sampling and no sigma noise: both runs see the same inputs at every step
    c = cfgs.simple_nerf_configs(precision='fp32', perturb=False, raw_noise_std=0.0)
    return get_model(c, model_configs=cfgs.model_configs()).to(cuda).train()


def _opt(kind, model):
    params = [p for p in model.parameters() if p.requires_grad]
    if kind == 'sgd':
        return torch.optim.SGD(params, lr=1e-2, momentum=0.9)
    if kind == 'adam_hip':
        from srf_hip.adam import AdamHip
        return AdamHip(params, lr=1e-3)
    return torch.optim.Adam(params, lr=1e-3, capturable=True, fused=kind == 'adam_fused')


# Parameter bounds after 5 steps.  The FP32 atomics of the weight-gradient GEMM make any two runs
# differ at rounding level; SGD passes that through (1e-5 relative).  Adam's m / sqrt(v) turns a
# rounding-level gradient of either sign into a full +-lr step, so an element may drift by up to
# lr per step (the loss trajectory above is still held to 1e-4).
@pytest.mark.parametrize('kind,param_tol', [('sgd', None), ('adam', 1e-3 * 5), ('adam_fused', 1e-3 * 5),
                                            ('adam_hip', 1e-3 * 5)])
def test_graph_replays_match_eager_steps(cuda, kind, param_tol):
    a = _model(cuda)
    b = copy.deepcopy(a)
    batch = synthetic.train_batch(512, seed=1, device=cuda)
    inp = {'pixel_id': batch['pixel_id'], 'num_frames': 3, 'iter_num': 0, 'sub_batch_index': 0}
    opt_a = _opt(kind, a)
    opt_b = _opt(kind, b)

    def step(model, opt):
        out = model(inp)
        loss = synthetic.train_loss(out, batch)
        loss.backward()
        opt.step()
        return loss

    # eager reference: 5 steps
    ref = []
    for _ in range(5):
        opt_a.zero_grad(set_to_none=True)
        ref.append(step(a, opt_a).item())
    # graph: 2 eager warm-up steps on a side stream, capture, then 3 replays
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

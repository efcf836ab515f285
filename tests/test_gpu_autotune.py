"""The launch-schedule knobs change no result beyond the order of FP32 atomic additions: the gather
backwards under other run lengths and grid caps against FP64 autograd, the weight gradients of the
NeRF MLP and of the colour network under other point chunks against the built-in schedule, and the
autotuner picks one of its candidates and leaves it set."""
import pytest
import torch
import torch.nn.functional as F

import oracle.tensorf as ot
from srf_hip import autotune, color, mlp as hm, tensorf

pytestmark = pytest.mark.gpu

RES = [40, 36, 30]
BBOX = torch.tensor([[-1.5, -1.0, -0.5], [1.5, 1.2, 1.0]])


@pytest.fixture(autouse=True)
def _reset_knobs():
    autotune.reset()
    yield
    autotune.reset()


def _vm_factors(comps, seed):
    g = torch.Generator().manual_seed(seed)
    planes, lines = [], []
    for p, c in enumerate(comps):
        a0, a1 = ot.MATRIX_AXES[p]
        planes.append(0.1 * torch.randn(1, c, RES[a1], RES[a0], generator=g))
        lines.append(0.1 * torch.randn(1, c, RES[ot.VECTOR_AXES[p]], 1, generator=g))
    return planes, lines


def _points(n, seed):
    """Points along 64 rays (consecutive samples share texel corners, as in a render)."""
    g = torch.Generator().manual_seed(seed)
    o = BBOX[0] + (BBOX[1] - BBOX[0]) * torch.rand(64, 3, generator=g)
    d = F.normalize(torch.randn(64, 3, generator=g), dim=-1) * 0.02
    t = torch.arange(n // 64, dtype=torch.float32)
    return (o[:, None] + d[:, None] * t[None, :, None]).reshape(-1, 3)


@pytest.mark.parametrize('knob,value', [('gather_run', 1), ('gather_run', 7), ('gather_run', 256),
                                        ('gather_blocks', 256), ('gather_blocks', 65536)])
def test_gather_backwards_under_schedule_knobs(cuda, knob, value):
    planes, lines = _vm_factors([16, 8, 8], 1)
    pts = _points(64 * 120, 2)
    g = torch.Generator().manual_seed(3)
    gs = torch.randn(pts.shape[0], generator=g)
    P64 = [p.double().requires_grad_(True) for p in planes]
    L64 = [l.double().requires_grad_(True) for l in lines]
    pn = ot.normalize_points(pts.double(), BBOX.double())
    valid = ot.bbox_validity(pts.double(), BBOX.double())
    feat = ot.vm_density_features(P64, L64, pn[valid])
    app = ot.vm_appearance_features(P64, L64, pn)
    ga = torch.randn(app.shape, generator=g)
    (F.relu(feat) * gs[valid].double()).sum().backward()
    (app * ga.double()).sum().backward()
    autotune.set_knob(knob, value)
    P = [tensorf.texel_major_param(t.shape[1], t.shape[2], t.shape[3], t, device=cuda) for t in planes]
    L = [tensorf.texel_major_param(t.shape[1], t.shape[2], t.shape[3], t, device=cuda) for t in lines]
    sig, _ = tensorf.vm_density(pts.to(cuda), P, L, BBOX.flatten().tolist())
    (sig * gs.to(cuda)).sum().backward()
    got = tensorf.vm_appearance(pts.to(cuda), P, L, BBOX.flatten().tolist())
    (got * ga.to(cuda)).sum().backward()
    torch.cuda.synchronize()
    for a, b in zip(P + L, P64 + L64):
        err = (a.grad.cpu().double() - b.grad).abs().max() / b.grad.abs().max()
        assert err < 1e-4, (knob, value, err.item())
    # CP under the same knob
    g2 = torch.Generator().manual_seed(4)
    cl = [0.3 * torch.randn(1, 16, RES[ot.VECTOR_AXES[i]], 1, generator=g2) for i in range(3)]
    C64 = [l.double().requires_grad_(True) for l in cl]
    (ot.cp_appearance_features(C64, pn) * ga[:, :16].double()).sum().backward()
    CL = [tensorf.texel_major_param(t.shape[1], t.shape[2], t.shape[3], t, device=cuda) for t in cl]
    (tensorf.cp_appearance(pts.to(cuda), CL, BBOX.flatten().tolist()) * ga[:, :16].to(cuda)).sum().backward()
    torch.cuda.synchronize()
    for a, b in zip(CL, C64):
        err = (a.grad.cpu().double() - b.grad).abs().max() / b.grad.abs().max()
        assert err < 1e-4, ('cp', knob, value, err.item())


def _mlp_grads(cuda, spec, ws, bs, pts, dirs, gs, gc):
    wd = [None if w is None else w.to(cuda).requires_grad_(True) for w in ws]
    bd = [None if b is None else b.to(cuda).requires_grad_(True) for b in bs]
    sig, rgb = hm.mlp_forward(spec, wd, bd, pts.to(cuda), dirs.to(cuda))
    ((sig * gs.to(cuda)).sum() + (rgb * gc.to(cuda)).sum()).backward()
    torch.cuda.synchronize()
    return [t.grad.cpu() for t in wd + bd if t is not None]


@pytest.mark.parametrize('value', [2048, 32768])
def test_weight_grads_under_dw_chunk(cuda, value):
    spec = hm.MLPSpec(prec='fp32')
    g = torch.Generator().manual_seed(5)
    ws, bs = [], []
    for s in spec.param_shapes():
        if s is None:
            ws.append(None)
            bs.append(None)
            continue
        bound = (6.0 / (s[0][1] + s[0][0])) ** 0.5
        ws.append((torch.rand(s[0], generator=g) * 2 - 1) * bound)
        bs.append((torch.rand(s[1], generator=g) * 2 - 1) * 0.05)
    n = 20000
    pts = torch.rand(n, 3, generator=g) * 2 - 1
    dirs = F.normalize(torch.randn(n, 3, generator=g), dim=-1)
    gs, gc = torch.randn(n, generator=g), torch.randn(n, 3, generator=g)
    ref = _mlp_grads(cuda, spec, ws, bs, pts, dirs, gs, gc)
    autotune.set_knob('dw_chunk', value)
    got = _mlp_grads(cuda, spec, ws, bs, pts, dirs, gs, gc)
    for a, b in zip(got, ref):
        assert ((a - b).abs().max() / b.abs().max().clamp_min(1e-30)).item() < 1e-5


@pytest.mark.parametrize('prec,value', [('fp32', 2048), ('fp32', 65536), ('bf16', 2304), ('bf16', 65536)])
def test_color_weight_grads_under_chunk(cuda, prec, value):
    g = torch.Generator().manual_seed(6)
    m, C = 30000, 72
    app = torch.randn(m, C, generator=g) * 0.3
    vd = F.normalize(torch.randn(m, 3, generator=g), dim=-1)
    lin = lambda o, i: ((torch.rand(o, i, generator=g) * 2 - 1) / i ** 0.5, (torch.rand(o, generator=g) * 2 - 1) * 0.1)
    basis = (torch.rand(27, C, generator=g) * 2 - 1) / C ** 0.5
    (w1, b1), (w2, b2), (w3, b3) = lin(128, 30), lin(128, 128), lin(3, 128)
    gout = torch.randn(m, 3, generator=g)
    spec = color.ColorSpec(n_features=C, prec=prec)

    def grads():
        pd = [t.to(cuda).requires_grad_(True) for t in (basis, w1, b1, w2, b2, w3, b3)]
        rgb = color.color_forward(spec, app.to(cuda), pd[0], pd[1:], view_dirs=vd.to(cuda))
        (rgb * gout.to(cuda)).sum().backward()
        torch.cuda.synchronize()
        return [p.grad.cpu() for p in pd]

    ref = grads()
    autotune.set_knob('color_dw_chunk', value)
    for a, b in zip(grads(), ref):
        assert ((a - b).abs().max() / b.abs().max().clamp_min(1e-30)).item() < 1e-5


def test_autotuner_picks_a_candidate(cuda, tmp_path):
    planes, lines = _vm_factors([16, 8, 8], 7)
    P = [tensorf.texel_major_param(t.shape[1], t.shape[2], t.shape[3], t, device=cuda) for t in planes]
    L = [tensorf.texel_major_param(t.shape[1], t.shape[2], t.shape[3], t, device=cuda) for t in lines]
    pts = _points(64 * 400, 8).to(cuda)

    def step():
        sig, _ = tensorf.vm_density(pts, P, L, BBOX.flatten().tolist())
        sig.sum().backward()

    tuner = autotune.Autotuner(tmp_path / 'tune.json')
    rec = tuner.tune('vm_density_25600', 'gather_run', step, candidates=[0, 4, 64], reps=2)
    assert rec['chosen'] in (0, 4, 64) and set(rec['ms']) == {'0', '4', '64'}
    assert autotune.get_knob('gather_run') == rec['chosen']
    autotune.reset()
    assert autotune.Autotuner(tmp_path / 'tune.json').apply('vm_density_25600')
    assert autotune.get_knob('gather_run') == rec['chosen']

"""GPU parity of the TensoRF kernels (VM gather fwd/bwd, alpha mask, compaction, volume render,
AABB sampler) against the oracle (torch grid_sample on the CPU in FP32 and FP64)."""
import pytest
import torch

import oracle.tensorf as ot
from srf_hip import compositing, tensorf

pytestmark = pytest.mark.gpu

RES = [40, 36, 30]


def factors(comps, res, seed):
    g = torch.Generator().manual_seed(seed)
    planes, lines = [], []
    for p, c in enumerate(comps):
        a0, a1 = ot.MATRIX_AXES[p]
        v = ot.VECTOR_AXES[p]
        planes.append(0.1 * torch.randn(1, c, res[a1], res[a0], generator=g))
        lines.append(0.1 * torch.randn(1, c, res[v], 1, generator=g))
    return planes, lines


def points(n, bbox, seed, outside=0.2):
    """Uniform points in a box 20% larger than bbox (so about a third fall outside)."""
    g = torch.Generator().manual_seed(seed)
    lo, hi = bbox[0], bbox[1]
    pad = (hi - lo) * outside
    return lo - pad + (hi - lo + 2 * pad) * torch.rand(n, 3, generator=g)


BBOX = torch.tensor([[-1.5, -1.0, -0.5], [1.5, 1.2, 1.0]])


def to_dev(ts, dev, texel_major):
    out = []
    for t in ts:
        if texel_major:
            out.append(tensorf.texel_major_param(t.shape[1], t.shape[2], t.shape[3], t, device=dev))
        else:
            out.append(torch.nn.Parameter(t.to(dev)))
    return out


def oracle_density(planes, lines, pts, bbox, alpha=None, kind='ReLU', offset=-10.0, dtype=torch.float64):
    P = [p.to(dtype) for p in planes]
    L = [l.to(dtype) for l in lines]
    pts = pts.to(dtype)
    bb = bbox.to(dtype)
    valid = ot.bbox_validity(pts, bb)
    if alpha is not None:
        a = ot.alpha_mask_sample(alpha[0].to(dtype), alpha[1].to(dtype), pts[valid])
        valid[valid.clone()] &= a > 0
    feat = torch.zeros(pts.shape[0], dtype=dtype)
    feat[valid] = ot.vm_density_features(P, L, ot.normalize_points(pts, bb)[valid])
    sigma = torch.zeros_like(feat)
    sigma[valid] = ot.density_predictor(feat[valid], kind, offset)
    return sigma, valid


@pytest.mark.parametrize('comps,texel_major', [([16, 4, 4], True), ([16, 16, 16], False), ([4, 4, 4], True),
                                               ([24, 24, 24], True)])
def test_density_forward(cuda, comps, texel_major):
    planes, lines = factors(comps, RES, 0)
    pts = points(20000, BBOX, 1)
    s64, v64 = oracle_density(planes, lines, pts, BBOX)
    s32, _ = oracle_density(planes, lines, pts, BBOX, dtype=torch.float32)
    P = to_dev(planes, cuda, texel_major)
    L = to_dev(lines, cuda, texel_major)
    sig, valid = tensorf.vm_density(pts.to(cuda), P, L, BBOX.flatten().tolist())
    torch.cuda.synchronize()
    assert torch.equal(valid.cpu().bool(), v64)
    # FP32 parity bar: 1e-5 abs against FP64 and against the FP32 oracle (sums of up to 72
    # products of N(0, 0.1) factors, |feature| ~ 0.1)
    assert (sig.cpu().double() - s64).abs().max() < 1e-5
    assert (sig.cpu() - s32).abs().max() < 1e-5


def test_density_alpha_mask_and_softplus(cuda):
    planes, lines = factors([16, 4, 4], RES, 2)
    g = torch.Generator().manual_seed(3)
    vol = (torch.rand(16, 16, 16, generator=g) < 0.4).float()
    abox = torch.tensor([[-1.4, -0.9, -0.4], [1.4, 1.1, 0.9]])
    pts = points(30000, BBOX, 4)
    s64, v64 = oracle_density(planes, lines, pts, BBOX, alpha=(vol, abox), kind='SoftPlus', offset=-2.0)
    mask = tensorf.AlphaMask(vol.to(cuda), abox.flatten().tolist())
    P = to_dev(planes, cuda, True)
    L = to_dev(lines, cuda, True)
    sig, valid = tensorf.vm_density(pts.to(cuda), P, L, BBOX.flatten().tolist(), alpha=mask,
                                    predictor=tensorf.SOFTPLUS, offset=-2.0)
    torch.cuda.synchronize()
    n_diff = (valid.cpu().bool() != v64).sum().item()
    # occupancy is a sign test on a trilinear sum: a point exactly on a zero/nonzero boundary may
    # round differently; none is expected for random points
    assert n_diff == 0
    assert (sig.cpu().double() - s64).abs().max() < 1e-5
    a = tensorf.alpha_mask_sample(mask, pts.to(cuda))
    ref = ot.alpha_mask_sample(vol.double(), abox.double(), pts.double())
    ref32 = ot.alpha_mask_sample(vol, abox, pts)
    # FP32 coordinates: the normalised position carries ~1e-7 relative rounding, i.e. ~1e-6 of a voxel
    assert (a.cpu().double() - ref).abs().max() < 1e-5
    assert (a.cpu() - ref32).abs().max() < 1e-5


@pytest.fixture(params=[True, False], ids=['cell_sorted', 'sample_order'])
def sorted_bwd(request, monkeypatch):
    """Both factor-backward schedules: samples ordered by plane cell (workspace) and in sample order."""
    monkeypatch.setattr(tensorf, 'SORTED_BACKWARD', request.param)
    return request.param


@pytest.mark.parametrize('comps,kind', [([16, 4, 4], 'ReLU'), ([8, 8, 8], 'SoftPlus')])
def test_density_backward_matches_autograd(cuda, comps, kind, sorted_bwd):
    planes, lines = factors(comps, RES, 5)
    pts = points(8000, BBOX, 6)
    g = torch.Generator().manual_seed(7)
    gs = torch.randn(8000, generator=g)
    # reference gradient: autograd through grid_sample in FP64 on the CPU
    P64 = [p.double().requires_grad_(True) for p in planes]
    L64 = [l.double().requires_grad_(True) for l in lines]
    s64, _ = oracle_density(P64, L64, pts, BBOX, kind=kind, offset=-1.0)
    (s64 * gs.double()).sum().backward()
    P = to_dev(planes, cuda, True)
    L = to_dev(lines, cuda, True)
    sig, _ = tensorf.vm_density(pts.to(cuda), P, L, BBOX.flatten().tolist(), predictor=tensorf.predictor_code(kind),
                                offset=-1.0)
    (sig * gs.to(cuda)).sum().backward()
    torch.cuda.synchronize()
    for got, ref in zip(P + L, P64 + L64):
        assert got.grad.shape == ref.grad.shape
        err = (got.grad.cpu().double() - ref.grad).abs().max() / ref.grad.abs().max()
        assert err < 1e-4, err


@pytest.mark.parametrize('comps', [[48, 12, 12], [16, 16, 16]])
def test_appearance_forward_backward(cuda, comps, sorted_bwd):
    planes, lines = factors(comps, RES, 8)
    pts = points(6000, BBOX, 9, outside=0.0)
    g = torch.Generator().manual_seed(10)
    idx = torch.nonzero(torch.rand(6000, generator=g) < 0.3).flatten().to(torch.int32)
    P64 = [p.double().requires_grad_(True) for p in planes]
    L64 = [l.double().requires_grad_(True) for l in lines]
    ref = ot.vm_appearance_features(P64, L64, ot.normalize_points(pts.double(), BBOX.double())[idx.long()])
    gout = torch.randn(ref.shape, generator=g)
    (ref * gout.double()).sum().backward()
    P = to_dev(planes, cuda, True)
    L = to_dev(lines, cuda, False)
    got = tensorf.vm_appearance(pts.to(cuda), P, L, BBOX.flatten().tolist(), idx=idx.to(cuda))
    (got * gout.to(cuda)).sum().backward()
    torch.cuda.synchronize()
    assert got.shape == ref.shape
    assert (got.detach().cpu().double() - ref.detach()).abs().max() < 1e-6
    for a, b in zip(P + L, P64 + L64):
        err = (a.grad.cpu().double() - b.grad).abs().max() / b.grad.abs().max()
        assert err < 1e-4, err


@pytest.mark.parametrize('n', [0, 1, 1000, 1024, 5000, 1 << 20])
def test_threshold_compact_is_boolean_indexing(cuda, n):
    g = torch.Generator().manual_seed(n)
    v = torch.rand(n, generator=g)
    idx, count = tensorf.threshold_compact(v.to(cuda), 0.7)
    torch.cuda.synchronize()
    ref = torch.nonzero(v > 0.7).flatten()
    c = int(count.item())
    assert c == ref.numel()
    assert torch.equal(idx[:c].cpu().long(), ref)


def test_volume_render_forward_backward(cuda):
    g = torch.Generator().manual_seed(11)
    w = torch.rand(64, 37, generator=g)
    v = torch.rand(64, 37, 3, generator=g)
    wd = w.to(cuda).requires_grad_(True)
    vd = v.to(cuda).requires_grad_(True)
    out = compositing.volume_render(vd, wd)
    go = torch.randn(64, 3, generator=g)
    (out * go.to(cuda)).sum().backward()
    w64 = w.double().requires_grad_(True)
    v64 = v.double().requires_grad_(True)
    ref = torch.sum(w64[..., None] * v64, dim=-2)
    (ref * go.double()).sum().backward()
    assert (out.detach().cpu().double() - ref.detach()).abs().max() < 1e-5
    assert (wd.grad.cpu().double() - w64.grad).abs().max() < 1e-5
    assert (vd.grad.cpu().double() - v64.grad).abs().max() < 1e-6


def test_aabb_samples_bit_exact(cuda):
    g = torch.Generator().manual_seed(12)
    o = torch.randn(500, 3, generator=g) * 3
    d = torch.randn(500, 3, generator=g)
    d[::7, 1] = 0.0
    u = torch.rand(500, generator=g)
    for uu in (None, u):
        ref = ot.aabb_z_vals(o, d, BBOX, near=0.05, far=8.0, step=0.0125, n_samples=96, u=uu)
        got = tensorf.aabb_samples(o.to(cuda), d.to(cuda), BBOX.flatten().tolist(), 0.05, 8.0, 0.0125, 96,
                                   u=None if uu is None else uu.to(cuda))
        torch.cuda.synchronize()
        assert torch.equal(got.cpu(), ref)


def test_composite_without_colour(cuda):
    """TensoRF composites weights first (rgb NULL), then colours with volume_render."""
    from srf_hip import synthetic
    import oracle.nerf as on
    r = synthetic.rays(32, seed=13)
    g = torch.Generator().manual_seed(13)
    z = on.stratified_z_vals(32, 50, 0., 1., u=torch.rand(32, 50, generator=g))
    sigma = torch.relu(torch.randn(32, 50, generator=g))
    ref = on.volume_rendering(sigma.double(), torch.zeros(32, 50, 3, dtype=torch.float64), z_vals_ndc=z.double(),
                              rays_o=r['rays_o'].double(), rays_d=r['rays_d'].double(),
                              rays_d_ndc=r['rays_d_ndc'].double(), ndc=True, distance_scale=25.0)
    sd = sigma.to(cuda).requires_grad_(True)
    out = compositing.composite(sd, None, z.to(cuda), r['rays_o'].to(cuda), r['rays_d'].to(cuda),
                                r['rays_d_ndc'].to(cuda), ndc=True, distance_scale=25.0)
    for k in ('weights', 'acc', 'depth_ndc'):
        assert (out[k].detach().cpu().double() - ref[k]).abs().max() < 1e-5, k
    out['weights'].sum().backward()
    assert torch.isfinite(sd.grad).all()


def test_sorted_backward_at_full_size_matches_sample_order(cuda):
    """At the 300^2 shape (line copies of 300 x 48 floats in LDS) and the 512^2 one (512 x 64, the
    largest LDS copy), with a device-side row count below the capacity: the cell-sorted appearance
    and density backwards equal the sample-order ones up to the order of FP32 additions."""
    for res, comps, n in ((300, [48, 48, 48], 200000), (512, [64, 64, 64], 120000)):
        planes, lines = factors(comps, [res, res, res], 31)
        pts = points(n, BBOX, 32, outside=0.0).to(cuda)
        P = to_dev(planes, cuda, True)
        L = to_dev(lines, cuda, True)
        idx = torch.randperm(n, generator=torch.Generator().manual_seed(33))[: n // 2].to(torch.int32).to(cuda)
        count = torch.tensor([n // 3], dtype=torch.int32, device=cuda)
        gout = torch.randn(n // 2, sum(comps), generator=torch.Generator().manual_seed(34)).to(cuda)
        grads = {}
        for mode in (True, False):
            tensorf.SORTED_BACKWARD = mode
            try:
                for t in P + L:
                    t.grad = None
                f = tensorf.vm_appearance(pts, P, L, BBOX.flatten().tolist(), idx=idx, n_rows=n // 2, count=count)
                (f[: n // 3] * gout[: n // 3]).sum().backward()
                sig, _ = tensorf.vm_density(pts, P, L, BBOX.flatten().tolist())
                (sig * gout[:, 0].repeat(2)[:n]).sum().backward()
                grads[mode] = [t.grad.detach().clone() for t in P + L]
            finally:
                tensorf.SORTED_BACKWARD = True
        for a, b in zip(grads[True], grads[False]):
            assert (a - b).abs().max().item() <= 1e-5 * b.abs().max().item()

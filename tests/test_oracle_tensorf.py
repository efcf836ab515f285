"""CPU pins of the TensoRF oracle (oracle/tensorf.py) against closed forms and library routines.

The reference ships no golden vectors for this path; these checks fix the axis assignment, the
align_corners=True sampling and the sampler against what the mathematics determines.
"""
import math

import pytest
import torch
import torch.nn.functional as F

import oracle.tensorf as ot


def rand_factors(comps, res, seed, dtype=torch.float64):
    """Planes [1,c,res[a1],res[a0]] and lines [1,c,res[v],1] as created at SimpleTensoRF09.py:1154-1165."""
    g = torch.Generator().manual_seed(seed)
    planes, lines = [], []
    for p, c in enumerate(comps):
        a0, a1 = ot.MATRIX_AXES[p]
        v = ot.VECTOR_AXES[p]
        planes.append(0.1 * torch.randn(1, c, res[a1], res[a0], generator=g, dtype=dtype))
        lines.append(0.1 * torch.randn(1, c, res[v], 1, generator=g, dtype=dtype))
    return planes, lines


def test_texel_centres_read_stored_values():
    """At a texel centre the density feature is sum_p sum_c M_p[c, y, x] V_p[c, z] exactly, with
    (x, y) = (pts[a0], pts[a1]) indexing (W, H) of plane p and pts[v] indexing line p."""
    res = [5, 7, 9]
    planes, lines = rand_factors([3, 2, 4], res, 0)
    idx = torch.tensor([[1, 2, 3], [4, 6, 8], [0, 0, 0]])
    pts_n = torch.stack([2 * idx[:, k].double() / (res[k] - 1) - 1 for k in range(3)], -1)
    got = ot.vm_density_features(planes, lines, pts_n)
    want = torch.zeros(3, dtype=torch.float64)
    for p in range(3):
        a0, a1 = ot.MATRIX_AXES[p]
        v = ot.VECTOR_AXES[p]
        for i in range(3):
            want[i] += (planes[p][0, :, idx[i, a1], idx[i, a0]] * lines[p][0, :, idx[i, v], 0]).sum()
    assert torch.allclose(got, want, rtol=0, atol=1e-14)


def test_separable_linear_factors_interpolate_exactly():
    """Factors that are linear in each coordinate are reproduced exactly by (bi)linear sampling,
    so the feature of a rank-1 field is the product of the three linear functions."""
    res = [6, 8, 10]
    lin = [lambda t, a=a: 0.5 + 0.25 * a + 0.3 * t for a in range(3)]  # functions of normalised coords
    planes, lines = [], []
    for p in range(3):
        a0, a1 = ot.MATRIX_AXES[p]
        v = ot.VECTOR_AXES[p]
        xs = torch.linspace(-1, 1, res[a0], dtype=torch.float64)
        ys = torch.linspace(-1, 1, res[a1], dtype=torch.float64)
        zs = torch.linspace(-1, 1, res[v], dtype=torch.float64)
        planes.append((lin[a1](ys)[:, None] * lin[a0](xs)[None, :])[None, None])
        lines.append(lin[v](zs)[None, None, :, None])
    g = torch.Generator().manual_seed(3)
    pts_n = torch.rand(200, 3, generator=g, dtype=torch.float64) * 2 - 1
    got = ot.vm_density_features(planes, lines, pts_n)
    want = 3 * lin[0](pts_n[:, 0]) * lin[1](pts_n[:, 1]) * lin[2](pts_n[:, 2])
    assert torch.allclose(got, want, rtol=0, atol=1e-12)


def test_grid_sample_matches_closed_form_bilinear():
    g = torch.Generator().manual_seed(4)
    img = torch.randn(3, 11, 7, generator=g, dtype=torch.float64)
    x = torch.rand(500, generator=g, dtype=torch.float64) * 2.2 - 1.1   # includes the zero-padded border
    y = torch.rand(500, generator=g, dtype=torch.float64) * 2.2 - 1.1
    ref = F.grid_sample(img[None], torch.stack([x, y], -1).view(1, -1, 1, 2), align_corners=True).view(3, -1)
    assert torch.allclose(ot.bilinear_align_corners(img, x, y), ref, rtol=0, atol=1e-13)


def test_appearance_features_are_channelwise_products():
    res = [5, 6, 7]
    planes, lines = rand_factors([4, 8, 4], res, 5)
    g = torch.Generator().manual_seed(6)
    pts_n = torch.rand(64, 3, generator=g, dtype=torch.float64) * 2 - 1
    app = ot.vm_appearance_features(planes, lines, pts_n)
    assert app.shape == (64, 16)
    # summing each plane's block of channels gives that plane's density contribution
    dens = ot.vm_density_features(planes, lines, pts_n)
    assert torch.allclose(app.sum(-1), dens, rtol=0, atol=1e-14)


def test_alpha_mask_trilinear_at_voxel_centres_and_rank1():
    g = torch.Generator().manual_seed(7)
    vol = (torch.rand(4, 5, 6, generator=g) > 0.5).double()
    bbox = torch.tensor([[-1.0, -2.0, 0.0], [1.0, 2.0, 3.0]], dtype=torch.float64)
    # voxel (z, y, x) centre in world coordinates
    zyx = torch.tensor([[1, 2, 3], [3, 4, 5], [0, 0, 0]])
    pts = torch.stack([bbox[0, k] + (bbox[1, k] - bbox[0, k]) * zyx[:, 2 - k].double() / (vol.shape[2 - k] - 1)
                       for k in range(3)], -1)
    got = ot.alpha_mask_sample(vol, bbox, pts)
    assert torch.equal(got, vol[zyx[:, 0], zyx[:, 1], zyx[:, 2]])
    # a trilinear (rank-1 affine) volume is reproduced exactly anywhere inside
    zs, ys, xs = (torch.linspace(0, 1, n, dtype=torch.float64) for n in vol.shape)
    lin = (1 + zs)[:, None, None] * (2 + ys)[None, :, None] * (3 - xs)[None, None, :]
    q = bbox[0] + (bbox[1] - bbox[0]) * torch.rand(100, 3, generator=g, dtype=torch.float64)
    t = (q - bbox[0]) / (bbox[1] - bbox[0])
    want = (1 + t[:, 2]) * (2 + t[:, 1]) * (3 - t[:, 0])
    assert torch.allclose(ot.alpha_mask_sample(lin, bbox, q), want, rtol=0, atol=1e-12)


def test_aabb_entry_closed_form():
    bbox = torch.tensor([[-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]])
    o = torch.tensor([[0.0, 0.0, -5.0], [3.0, 0.5, 0.0], [0.0, 0.0, 0.0]])
    d = torch.tensor([[0.0, 0.0, 1.0], [-1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    z = ot.aabb_z_vals(o, d, bbox, near=0.1, far=6.0, step=0.25, n_samples=4)
    # ray 0 enters at t = 4; ray 1 at t = 2 (the zero components of d become 1e-6 and give huge
    # negative or positive rates that the max over axes ignores); ray 2 starts inside -> near
    assert torch.allclose(z[:, 0], torch.tensor([4.0, 2.0, 0.1]))
    assert torch.allclose(z[0], torch.tensor([4.0, 4.25, 4.5, 4.75]))
    u = torch.tensor([0.5, 0.0, 0.25])
    zu = ot.aabb_z_vals(o, d, bbox, near=0.1, far=6.0, step=0.25, n_samples=4, u=u)
    assert torch.allclose(zu - z, (0.25 * u)[:, None].expand(3, 4))


def test_num_samples_rule():
    """compute_num_samples = min(max, round(|res| / 0.5)): a 128^3 grid gives round(443.4) = 443."""
    res = torch.tensor([128, 128, 128])
    assert ot.compute_num_samples(res, 0.5, 1000) == 443
    assert ot.compute_num_samples(res, 0.5, 300) == 300
    assert round(math.sqrt(3) * 128 / 0.5) == 443


def rand_cp_lines(rank, res, seed, dtype=torch.float64):
    """Lines [1,R,res[v_i],1] for v = (2, 1, 0), as created at SimpleTensoRF09.py:989-996."""
    g = torch.Generator().manual_seed(seed)
    return [0.1 * torch.randn(1, rank, res[ot.VECTOR_AXES[i]], 1, generator=g, dtype=dtype) for i in range(3)]


def test_cp_texel_centres_read_stored_values():
    """At a texel centre the CP feature is sum_r L0[r, z] L1[r, y] L2[r, x] exactly: line 0 runs
    along pts[2], line 1 along pts[1], line 2 along pts[0]."""
    res = [5, 7, 9]
    lines = rand_cp_lines(6, res, 11)
    idx = torch.tensor([[1, 2, 3], [4, 6, 8], [0, 0, 0], [2, 5, 7]])
    pts_n = torch.stack([2 * idx[:, k].double() / (res[k] - 1) - 1 for k in range(3)], -1)
    got = ot.cp_appearance_features(lines, pts_n)
    want = lines[0][0, :, idx[:, 2], 0] * lines[1][0, :, idx[:, 1], 0] * lines[2][0, :, idx[:, 0], 0]
    assert torch.allclose(got, want.T, rtol=0, atol=1e-15)
    assert torch.allclose(ot.cp_density_features(lines, pts_n), want.sum(0), rtol=0, atol=1e-15)


def test_cp_separable_linear_lines_interpolate_exactly():
    """Lines linear in their coordinate are reproduced exactly, so a rank-1 CP field is the
    product of the three linear functions everywhere in the box."""
    res = [6, 8, 10]
    lin = [lambda t, a=a: 0.5 + 0.25 * a + 0.3 * t for a in range(3)]
    lines = []
    for i in range(3):
        v = ot.VECTOR_AXES[i]
        lines.append(lin[v](torch.linspace(-1, 1, res[v], dtype=torch.float64))[None, None, :, None])
    g = torch.Generator().manual_seed(12)
    pts_n = torch.rand(300, 3, generator=g, dtype=torch.float64) * 2 - 1
    want = lin[0](pts_n[:, 0]) * lin[1](pts_n[:, 1]) * lin[2](pts_n[:, 2])
    assert torch.allclose(ot.cp_density_features(lines, pts_n), want, rtol=0, atol=1e-13)


def test_cp_equals_vm_with_outer_product_plane():
    """A CP component L0(z) L1(y) L2(x) is the VM component with plane 0 = L2 (x) L1 (an outer
    product on the (x, y) plane, whose bilinear sample is the product of the two linear samples)
    and line 0 = L0, the other planes zero: the CP path agrees with the pinned VM path."""
    res = [6, 7, 8]
    rank = 5
    lines = rand_cp_lines(rank, res, 13)
    plane0 = (lines[1][0, :, :, 0][:, :, None] * lines[2][0, :, :, 0][:, None, :])[None]   # [1,R,H=res1,W=res0]
    planes = [plane0, torch.zeros(1, rank, res[2], res[0], dtype=torch.float64),
              torch.zeros(1, rank, res[2], res[1], dtype=torch.float64)]
    vm_lines = [lines[0], torch.zeros(1, rank, res[1], 1, dtype=torch.float64),
                torch.zeros(1, rank, res[0], 1, dtype=torch.float64)]
    g = torch.Generator().manual_seed(14)
    pts_n = torch.rand(400, 3, generator=g, dtype=torch.float64) * 2 - 1
    cp = ot.cp_appearance_features(lines, pts_n)
    vm = ot.vm_appearance_features(planes, vm_lines, pts_n)[:, :rank]
    assert torch.allclose(cp, vm, rtol=0, atol=1e-15)


def test_ragged_render_gradient_matches_finite_differences():
    """The oracle's ragged non-NDC TensoRF render (every ray rendered on its own) differentiated by
    autograd (grid_sample's backward, cumprod, the colour MLP) agrees with central differences in
    FP64 on a tiny scene: this pins the reference gradient the GPU tests compare against."""
    g = torch.Generator().manual_seed(21)
    res = [5, 6, 7]
    planes_d, lines_d = rand_factors([4, 4, 4], res, 22)
    planes_c, lines_c = rand_factors([4, 4, 4], res, 23)
    planes_d = [p * 8 for p in planes_d]
    lin = lambda o, i: (torch.randn(o, i, generator=g, dtype=torch.float64) / i ** 0.5,
                        0.1 * torch.randn(o, generator=g, dtype=torch.float64))
    (w1, b1), (w2, b2), (w3, b3) = lin(16, 30), lin(16, 16), lin(3, 16)
    model = {'bbox': torch.tensor([[-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]], dtype=torch.float64),
             'planes_density': planes_d, 'lines_density': lines_d, 'planes_color': planes_c, 'lines_color': lines_c,
             'basis': torch.randn(27, 12, generator=g, dtype=torch.float64) / 12 ** 0.5,
             'color_mlp': {'mlp.0.weight': w1, 'mlp.0.bias': b1, 'mlp.2.weight': w2, 'mlp.2.bias': b2,
                           'mlp.4.weight': w3, 'mlp.4.bias': b3}}
    counts = [3, 7, 1, 5]
    off = torch.tensor([0] + list(torch.cumsum(torch.tensor(counts), 0)), dtype=torch.int32)
    o = torch.rand(4, 3, generator=g, dtype=torch.float64) * 0.8 - 0.4
    d = torch.nn.functional.normalize(torch.randn(4, 3, generator=g, dtype=torch.float64), dim=-1)
    ray = torch.repeat_interleave(torch.arange(4), torch.tensor(counts))
    k = torch.arange(sum(counts)) - off[:-1].long()[ray]
    z = 0.05 + 0.15 * (k.double() + 0.3)
    pts = o[ray] + d[ray] * z[:, None]
    gr = torch.randn(4, 3, generator=g, dtype=torch.float64)

    def loss():
        out = ot.tensorf_render_ragged(model, pts, z, off, o, d, d, distance_scale=25.0, weight_threshold=1e-4)
        return (out['rgb'] * gr).sum() + out['acc'].sum()

    targets = [(planes_d[0], (0, 1, 2, 3)), (lines_d[1], (0, 2, 3, 0)), (planes_c[2], (0, 3, 4, 2)),
               (model['basis'], (5, 7)), (w1, (3, 11))]
    for t, _ in targets:
        t.requires_grad_(True)
    base = loss()
    assert (ot.tensorf_render_ragged(model, pts, z, off, o, d, d)['weights'] > 1e-4).sum() >= 6
    base.backward()
    h = 1e-6
    for t, i in targets:
        with torch.no_grad():
            t[i] += h
            up = loss().item()
            t[i] -= 2 * h
            dn = loss().item()
            t[i] += h
        fd = (up - dn) / (2 * h)
        assert abs(t.grad[i].item() - fd) <= 1e-6 * max(1.0, abs(fd)), (i, t.grad[i].item(), fd)

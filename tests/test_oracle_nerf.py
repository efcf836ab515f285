"""CPU pins of the NeRF oracle: closed forms, brute force, library routines, invariants.

Each test is chosen so that a plausible restatement mistake (dropped term, wrong index,
swapped operand, off-by-one in the CDF) fails it; none re-types the oracle's formula."""
import math

import numpy as np
import pytest
import torch

import oracle.nerf as on

torch.set_default_dtype(torch.float32)


# ------------------------------------------------------------------ linspace formula
def _linspace_formula(start, end, steps):
    """The closed form the HIP sampler uses for torch.linspace (fma with a single rounding)."""
    step = np.float32(np.float32(end - start) / np.float32(steps - 1))
    half = steps // 2
    out = []
    for i in range(steps):
        if i < half:
            v = np.float64(step) * i + np.float64(np.float32(start))
        else:
            v = np.float64(np.float32(end)) - np.float64(step) * (steps - 1 - i)
        out.append(np.float32(v))
    return np.array(out, dtype=np.float32)


@pytest.mark.parametrize('steps', [2, 3, 5, 63, 64, 128, 192, 443, 1000])
@pytest.mark.parametrize('start,end', [(0., 1.), (2., 6.), (0.5, 40.)])
def test_linspace_formula_matches_torch(steps, start, end):
    ref = torch.linspace(start, end, steps=steps).numpy()
    assert np.array_equal(_linspace_formula(start, end, steps), ref)


# -------------------------------------------------------------------- stratified
def test_stratified_deterministic_is_linspace():
    z = on.stratified_z_vals(4, 64, 0., 1.)
    assert torch.equal(z[0], torch.linspace(0, 1, 64))
    assert torch.equal(z[3], z[0])


def test_stratified_u_bounds_and_sorted():
    n, s = 50, 64
    z_lo = on.stratified_z_vals(n, s, 0., 1., u=torch.zeros(n, s))
    z_hi = on.stratified_z_vals(n, s, 0., 1., u=torch.ones(n, s) * (1 - 1e-7))
    t = torch.linspace(0, 1, s)
    mids = 0.5 * (t[1:] + t[:-1])
    assert torch.allclose(z_lo[0], torch.cat([t[:1], mids]))       # u=0 -> lower bounds
    assert torch.allclose(z_hi[0], torch.cat([mids, t[-1:]]), atol=1e-6)
    u = torch.rand(n, s, generator=torch.Generator().manual_seed(0))
    z = on.stratified_z_vals(n, s, 0., 1., u=u)
    assert bool((z[:, 1:] >= z[:, :-1]).all()) and z.min() >= 0 and z.max() <= 1


def test_stratified_lindisp_inverse_depth_uniform():
    z = on.stratified_z_vals(1, 9, 1., 5., lindisp=True, dtype=torch.float64)[0]
    inv = 1 / z
    assert torch.allclose(inv[1:] - inv[:-1], torch.full((8,), (1 / 5 - 1) / 8, dtype=torch.float64))


# -------------------------------------------------------------------- sample_pdf
def _brute_inds(cdf, u):
    return (cdf[:, None, :] <= u[:, :, None]).sum(-1)


def test_searchsorted_is_count_of_le():
    g = torch.Generator().manual_seed(1)
    w = torch.rand(200, 62, generator=g)
    w[::3, 10:30] = 0
    cdf = on.pdf_cdf(w)
    u = torch.rand(200, 128, generator=g)
    u[:, 0] = 0.0
    u[:, 1] = 1.0
    u[5, 2:10] = cdf[5, 7]                         # exact ties
    _, inds, _ = on.sample_pdf(torch.sort(torch.rand(200, 63, generator=g), -1)[0], w, 128, u=u, cdf=cdf)
    assert torch.equal(inds, _brute_inds(cdf, u))


def test_cdf_construction():
    w = torch.tensor([[0., 1., 3.]], dtype=torch.float64)
    cdf = on.pdf_cdf(w)[0]
    ww = w[0] + 1e-5
    expect = torch.cat([torch.zeros(1, dtype=torch.float64), torch.cumsum(ww / ww.sum(), 0)])
    assert torch.allclose(cdf, expect) and cdf[0] == 0 and abs(cdf[-1] - 1) < 1e-12


def test_flat_pdf_inverts_linearly():
    bins = torch.linspace(0.2, 0.8, 63, dtype=torch.float64)[None]
    w = torch.zeros(1, 62, dtype=torch.float64)
    u = torch.rand(1, 500, generator=torch.Generator().manual_seed(2), dtype=torch.float64)
    s, _, _ = on.sample_pdf(bins, w, 500, u=u)
    assert torch.allclose(s, 0.2 + 0.6 * u, atol=1e-10)


def test_spike_concentrates_samples():
    bins = torch.linspace(0., 1., 63, dtype=torch.float64)[None]
    w = torch.zeros(1, 62, dtype=torch.float64)
    w[0, 40] = 1e3
    s, _, _ = on.sample_pdf(bins, w, 256)
    inside = ((s >= bins[0, 40]) & (s <= bins[0, 41])).double().mean()
    assert inside > 0.99


def test_det_endpoints_and_denom_rule():
    bins = torch.linspace(0., 1., 63, dtype=torch.float64)[None]
    w = torch.rand(1, 62, generator=torch.Generator().manual_seed(3), dtype=torch.float64)
    s, inds, cdf = on.sample_pdf(bins, w, 128)
    assert inds[0, 0] == 1 and s[0, 0] == 0.0              # u=0 -> first bin start
    # u=1: index 63 (clamped to 62/62, denom 0 -> 1) or 62 if rounding left cdf[-1] above 1
    assert inds[0, -1] == (cdf[0] <= 1.0).sum() and abs(s[0, -1] - bins[0, 62]) < 1e-12
    # zero-width cdf step (< 1e-5): t = u - cdf_below, not a division by ~0
    cdf2 = torch.tensor([[0., 0.5, 0.5 + 1e-7, 1.]], dtype=torch.float64)
    s2, _, _ = on.sample_pdf(torch.tensor([[0., 1., 2., 3.]], dtype=torch.float64), None, 1,
                             u=torch.tensor([[0.5 + 5e-8]], dtype=torch.float64), cdf=cdf2)
    assert abs(s2.item() - (1.0 + 5e-8 * 1.0)) < 1e-12


def test_fine_z_vals_is_sorted_union():
    g = torch.Generator().manual_seed(4)
    zc = on.stratified_z_vals(30, 64, 0., 1., u=torch.rand(30, 64, generator=g))
    w = torch.rand(30, 64, generator=g)
    zm, zf = on.fine_z_vals(zc, w, 128, u=torch.rand(30, 128, generator=g))
    assert zm.shape == (30, 192)
    assert bool((zm[:, 1:] >= zm[:, :-1]).all())
    assert torch.equal(torch.sort(torch.cat([zc, zf], -1), -1)[0], zm)
    mids = 0.5 * (zc[:, 1:] + zc[:, :-1])
    assert bool((zf >= mids[:, :1] - 1e-7).all() and (zf <= mids[:, -1:] + 1e-7).all())


# ---------------------------------------------------------------------------- PE
@pytest.mark.parametrize('deg,dim', [(10, 63), (4, 27), (3, 21), (0, 3)])
def test_positional_encoding_closed_form(deg, dim):
    x = torch.tensor([[0.3, -1.2, 0.7]], dtype=torch.float64)
    e = on.positional_encoding(x, deg)[0]
    assert e.shape == (dim,)
    assert torch.equal(e[:3], x[0])
    for k in range(deg):
        for c in range(3):
            assert math.isclose(e[3 + 6 * k + c].item(), math.sin(2 ** k * x[0, c].item()), abs_tol=1e-12)
            assert math.isclose(e[6 + 6 * k + c].item(), math.cos(2 ** k * x[0, c].item()), abs_tol=1e-12)


def test_points_augmented_split_is_prefix():
    """The 21-dim trunk input of the points-augmented model is the L=3 encoding."""
    x = torch.rand(10, 3, dtype=torch.float64)
    assert torch.equal(on.positional_encoding(x, 10)[:, :21], on.positional_encoding(x, 3))


# ---------------------------------------------------------------------- MLP shapes
def test_mlp_param_shapes_match_reference_geometry():
    s = on.mlp_param_shapes()
    assert s['pts_linears.0.weight'] == (256, 63) and s['pts_linears.5.weight'] == (256, 319)
    assert s['views_linears.0.weight'] == (128, 283) and s['pts_output_linear.weight'] == (1, 256)
    s = on.mlp_param_shapes(sigma_pe_degree=3)
    assert s['pts_linears.0.weight'] == (256, 21) and s['pts_linears.5.weight'] == (256, 277)
    assert s['views_linears.0.weight'] == (128, 325)
    s = on.mlp_param_shapes(use_view_dirs=False, view_dependent_rgb=False)
    assert s['pts_output_linear.weight'] == (4, 256) and 'feature_linear.weight' not in s
    macs = lambda sh: sum(v[0] * v[1] for k, v in sh.items() if k.endswith('weight'))
    assert macs(on.mlp_param_shapes()) == 593408
    assert macs(on.mlp_param_shapes(sigma_pe_degree=3)) == 577280
    assert macs(on.mlp_param_shapes(use_view_dirs=False, view_dependent_rgb=False)) == 492032


def _rand_params(shapes, seed):
    g = torch.Generator().manual_seed(seed)
    return {k: torch.randn(v, generator=g, dtype=torch.float64) * 0.2 for k, v in shapes.items()}


def test_mlp_gradients_vs_finite_differences():
    """Autograd of the oracle MLP against FP64 finite differences: full gradcheck on the head
    biases, directional derivatives (one random direction per tensor) on the large weights."""
    p = _rand_params(on.mlp_param_shapes(), 0)
    gx = torch.Generator().manual_seed(11)   # own generator: the draw must not depend on test order
    x = torch.rand(3, 3, dtype=torch.float64, generator=gx)
    d = torch.nn.functional.normalize(torch.rand(3, 3, dtype=torch.float64, generator=gx), dim=-1)

    def loss(q):
        o = on.mlp_forward(q, x, d)
        return o['sigma'].sum() + o['rgb'].sum()

    names = ['pts_output_linear.bias', 'views_output_linear.bias', 'views_linears.0.bias']
    def f(*ts):
        q = dict(p)
        q.update(dict(zip(names, ts)))
        return loss(q)
    assert torch.autograd.gradcheck(f, [p[n].clone().requires_grad_(True) for n in names], eps=1e-6, atol=1e-6)
    g = torch.Generator().manual_seed(1)
    for n in ['pts_linears.0.weight', 'pts_linears.5.weight', 'feature_linear.weight', 'views_linears.0.weight']:
        q = {k: v.clone().requires_grad_(k == n) for k, v in p.items()}
        grad, = torch.autograd.grad(loss(q), [q[n]])
        v = torch.randn(p[n].shape, generator=g, dtype=torch.float64)
        h = 1e-6
        qp, qm = dict(p), dict(p)
        qp[n] = p[n] + h * v
        qm[n] = p[n] - h * v
        fd = (loss(qp) - loss(qm)) / (2 * h)
        assert abs(fd.item() - (grad * v).sum().item()) < 1e-6 * max(1.0, abs(fd.item())), n


# -------------------------------------------------------------------- compositing
def _rays(n, seed=0, dtype=torch.float64):
    from srf_hip import synthetic
    r = synthetic.rays(n, seed=seed)
    return {k: v.to(dtype) if v.is_floating_point() else v for k, v in r.items()}


def _vr(sigma, rgb, z, r, **kw):
    return on.volume_rendering(sigma, rgb, z_vals_ndc=z, rays_o=r['rays_o'], rays_d=r['rays_d'],
                               rays_d_ndc=r['rays_d_ndc'], **kw)


def test_zero_density_is_empty():
    r = _rays(4)
    z = on.stratified_z_vals(4, 16, 0., 1., dtype=torch.float64)
    out = _vr(torch.zeros(4, 16, dtype=torch.float64), torch.rand(4, 16, 3, dtype=torch.float64), z, r)
    assert torch.all(out['weights'] == 0) and torch.all(out['acc'] == 0) and torch.all(out['rgb'] == 0)
    out = _vr(torch.zeros(4, 16, dtype=torch.float64), torch.rand(4, 16, 3, dtype=torch.float64), z, r,
              white_bkgd=True)
    assert torch.allclose(out['rgb'], torch.ones(4, 3, dtype=torch.float64))


def test_opaque_sample_gives_one_hot_weights():
    r = _rays(3)
    z = on.stratified_z_vals(3, 10, 0., 1., dtype=torch.float64)
    sigma = torch.zeros(3, 10, dtype=torch.float64)
    sigma[:, 4] = 1e6
    rgb = torch.rand(3, 10, 3, dtype=torch.float64)
    out = _vr(sigma, rgb, z, r)
    assert torch.allclose(out['weights'][:, 4], torch.ones(3, dtype=torch.float64))
    assert torch.allclose(out['rgb'], rgb[:, 4])
    assert torch.allclose(out['depth_ndc'], z[:, 4] / (1 + 1e-6), rtol=1e-9)
    assert torch.all(out['weights'][:, 5:].abs() < 1e-9)


def test_constant_density_geometric_weights():
    """Uniform spacing and constant sigma: w_k = (1-a) (a + 1e-10)^k with a = exp(-sigma delta)."""
    n, s = 2, 12
    z = torch.linspace(0., 1., s, dtype=torch.float64).expand(n, s).contiguous()
    r = _rays(n)
    sig = torch.full((n, s), 3.0, dtype=torch.float64)
    out = _vr(sig, torch.rand(n, s, 3, dtype=torch.float64), z, r)
    dn = torch.norm(r['rays_d_ndc'], dim=-1)
    a = torch.exp(-3.0 * (z[0, 1] - z[0, 0]) * dn)
    k = torch.arange(s - 1, dtype=torch.float64)
    expect = (1 - a[:, None]) * (a[:, None] + 1e-10) ** k
    assert torch.allclose(out['weights'][:, :-1], expect, rtol=1e-12, atol=1e-15)
    assert torch.all(out['weights'][:, -1] == 0)         # last delta is (1 - 1) * |d| = 0


def test_transmittance_telescoping_identity():
    """sum_k w_k = 1 - T_S q... : acc = 1 - T_last*(1-alpha_last) + 1e-10 sum_k T_k (exact algebra)."""
    r = _rays(16)
    g = torch.Generator().manual_seed(5)
    z = on.stratified_z_vals(16, 40, 0., 1., u=torch.rand(16, 40, generator=g, dtype=torch.float64),
                             dtype=torch.float64)
    sig = torch.rand(16, 40, generator=g, dtype=torch.float64) * 30
    out = _vr(sig, torch.rand(16, 40, 3, generator=g, dtype=torch.float64), z, r)
    T, al = out['visibility'], out['alpha']
    T_end = T[:, -1] * (1 - al[:, -1] + 1e-10)
    assert torch.allclose(out['acc'], 1 - T_end + 1e-10 * T.sum(-1), atol=1e-12)
    assert bool((out['acc'] <= 1 + 1e-9).all())


def test_depth_ndc_round_trip():
    r = _rays(20)
    zm = torch.rand(20, 1, dtype=torch.float64) * 10 + 1.5
    zn = on.convert_depth_to_ndc(zm, r['rays_o'], r['rays_d'], near=1.0)
    back = on.convert_depth_from_ndc(zn, r['rays_o'], r['rays_d'])
    assert torch.allclose(back, zm, rtol=1e-9)
    # z_ndc == 1 uses the +1e-3 guard instead of dividing by zero
    inf = on.convert_depth_from_ndc(torch.ones(20, 1, dtype=torch.float64), r['rays_o'], r['rays_d'])
    assert torch.isfinite(inf).all()


def test_volume_rendering_gradcheck():
    r = _rays(3)
    g = torch.Generator().manual_seed(6)
    z = on.stratified_z_vals(3, 8, 0., 1., u=torch.rand(3, 8, generator=g, dtype=torch.float64), dtype=torch.float64)
    sig = (torch.rand(3, 8, generator=g, dtype=torch.float64) * 5).requires_grad_(True)
    rgb = torch.rand(3, 8, 3, generator=g, dtype=torch.float64).requires_grad_(True)
    def f(s, c):
        o = _vr(s, c, z, r)
        return o['rgb'].sum() + o['depth'].sum() * 0.1 + o['depth_var'].sum() * 0.01 + o['acc'].sum()
    assert torch.autograd.gradcheck(f, (sig, rgb), eps=1e-6, atol=1e-5)


# -------------------------------------------------------------------------- rays
def test_rays_pinhole_and_ndc_geometry():
    from srf_hip import synthetic
    cam = synthetic.camera(seed=0)
    pid = torch.tensor([[0, 252, 189]], dtype=torch.int32)          # principal point
    K = cam['K'][None]
    E = torch.eye(4)[None]
    ro, rd = on.get_rays(pid, K, E)
    assert torch.allclose(rd, torch.tensor([[0., 0., -1.]]), atol=1e-6)     # looks down -z
    pid = torch.tensor([[0, 252 + 40, 189]], dtype=torch.int32)
    _, rd = on.get_rays(pid, K, E)
    assert torch.allclose(rd[0, 0], torch.tensor(40. / K[0, 0, 0]))
    # NDC: the point on the ray at the near plane maps to z_ndc = -1 ... and z -> -inf maps to +1
    ro = torch.zeros(1, 3)
    o_ndc, d_ndc = on.get_ndc_rays(ro, rd, (378, 504), K, near=1.0)
    assert torch.allclose(o_ndc[0, 2], torch.tensor(-1.0))
    assert torch.allclose(o_ndc[0, 2] + d_ndc[0, 2], torch.tensor(1.0))

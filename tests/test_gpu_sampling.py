"""GPU parity of the sampling kernels against the oracle (bit-exact where the reference's
arithmetic is reproducible op by op)."""
import pytest
import torch

import oracle.nerf as on
from srf_hip import sampling

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('n_samples', [64, 128, 192, 443, 5, 1])
@pytest.mark.parametrize('perturb', [False, True])
def test_stratified_bit_exact(cuda, n_samples, perturb):
    n_rays = 257
    g = torch.Generator().manual_seed(n_samples)
    u = torch.rand(n_rays, n_samples, generator=g) if perturb else None
    ref = on.stratified_z_vals(n_rays, n_samples, 0., 1., u=u)
    got = sampling.stratified_z_vals(n_rays, n_samples, 0., 1., u=None if u is None else u.to(cuda), device=cuda)
    assert torch.equal(got.cpu(), ref)


@pytest.mark.parametrize('near,far,lindisp', [(2.0, 6.0, False), (0.5, 40.0, True)])
def test_stratified_metric_range_bit_exact(cuda, near, far, lindisp):
    u = torch.rand(100, 64, generator=torch.Generator().manual_seed(3))
    ref = on.stratified_z_vals(100, 64, near, far, lindisp=lindisp, u=u)
    got = sampling.stratified_z_vals(100, 64, near, far, lindisp=lindisp, u=u.to(cuda))
    assert torch.equal(got.cpu(), ref)


def _coarse_weights(n_rays, n, seed):
    g = torch.Generator().manual_seed(seed)
    w = torch.rand(n_rays, n, generator=g) ** 4
    w[::7] = 0.0                    # flat pdf rows (all 1e-5)
    w[1::11, 20] = 50.0             # spikes
    w[2::13, :30] = 0.0
    return w


@pytest.mark.parametrize('det', [True, False])
def test_sample_pdf_given_cdf_bit_exact(cuda, det):
    n_rays, n_bins, n_fine = 1031, 63, 128
    g = torch.Generator().manual_seed(1)
    bins = torch.sort(torch.rand(n_rays, n_bins, generator=g), -1)[0]
    w = _coarse_weights(n_rays, n_bins - 1, 2)
    cdf = on.pdf_cdf(w)
    u = None if det else torch.rand(n_rays, n_fine, generator=g)
    if u is not None:
        u[:5, :3] = 0.0
        u[5:9, -1] = 1.0
        u[9, :] = cdf[9, 10]        # ties with a cdf entry (right=True semantics)
    ref_s, ref_i, _ = on.sample_pdf(bins, w, n_fine, u=u, cdf=cdf)
    s, i = sampling.sample_pdf(bins.to(cuda), None, n_fine, u=None if u is None else u.to(cuda), cdf=cdf.to(cuda),
                               return_inds=True)
    assert torch.equal(i.cpu(), ref_i)
    assert torch.equal(s.cpu(), ref_s)


def test_sample_pdf_internal_cdf(cuda):
    """With its own CDF the kernel is as close to the FP64 result as the FP32 oracle is.
    (Bit-exactness is defined given the same CDF, tested above; here torch's float
    normaliser sum and the kernel's may differ by an ulp, which narrow CDF steps amplify.)"""
    n_rays, n_bins, n_fine = 2048, 63, 128
    g = torch.Generator().manual_seed(4)
    bins = torch.sort(torch.rand(n_rays, n_bins, generator=g), -1)[0]
    w = _coarse_weights(n_rays, n_bins - 1, 5)
    u = torch.rand(n_rays, n_fine, generator=g)
    ref_s, ref_i, ref_cdf = on.sample_pdf(bins, w, n_fine, u=u)
    s64, _, _ = on.sample_pdf(bins.double(), w.double(), n_fine, u=u.double())
    s, i, cdf = sampling.sample_pdf(bins.to(cuda), w.to(cuda), n_fine, u=u.to(cuda), return_inds=True,
                                    return_cdf=True)
    # cumsum accumulates in double as torch's CPU kernel does; the normaliser may differ by 1 ulp
    assert (cdf.cpu() - ref_cdf).abs().max().item() <= 5e-7
    assert (i.cpu() != ref_i).float().mean().item() < 1e-3   # only u within an ulp of a cdf entry
    err_k = (s.cpu().double() - s64).abs()
    err_o = (ref_s.double() - s64).abs()
    assert err_k.max() <= 2 * err_o.max() + 1e-6
    assert err_k.mean() <= 2 * err_o.mean() + 1e-8


@pytest.mark.parametrize('det', [True, False])
@pytest.mark.parametrize('n_coarse,n_fine', [(64, 128), (64, 64), (32, 16), (192, 96)])
def test_fine_z_vals(cuda, det, n_coarse, n_fine):
    n_rays = 1500
    g = torch.Generator().manual_seed(7)
    z = on.stratified_z_vals(n_rays, n_coarse, 0., 1., u=torch.rand(n_rays, n_coarse, generator=g))
    w = _coarse_weights(n_rays, n_coarse, 8)
    u = None if det else torch.rand(n_rays, n_fine, generator=g)
    ref_m, ref_f = on.fine_z_vals(z, w, n_fine, u=u)
    _, f64 = on.fine_z_vals(z.double(), w.double(), n_fine, u=None if u is None else u.double())
    m, f = sampling.fine_z_vals(z.to(cuda), w.to(cuda), n_fine, u=None if u is None else u.to(cuda),
                                return_fine=True)
    m, f = m.cpu(), f.cpu()
    # sortedness and multiset are exact by construction
    assert bool((m[:, 1:] >= m[:, :-1]).all())
    assert torch.equal(torch.sort(torch.cat([z, f], -1), -1)[0], m)
    # the fine samples are as close to FP64 as the FP32 oracle's (see test_sample_pdf_internal_cdf)
    err_k = (f.double() - f64).abs()
    err_o = (ref_f.double() - f64).abs()
    assert err_k.max() <= 2 * err_o.max() + 1e-6
    assert err_k.mean() <= 2 * err_o.mean() + 1e-8
    # ... and given the oracle's CDF the merge reproduces torch.sort(cat) of the oracle bit for bit
    mids = .5 * (z[..., 1:] + z[..., :-1])
    cdf = on.pdf_cdf(w[..., 1:-1])
    fs = sampling.sample_pdf(mids.to(cuda), None, n_fine, u=None if u is None else u.to(cuda), cdf=cdf.to(cuda))
    assert torch.equal(fs.cpu(), ref_f)


def test_fine_z_vals_given_same_samples_is_sort(cuda):
    """When the fine samples coincide, the merge equals torch.sort bit for bit (incl. ties)."""
    n_rays = 300
    z = on.stratified_z_vals(n_rays, 64, 0., 1.)            # deterministic -> ties with det u
    w = torch.zeros(n_rays, 64)                              # flat pdf -> fine samples on the grid
    m, f = sampling.fine_z_vals(z.to(cuda), w.to(cuda), 128, return_fine=True)
    assert torch.equal(m.cpu(), torch.sort(torch.cat([z, f.cpu()], -1), -1)[0])


def test_empty_batch(cuda):
    z = sampling.stratified_z_vals(0, 64, 0., 1., device=cuda)
    assert z.shape == (0, 64)

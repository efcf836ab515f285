"""The oracle's NeRF MLP and TensoRF colour network against vectors written by the reference's own
classes (tests/golden/mlp_golden.npz, made by tests/golden/make_golden.py from
src/models/SimpleNeRF17.py:616-785 and src/models/SimpleTensoRF09.py:1370-1421).

Layer shapes and gradchecks cannot tell the concatenation orders apart; these vectors can.  The
second group of tests shows it: the oracle run with each concatenation's operands swapped (done by
permuting the consuming layer's weight columns, which is the same function) misses the vectors by
orders of magnitude more than the tolerance.
"""
import pathlib

import numpy as np
import pytest
import torch

from oracle import nerf as onerf
from oracle import tensorf as otensorf

GOLDEN = pathlib.Path(__file__).resolve().parent / 'golden' / 'mlp_golden.npz'

NERF_VARIANTS = {
    'main': dict(),
    'points_aug': dict(sigma_pe_degree=3),
    'views_aug': dict(use_view_dirs=False, view_dependent_rgb=False),
}
COLOR_VARIANTS = {'shipped': (0, 0), 'pe': (2, 2)}
TOL = {torch.float32: 1e-6, torch.float64: 1e-12}


@pytest.fixture(scope='module')
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _params(golden, prefix, dtype):
    head = f'{prefix}/param/'
    return {k[len(head):]: torch.from_numpy(v).to(dtype) for k, v in golden.items() if k.startswith(head)}


def _nerf(golden, variant, dtype, params=None):
    p = params if params is not None else _params(golden, f'nerf_{variant}', dtype)
    pts = torch.from_numpy(golden['pts']).to(dtype)
    dirs = torch.from_numpy(golden['dirs']).to(dtype)
    return onerf.mlp_forward(p, pts, dirs, **NERF_VARIANTS[variant])


def _color(golden, variant, dtype, params=None):
    p = params if params is not None else _params(golden, f'color_{variant}', dtype)
    fpe, vpe = COLOR_VARIANTS[variant]
    feats = torch.from_numpy(golden['features']).to(dtype)
    dirs = torch.from_numpy(golden['dirs']).to(dtype)
    return otensorf.color_mlp_forward(p, feats, dirs, features_pe_degree=fpe, views_pe_degree=vpe)


@pytest.mark.parametrize('dtype', [torch.float32, torch.float64], ids=['f32', 'f64'])
@pytest.mark.parametrize('variant', list(NERF_VARIANTS))
def test_nerf_mlp_matches_reference_vectors(golden, variant, dtype):
    tag = 'f32' if dtype == torch.float32 else 'f64'
    out = _nerf(golden, variant, dtype)
    for key in ('sigma', 'rgb'):
        ref = torch.from_numpy(golden[f'nerf_{variant}/{tag}/{key}'])
        assert out[key].shape == ref.shape
        err = (out[key] - ref).abs().max().item()
        assert err <= TOL[dtype], (variant, key, err)
    # the vectors are informative: densities are not all clamped and colours are not saturated
    assert (out['sigma'] > 0).float().mean() > 0.2
    assert out['rgb'].std() > 1e-2


@pytest.mark.parametrize('dtype', [torch.float32, torch.float64], ids=['f32', 'f64'])
@pytest.mark.parametrize('variant', list(COLOR_VARIANTS))
def test_color_mlp_matches_reference_vectors(golden, variant, dtype):
    tag = 'f32' if dtype == torch.float32 else 'f64'
    out = _color(golden, variant, dtype)
    ref = torch.from_numpy(golden[f'color_{variant}/{tag}/color'])
    assert out.shape == ref.shape
    assert (out - ref).abs().max().item() <= TOL[dtype]
    assert out.std() > 1e-2


def _swap_columns(w, len_a, len_b):
    """w' such that w' applied to cat([a, b]) computes what w applied to the swapped cat([b, a])
    does: the oracle run on w' is the oracle with that concatenation written the other way round."""
    assert w.shape[1] == len_a + len_b
    return torch.cat([w[:, len_b:], w[:, :len_b]], dim=1)


def _miss(out, ref):
    return (out - torch.from_numpy(ref)).abs().max().item()


def test_vectors_reject_swapped_skip_concat(golden):
    """cat([x_in, h]) at the skip (SimpleNeRF17.py:732-733) against cat([h, x_in])."""
    for variant, trunk in (('main', 63), ('points_aug', 21), ('views_aug', 63)):
        p = _params(golden, f'nerf_{variant}', torch.float64)
        p['pts_linears.5.weight'] = _swap_columns(p['pts_linears.5.weight'], trunk, 32)
        out = _nerf(golden, variant, torch.float64, p)
        assert _miss(out['sigma'], golden[f'nerf_{variant}/f64/sigma']) > 1e-3 or \
            _miss(out['rgb'], golden[f'nerf_{variant}/f64/rgb']) > 1e-3, variant


def test_vectors_reject_swapped_view_head_concats(golden):
    """cat([feature, PE_dir]) (:764) and, for the points-augmented model, cat([feature, PE_hi]) (:701-703)."""
    p = _params(golden, 'nerf_main', torch.float64)
    p['views_linears.0.weight'] = _swap_columns(p['views_linears.0.weight'], 32, 27)
    assert _miss(_nerf(golden, 'main', torch.float64, p)['rgb'], golden['nerf_main/f64/rgb']) > 1e-3
    p = _params(golden, 'nerf_points_aug', torch.float64)
    w = p['views_linears.0.weight']
    # [feature 32, PE_hi 42, PE_dir 27] -> [PE_hi 42, feature 32, PE_dir 27]
    p['views_linears.0.weight'] = torch.cat([_swap_columns(w[:, :74], 32, 42), w[:, 74:]], dim=1)
    assert _miss(_nerf(golden, 'points_aug', torch.float64, p)['rgb'], golden['nerf_points_aug/f64/rgb']) > 1e-3
    # [feature, PE_hi, PE_dir] -> [feature, PE_dir, PE_hi]
    p = _params(golden, 'nerf_points_aug', torch.float64)
    w = p['views_linears.0.weight']
    p['views_linears.0.weight'] = torch.cat([w[:, :32], _swap_columns(w[:, 32:], 42, 27)], dim=1)
    assert _miss(_nerf(golden, 'points_aug', torch.float64, p)['rgb'], golden['nerf_points_aug/f64/rgb']) > 1e-3


def test_vectors_reject_swapped_views_aug_channels(golden):
    """Views-augmented head: sigma is channel 0 and rgb channels 1..3 (:738-750)."""
    p = _params(golden, 'nerf_views_aug', torch.float64)
    for k in ('pts_output_linear.weight', 'pts_output_linear.bias'):
        p[k] = torch.cat([p[k][3:], p[k][:3]], dim=0)
    out = _nerf(golden, 'views_aug', torch.float64, p)
    assert _miss(out['sigma'], golden['nerf_views_aug/f64/sigma']) > 1e-3


def test_vectors_reject_swapped_color_input_order(golden):
    """[PE(features), PE(view)] (SimpleTensoRF09.py:1415-1418) against [PE(view), PE(features)]."""
    for variant, (fpe, vpe) in COLOR_VARIANTS.items():
        p = _params(golden, f'color_{variant}', torch.float64)
        p['mlp.0.weight'] = _swap_columns(p['mlp.0.weight'], 27 * (2 * fpe + 1), 3 * (2 * vpe + 1))
        assert _miss(_color(golden, variant, torch.float64, p), golden[f'color_{variant}/f64/color']) > 1e-3

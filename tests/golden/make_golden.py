"""Writes tests/golden/mlp_golden.npz: input/output vectors of the reference's own NeRF MLP and TensoRF
colour network, for the oracle functions that no closed form pins (their concatenation orders).

The reference classes are imported from the reference tree in THIS container only, instantiated with
seeded weights (narrower than the shipped 256/128 widths, so the fixture stays small; the wiring under
test does not depend on the width), put in eval() (no sigma noise) and evaluated on 64 seeded points
in float32 and float64.  The fixture holds only data: every variant's state_dict, the inputs and the
outputs, as plain arrays (numpy.load with allow_pickle=False reads it).

Variants (reference classes and configs):
  nerf_main        MLP, src/models/SimpleNeRF17.py:616-785, points PE 10, views PE 4, view-dependent rgb
  nerf_points_aug  same class with points_sigma_positional_encoding_degree 3 (the 21/42 split,
                   src/NerfLlffTrainerTester10.py:321-333)
  nerf_views_aug   same class with use_view_dirs False, view_dependent_rgb False (:347-358)
  color_shipped    MlpFeaturesColorPredictor, src/models/SimpleTensoRF09.py:1370-1421, features 27,
                   features PE 0, views PE 0 (src/NerfLlffTrainerTester10.py:958-961)
  color_pe         same class with features PE 2 and views PE 2 (both encoders and their order)

    python tests/golden/make_golden.py [/path/to/reference]
"""
import pathlib
import sys

import numpy as np
import torch

HERE = pathlib.Path(__file__).resolve().parent
N_POINTS = 64
NERF_WIDTH, NERF_VIEWS_WIDTH = 32, 16
COLOR_UNITS = 32


def nerf_mlp_configs(variant: str) -> dict:
    cfg = {'points_net_depth': 8, 'views_net_depth': 1, 'points_net_width': NERF_WIDTH,
           'views_net_width': NERF_VIEWS_WIDTH, 'points_positional_encoding_degree': 10,
           'views_positional_encoding_degree': 4, 'use_view_dirs': True, 'view_dependent_rgb': True,
           'predict_visibility': False}
    if variant == 'points_aug':
        cfg['points_sigma_positional_encoding_degree'] = 3
    elif variant == 'views_aug':
        cfg['use_view_dirs'] = False
        cfg['view_dependent_rgb'] = False
    return cfg


def color_tensor_configs(features_pe: int, views_pe: int) -> dict:
    return {'features_positional_encoding_degree': features_pe, 'views_positional_encoding_degree': views_pe,
            'use_view_dirs': True, 'view_dependent_color': True}


def seeded_(module: torch.nn.Module, seed: int) -> None:
    """Draws every parameter from one seeded generator: weights uniform with the variance 2/fan_in
    that keeps activations from shrinking through the ReLU trunk, biases uniform in [-0.1, 0.1], so
    that the biases are not zero and every input channel moves the outputs."""
    g = torch.Generator().manual_seed(seed)
    with torch.no_grad():
        for p in module.parameters():
            scale = (6.0 / p.shape[-1]) ** 0.5 if p.ndim == 2 else 0.1
            p.copy_((torch.rand(p.shape, generator=g, dtype=torch.float64) * 2 - 1) * scale)


def main(ref_root: str = '/root/reference') -> None:
    sys.path.insert(0, str(pathlib.Path(ref_root) / 'src'))
    from models.SimpleNeRF17 import MLP
    from models.SimpleTensoRF09 import MlpFeaturesColorPredictor

    g = torch.Generator().manual_seed(20240620)
    pts = (torch.rand(N_POINTS, 3, generator=g, dtype=torch.float64) * 2 - 1) * 1.2
    dirs = torch.randn(N_POINTS, 3, generator=g, dtype=torch.float64)
    dirs = dirs / dirs.norm(dim=-1, keepdim=True)
    features = torch.randn(N_POINTS, 27, generator=g, dtype=torch.float64) * 0.7
    arrays = {'pts': pts.numpy(), 'dirs': dirs.numpy(), 'features': features.numpy()}

    configs = {'model': {'raw_noise_std': 1.0}}
    for seed, variant in enumerate(('main', 'points_aug', 'views_aug')):
        mlp = MLP('coarse_model', configs, nerf_mlp_configs(variant), {})
        seeded_(mlp, 110 + seed)
        mlp.eval()
        for p in mlp.state_dict().items():
            arrays[f'nerf_{variant}/param/{p[0]}'] = p[1].numpy().astype(np.float32)
        for dtype, tag in ((torch.float32, 'f32'), (torch.float64, 'f64')):
            m = mlp.to(dtype)
            with torch.no_grad():
                out = m({'pts': pts.to(dtype), 'view_dirs': dirs.to(dtype)})
            arrays[f'nerf_{variant}/{tag}/sigma'] = out['sigma'].numpy()
            arrays[f'nerf_{variant}/{tag}/rgb'] = out['rgb'].numpy()
        mlp.float()

    for seed, (name, fpe, vpe) in enumerate((('shipped', 0, 0), ('pe', 2, 2))):
        cp = MlpFeaturesColorPredictor({}, color_tensor_configs(fpe, vpe), {}, 27, fpe, vpe, COLOR_UNITS)
        seeded_(cp, 200 + seed)
        cp.eval()
        for k, v in cp.state_dict().items():
            arrays[f'color_{name}/param/{k}'] = v.numpy().astype(np.float32)
        for dtype, tag in ((torch.float32, 'f32'), (torch.float64, 'f64')):
            m = cp.to(dtype)
            with torch.no_grad():
                out = m({'features': features.to(dtype), 'view_dirs': dirs.to(dtype)})
            arrays[f'color_{name}/{tag}/color'] = out['color'].numpy()
        cp.float()

    out_path = HERE / 'mlp_golden.npz'
    np.savez_compressed(out_path, **arrays)
    print('wrote', out_path, sum(a.nbytes for a in arrays.values()), 'bytes of arrays')


if __name__ == '__main__':
    main(*sys.argv[1:])

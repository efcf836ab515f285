"""The drop-in models built from the reference's own recorded run configs (tests/golden, written by
make_run_fixtures.py from runs/training/train1061 and train1546): only the model name changes.

The reference ships no golden vectors for the hot path; these fixtures are its real configuration
and camera files (NeRF-LLFF fern, 756 x 1008, f = 815.13).  The expected shapes follow from the
configs by hand: e.g. 2,097,156 voxels in the 3.0 x 3.34 x 2.0 box give a voxel side of
0.021221, hence 141 x 157 x 94 voxels and round(|(141, 157, 94)| / 0.5) = 462 samples.
"""
import copy
import json
import pathlib

import pytest
import torch

from models.ModelFactory02 import get_model
from srf_hip import configs as cfgs

GOLDEN = pathlib.Path(__file__).resolve().parent / 'golden'


def load(name, model_name):
    d = json.loads((GOLDEN / name).read_text())
    c = copy.deepcopy(d['configs'])
    c['model']['name'] = model_name
    return c, d['model_configs']


def test_nerf_dropin_builds_from_the_recorded_run():
    c, mc = load('llff_fern_nerf.json', 'SimpleNeRFHip17')
    model = get_model(c, model_configs=mc)
    sd = model.state_dict()
    # 8 x 256 trunk, skip after layer 4, 1 x 128 view head (SimpleNeRF17.py:644-666)
    for prefix, d_in, view_in in (('coarse_model.', 63, 283), ('fine_model.', 63, 283),
                                  ('augmented_models_nn.0.coarse_model.', 21, 325)):
        assert tuple(sd[f'{prefix}pts_linears.0.weight'].shape) == (256, d_in)
        assert tuple(sd[f'{prefix}pts_linears.5.weight'].shape) == (256, 256 + d_in)
        assert tuple(sd[f'{prefix}views_linears.0.weight'].shape) == (128, view_in)
    # views-augmented: sigma + rgb straight from the trunk (SimpleNeRF17.py:654-666)
    assert tuple(sd['augmented_models_nn.1.coarse_model.pts_output_linear.weight'].shape) == (4, 256)
    assert not any(k.startswith('augmented_models_nn.1.coarse_model.views_linears') for k in sd)
    # the synthetic configuration the benchmark runs is this run's model configuration
    ours = cfgs.simple_nerf_configs()['model']
    for k, v in c['model'].items():
        if k in ('name', 'chunk', 'netchunk'):
            continue
        assert ours[k] == v, k


def test_tensorf_dropin_builds_from_the_recorded_run():
    c, mc = load('llff_fern_tensorf.json', 'SimpleTensoRFHip09')
    model = get_model(c, model_configs=mc)
    tm = model.coarse_model
    assert tm.resolution.tolist() == [141, 157, 94]
    assert int(tm.num_samples) == 462
    aug = model.augmented_models[0]['coarse_model']
    assert aug.resolution.tolist() == [77, 86, 38]
    assert int(aug.num_samples) == 243
    sd = model.state_dict()
    assert tuple(sd['coarse_model.matrices_density.0'].shape) == (1, 16, 157, 141)
    assert tuple(sd['coarse_model.matrices_color.2'].shape) == (1, 12, 94, 157)
    assert tuple(sd['coarse_model.vectors_color.0'].shape) == (1, 48, 94, 1)
    assert tuple(sd['coarse_model.basis_matrix_color.weight'].shape) == (27, 72)
    assert tuple(sd['coarse_model.color_predictor.mlp.0.weight'].shape) == (128, 30)
    assert tuple(sd['augmented_models_nn.0.coarse_model.matrices_density.1'].shape) == (1, 4, 38, 77)


def test_recorded_cameras_are_llff_fern():
    for name in ('llff_fern_nerf.json', 'llff_fern_tensorf.json'):
        mc = json.loads((GOLDEN / name).read_text())['model_configs']
        K = torch.tensor(mc['intrinsics'])
        assert mc['resolution'] == [756, 1008]
        assert torch.allclose(K[:, 0, 0], torch.tensor(815.131591796875)) and (K[:, :2, 2] == torch.tensor([504., 378.])).all()
        E = torch.tensor(mc['extrinsics'])
        assert E.shape == (len(mc['train_frame_nums']), 4, 4)
        R = E[:, :3, :3].double()
        assert torch.allclose(R @ R.transpose(1, 2), torch.eye(3, dtype=torch.float64).expand_as(R), atol=1e-5)
        assert mc['near'] == 1.0 and mc['near_ndc'] == 0.0 and mc['far_ndc'] == 1.0


@pytest.mark.parametrize('name', ['llff_fern_nerf.json', 'llff_fern_tensorf.json'])
def test_fixture_keeps_the_run_configuration_whole(name):
    d = json.loads((GOLDEN / name).read_text())
    assert d['configs']['data_loader']['ndc'] is True
    assert d['configs']['database'] == 'NeRF_LLFF'

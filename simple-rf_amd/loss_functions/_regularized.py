"""The models a loss regularises, in the reference's order (MSE14.py:29-79, SparseDepthMSE14.py:30-80):
main coarse, main fine, then each augmentation's coarse and fine model, each with its weight from
loss_configs['models_to_regularize'] and present only when the model configuration has it."""
from __future__ import annotations


def regularized_outputs(configs: dict, models_to_regularize: dict, key: str):
    """[(suffix, output key, weight)] for output `key` ('rgb' or 'depth')."""
    out = []
    model = configs['model']
    for stage in ('coarse', 'fine'):
        if f'{stage}_model' in model and f'main_{stage}' in models_to_regularize:
            out.append((stage, f'{key}_{stage}', models_to_regularize[f'main_{stage}']))
    for aug in model.get('augmentations', []):
        name = aug['name']
        for stage in ('coarse', 'fine'):
            if f'{stage}_model' in aug and f'{name}_{stage}' in models_to_regularize:
                out.append((f'{name}_{stage}', f'{name}_{key}_{stage}', models_to_regularize[f'{name}_{stage}']))
    return out

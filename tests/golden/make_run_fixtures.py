"""Writes the run fixtures of tests/golden from the reference's recorded training runs.

runs/training/train1061 is a Simple-NeRF run and runs/training/train1546 a Simple-TensoRF run, both
on NeRF-LLFF.  For each, the fixture keeps the run's own configs (database, data_loader, model) and
the fern scene's camera configs (intrinsics, extrinsics, bounds, near/far, bounding box) as the
reference trainer wrote them.  Data only: nothing of the reference's code is read or run.

    python tests/golden/make_run_fixtures.py [/path/to/reference]
"""
import json
import pathlib
import sys

HERE = pathlib.Path(__file__).resolve().parent
RUNS = {'llff_fern_nerf.json': 'train1061', 'llff_fern_tensorf.json': 'train1546'}


def main(ref_root: str = '/root/reference') -> None:
    runs = pathlib.Path(ref_root) / 'runs' / 'training'
    for out_name, run in RUNS.items():
        configs = json.loads((runs / run / 'Configs.json').read_text())
        model_configs = json.loads((runs / run / 'fern' / 'ModelConfigs.json').read_text())
        fixture = {
            'source': f'runs/training/{run}/Configs.json and runs/training/{run}/fern/ModelConfigs.json',
            'configs': {k: configs[k] for k in ('database', 'data_loader', 'model')},
            'model_configs': model_configs,
        }
        (HERE / out_name).write_text(json.dumps(fixture, indent=1) + '\n')
        print('wrote', HERE / out_name)


if __name__ == '__main__':
    main(*sys.argv[1:])

"""Name-based model loader with the reference's semantics (src/models/ModelFactory02.py:10-22):
module models.<name>, class <name without its 2-digit version suffix>."""
import importlib
import inspect


def get_model(configs: dict, **kwargs):
    filename = configs['model']['name']
    classname = f'{filename[:-2]}'
    module = importlib.import_module(f'models.{filename}')
    for name, cls in inspect.getmembers(module, inspect.isclass):
        if name == classname:
            return cls(configs, **kwargs)
    raise RuntimeError(f'Unknown model: {filename}')

"""Model plugins loaded by name, as the reference's ModelFactory02 does
(src/models/ModelFactory02.py:10-22): configs['model']['name'] = 'SimpleNeRFHip17' selects
models/SimpleNeRFHip17.py, class SimpleNeRFHip."""

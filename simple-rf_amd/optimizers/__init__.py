"""Drop-in optimiser factory on the fused HIP Adam, loaded like the reference's
(src/optimizers/OptimizerFactory02.py)."""

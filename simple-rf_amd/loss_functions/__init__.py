"""Drop-in loss plugins on the fused HIP loss kernels, loaded by config name as the reference's
LossComputer loads its own (src/loss_functions/LossComputer03.py:15-26: module loss_functions.<name>,
class <name>[:-2])."""

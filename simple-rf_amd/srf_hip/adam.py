"""Adam on the fused HIP kernel (csrc/adam.hip): the optimiser the reference builds in
src/optimizers/OptimizerFactory02.py:9-21 (torch.optim.Adam with the config's betas and lr).

AdamHip is a torch.optim.Optimizer with torch.optim.Adam's parameter groups, hyperparameters and
state ('step' as a float32 device tensor, 'exp_avg', 'exp_avg_sq'), so the trainer's learning-rate
decayer, the TensoRF optimiser reconfiguration and checkpointed optimizer state dicts work
unchanged in both directions.  step() updates every parameter with a gradient in one launch per 32
tensors instead of torch's per-operation launches.  Under graph capture the learning rates are
those at capture time (they are kernel arguments), as for torch's Adam with float learning rates.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib


def _dense(t: torch.Tensor) -> bool:
    """Whether t covers its storage span exactly (any dimension order): a flat walk then visits
    every element once."""
    if t.numel() == 0:
        return True
    span = 1 + sum((s - 1) * abs(st) for s, st in zip(t.shape, t.stride()))
    return span == t.numel()


class AdamHip(torch.optim.Optimizer):
    def __init__(self, params, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8, weight_decay: float = 0.0):
        if not 0.0 <= lr:
            raise ValueError(f'Invalid learning rate: {lr}')
        if not 0.0 <= betas[0] < 1.0 or not 0.0 <= betas[1] < 1.0:
            raise ValueError(f'Invalid beta parameters: {betas}')
        defaults = dict(lr=lr, betas=betas, eps=eps, weight_decay=weight_decay, amsgrad=False, maximize=False)
        super().__init__(params, defaults)

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        by_key = {}
        for group in self.param_groups:
            if group.get('amsgrad') or group.get('maximize'):
                raise _lib.SrfError('AdamHip implements Adam without amsgrad and maximize, as the reference uses it')
            for p in group['params']:
                if p.grad is None:
                    continue
                g = p.grad
                if p.dtype != torch.float32 or g.dtype != torch.float32 or p.device.type != 'cuda':
                    raise _lib.SrfError('AdamHip updates FP32 GPU parameters')
                state = self.state[p]
                if len(state) == 0:
                    state['step'] = torch.zeros((), dtype=torch.float32, device=p.device)
                    state['exp_avg'] = torch.zeros_like(p, memory_format=torch.preserve_format)
                    state['exp_avg_sq'] = torch.zeros_like(p, memory_format=torch.preserve_format)
                elif state['step'].device != p.device:   # a state dict of torch's Adam keeps it on the host
                    state['step'] = state['step'].to(device=p.device, dtype=torch.float32)
                m, v = state['exp_avg'], state['exp_avg_sq']
                if not (p.stride() == g.stride() == m.stride() == v.stride() and _dense(p)):
                    raise _lib.SrfError('AdamHip needs parameter, gradient and moments in one dense layout')
                key = (group['betas'][0], group['betas'][1], group['eps'], p.device)
                by_key.setdefault(key, []).append((p, g, m, v, state['step'], group['lr'], group['weight_decay']))
        for (b1, b2, eps, dev), items in by_key.items():
            n = len(items)
            ptrs = lambda k: (C.c_void_p * n)(*[it[k].data_ptr() for it in items])
            numel = (C.c_int64 * n)(*[it[0].numel() for it in items])
            lrs = (C.c_float * n)(*[float(it[5]) for it in items])
            wds = (C.c_float * n)(*[float(it[6]) for it in items])
            _lib.call('srf_adam_step', n, ptrs(0), ptrs(1), ptrs(2), ptrs(3), ptrs(4), numel, lrs, wds, float(b1),
                      float(b2), float(eps), _lib.stream_ptr(dev))
        return loss

"""Optimizer stage of the stroke-field training step, fused into one HIP pass per tensor.

Drop-in for what the reference does to the stroke parameters after backward:

* ``clip_gradients`` (``source/utils/training.py:67-77``): optional global-norm
  clipping, optional value clipping, then ``grad.nan_to_num_()``;
* ``torch.optim.Adam`` as built by ``create_optimizer`` (``training.py:119-136``,
  betas (0.9, 0.99), eps 1e-15 from ``source/configs.py:82-84``);
* ``StrokeField.clip_params`` (``source/models.py:576-584``): per-column clamp of
  shape / colour parameters to their ranges, density >= 0.

``Adam`` is a ``torch.optim.Optimizer`` with torch's Adam state and update
formulas (same operation order, float32); a parameter group may carry
``ranges`` (one ``(lo, hi)`` per last-dimension column, ``None`` for an open
side) and ``min_value`` / ``max_value`` (bounds for every entry).  With
``capturable=True`` the step count lives on the device, and a group's ``lr``
may be a one-element device tensor, so a step captured in a HIP graph
replays with the current count and rate.

``learning_rate_decay`` is the reference's per-step rate schedule
(``training.py:80-116``).
"""
import math

import torch

from . import _lib


def learning_rate_decay(step, lr_init, lr_final, max_steps, lr_delay_steps=0, lr_delay_mult=1):
    """Log-linear decay from lr_init (step 0) to lr_final (max_steps), eased in by a
    reverse-cosine delay over the first lr_delay_steps (training.py:80-116)."""
    if lr_init <= 0 or lr_final <= 0:
        raise ValueError(f'Interpolants {lr_init} and {lr_final} must be positive.')
    if lr_delay_steps > 0:
        delay = lr_delay_mult + (1 - lr_delay_mult) * math.sin(0.5 * math.pi * min(max(step / lr_delay_steps, 0), 1))
    else:
        delay = 1.0
    t = min(max(step / max_steps, 0.0), 1.0)
    return delay * math.exp(t * (math.log(lr_final) - math.log(lr_init)) + math.log(lr_init))


class Adam(torch.optim.Optimizer):

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.99), eps=1e-15, grad_max_norm=0.0, grad_max_val=0.0,
                 capturable=False):
        defaults = dict(lr=lr, betas=betas, eps=eps, ranges=None, min_value=None, max_value=None)
        super().__init__(params, defaults)
        self.grad_max_norm = float(grad_max_norm)
        self.grad_max_val = float(grad_max_val)
        self.capturable = capturable
        self._bounds = {}
        self._norm = {}  # device -> (squared-norm accumulator, left at zero by each step; clip coefficient)

    def _col_bounds(self, group, p):
        ranges = group.get('ranges')
        if not ranges:
            return None, None
        key = (id(group), p.device)
        if key not in self._bounds:
            if len(ranges) != p.shape[-1]:
                raise RuntimeError(f'{len(ranges)} column ranges for a parameter of width {p.shape[-1]}')
            lo = torch.tensor([-math.inf if r[0] is None else float(r[0]) for r in ranges], dtype=torch.float32)
            hi = torch.tensor([math.inf if r[1] is None else float(r[1]) for r in ranges], dtype=torch.float32)
            self._bounds[key] = (lo.to(p.device), hi.to(p.device))
        return self._bounds[key]

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        params = [p for g in self.param_groups for p in g['params'] if p.grad is not None]
        if not params:
            return loss
        lib = _lib.lib()
        coef = None
        if self.grad_max_norm > 0:
            dev = params[0].device
            if dev not in self._norm:
                self._norm[dev] = (torch.zeros(1, dtype=torch.float64, device=dev),
                                   torch.empty(1, dtype=torch.float32, device=dev))
            acc, coef = self._norm[dev]
            for p in params:
                _lib.require_cuda('grad', p.grad)
                _lib.check(lib.n3ds_grad_sqnorm_accumulate(_lib.ptr(p.grad), p.grad.numel(), _lib.ptr(acc),
                                                           _lib.stream()))
            _lib.check(lib.n3ds_grad_clip_coef(_lib.ptr(acc), self.grad_max_norm, _lib.ptr(coef), _lib.stream()))
        for group in self.param_groups:
            b1, b2 = group['betas']
            for p in group['params']:
                if p.grad is None:
                    continue
                _lib.require_cuda('param', p)
                _lib.require_cuda('grad', p.grad)
                st = self.state[p]
                if not st:
                    st['step'] = torch.zeros((), dtype=torch.float32, device=p.device if self.capturable else 'cpu')
                    st['exp_avg'] = torch.zeros_like(p, memory_format=torch.preserve_format)
                    st['exp_avg_sq'] = torch.zeros_like(p, memory_format=torch.preserve_format)
                st['step'] += 1
                lo, hi = self._col_bounds(group, p)
                ncol = p.shape[-1] if p.dim() > 1 else 1
                lo_all = -math.inf if group.get('min_value') is None else float(group['min_value'])
                hi_all = math.inf if group.get('max_value') is None else float(group['max_value'])
                step_dev = st['step'] if self.capturable else None
                step = 0 if self.capturable else int(st['step'].item())
                lr = group['lr']
                lr_dev = None
                if torch.is_tensor(lr):
                    if not self.capturable:
                        lr = float(lr)
                    else:
                        _lib.require_cuda('lr', lr)
                        lr_dev, lr = lr, 0.0
                _lib.check(lib.n3ds_adam_step(
                    _lib.ptr(p), _lib.ptr(p.grad), _lib.ptr(st['exp_avg']), _lib.ptr(st['exp_avg_sq']),
                    p.numel() // ncol, ncol, _lib.ptr(lo), _lib.ptr(hi), float(lr), _lib.ptr(lr_dev), float(b1), float(b2),
                    float(group['eps']), step, _lib.ptr(step_dev), self.grad_max_val, _lib.ptr(coef), lo_all, hi_all,
                    _lib.stream()))
        return loss

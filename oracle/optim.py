"""Oracle restatement of the stroke-parameter update (test infrastructure only).

Follows ``source/utils/training.py:67-77`` (clip_gradients: global-norm and
value clipping, then ``nan_to_num_``), ``training.py:119-136`` (Adam with the
config's betas / eps) and ``source/models.py:576-584`` (clip_params).  Adam is
restated with torch's own update formulas and operation order
(``m.lerp_(g, 1 - b1)``, ``v.mul_(b2).addcmul_(g, g, value=1 - b2)``,
``denom = sqrt(v) / sqrt(1 - b2^t) + eps``, ``p.addcdiv_(m, denom, value=-lr / (1 - b1^t))``).

Pinned by ``tests/test_oracle_optim.py`` against ``torch.optim.Adam`` and
``torch.nn.utils.clip_grad_norm_`` / ``clip_grad_value_`` themselves.
"""
import math

import torch


def clip_gradients(grads, max_norm=0.0, max_val=0.0):
    """training.py:67-77 on a list of gradient tensors (returns new tensors)."""
    grads = [g.clone() for g in grads]
    if max_norm > 0:
        norms = torch.stack([torch.linalg.vector_norm(g, 2) for g in grads])
        total = torch.linalg.vector_norm(norms, 2)
        coef = torch.clamp(max_norm / (total + 1e-6), max=1.0)
        grads = [g * coef for g in grads]
    if max_val > 0:
        grads = [g.clamp(-max_val, max_val) for g in grads]
    return [torch.nan_to_num(g) for g in grads]


def adam_step(p, g, m, v, step, lr, betas=(0.9, 0.99), eps=1e-15):
    """One torch.optim.Adam update (step counted from 1); returns new (p, m, v)."""
    b1, b2 = betas
    m = m.lerp(g, 1 - b1)
    v = v.mul(b2).addcmul(g, g, value=1 - b2)
    bc1 = 1 - b1 ** step
    bc2 = 1 - b2 ** step
    denom = (v.sqrt() / math.sqrt(bc2)).add(eps)
    p = p.addcdiv(m, denom, value=-lr / bc1)
    return p, m, v


def clip_params(p, ranges=None, min_value=None, max_value=None):
    """models.py:576-584: per-column clamp (None = open side), then bounds for all entries."""
    p = p.clone()
    if ranges:
        for i, (lo, hi) in enumerate(ranges):
            if lo is not None or hi is not None:
                p[..., i] = p[..., i].clamp(lo, hi)
    if min_value is not None or max_value is not None:
        p = p.clamp(min_value, max_value)
    return p

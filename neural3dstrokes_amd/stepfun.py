"""Proposal-level interval resampling: drop-in for ``source/utils/stepfun.py:sample_intervals``.

``sample_intervals(rand, t, w_logits, num_samples, single_jitter, domain)`` keeps the
reference's arguments, random-number use and results (``stepfun.py:216-317``: the
``linspace`` of sample positions and the ``torch.rand`` jitter are drawn here with the
same calls, in the same order); the softmax, CDF, inverse CDF and interval
construction run in one gfx950 kernel, ``n3ds_sample_intervals``, one wavefront per
ray.  ``max_dilate_weights`` (``stepfun.py:120-146``, applied at ``models.py:159-166``)
runs as ``n3ds_max_dilate_weights``.  As with the reference's default ``stop_level_grad=True`` (``models.py:183-184``)
the result carries no gradient.  ``resample_logits`` is the logit expression of
``models.py:168-171``.
"""
import torch

from . import _lib


def resample_logits(sdist, weights, resample_padding=0.0):
    """models.py:168-171: log(w + padding) on bins of positive width, -inf elsewhere."""
    return torch.where(sdist[..., 1:] > sdist[..., :-1], torch.log(weights + resample_padding),
                       torch.full_like(sdist[..., :-1], -torch.inf))


def sample_intervals(rand, t, w_logits, num_samples, single_jitter=False, domain=(-torch.inf, torch.inf)):
    """stepfun.py:280-317: [..., num_samples + 1] interval fenceposts drawn from the step
    function (t [..., m + 1] sorted, w_logits [..., m])."""
    if num_samples <= 1:
        raise ValueError(f'num_samples must be > 1, is {num_samples}.')
    if t.shape[:-1] != w_logits.shape[:-1] or t.shape[-1] != w_logits.shape[-1] + 1:
        raise RuntimeError(f't {tuple(t.shape)} and w_logits {tuple(w_logits.shape)} do not describe one step '
                           'function per ray')
    device = t.device
    eps = torch.finfo(t.dtype).eps
    # stepfun.py:234-250 with deterministic_center=True
    if not rand:
        pad = 1 / (2 * num_samples)
        u = torch.linspace(pad, 1. - pad - eps, num_samples, device=device)
        jitter, max_jitter = None, 0.0
    else:
        u_max = eps + (1 - eps) / num_samples
        max_jitter = (1 - u_max) / (num_samples - 1) - eps
        d = 1 if single_jitter else num_samples
        u = torch.linspace(0, 1 - u_max, num_samples, device=device)
        jitter = torch.rand(t.shape[:-1] + (d,), device=device)
    pre = t.shape[:-1]
    m = t.shape[-1] - 1
    tf = t.detach().reshape(-1, m + 1).float().contiguous()
    lf = w_logits.detach().reshape(-1, m).float().contiguous()
    jf = None if jitter is None else jitter.reshape(tf.shape[0], jitter.shape[-1]).float().contiguous()
    u = u.float().contiguous()
    for name, x in (('t', tf), ('w_logits', lf), ('jitter', jf)):
        _lib.require_cuda(name, x)
    out = torch.empty(tf.shape[0], num_samples + 1, dtype=torch.float32, device=device)
    minval, maxval = (float(v) for v in domain)
    _lib.check(_lib.lib().n3ds_sample_intervals(
        _lib.ptr(out), _lib.ptr(tf), _lib.ptr(lf), _lib.ptr(u), _lib.ptr(jf), 0 if jf is None else jf.shape[1],
        float(max_jitter), tf.shape[0], m, num_samples, minval, maxval, _lib.stream()))
    return out.reshape(*pre, num_samples + 1).to(t.dtype)


def max_dilate_weights(t, w, dilation, domain=(-torch.inf, torch.inf), renormalize=False):
    """stepfun.py:136-146: (t_dilate [..., 3m + 1], w_dilate [..., 3m]) for t [..., m + 1]
    sorted and w [..., m] non-negative."""
    if t.shape[:-1] != w.shape[:-1] or t.shape[-1] != w.shape[-1] + 1:
        raise RuntimeError(f't {tuple(t.shape)} and w {tuple(w.shape)} do not describe one step function per ray')
    pre, m = t.shape[:-1], w.shape[-1]
    tf = t.detach().reshape(-1, m + 1).float().contiguous()
    wf = w.detach().reshape(-1, m).float().contiguous()
    for name, x in (('t', tf), ('w', wf)):
        _lib.require_cuda(name, x)
    t_out = torch.empty(tf.shape[0], 3 * m + 1, dtype=torch.float32, device=t.device)
    w_out = torch.empty(tf.shape[0], 3 * m, dtype=torch.float32, device=t.device)
    minval, maxval = (float(v) for v in domain)
    _lib.check(_lib.lib().n3ds_max_dilate_weights(
        _lib.ptr(t_out), _lib.ptr(w_out), _lib.ptr(tf), _lib.ptr(wf), tf.shape[0], m, float(dilation), minval, maxval,
        int(bool(renormalize)), _lib.stream()))
    return t_out.reshape(*pre, 3 * m + 1).to(t.dtype), w_out.reshape(*pre, 3 * m).to(w.dtype)

"""Oracle restatement of the proposal-level interval sampler (test infrastructure only).

Follows ``source/utils/stepfun.py``: ``sorted_interp`` (:5-24, the masked max / min
interval search, restated literally), ``max_dilate_weights`` (:109-146, sort and
masked max restated literally), ``integrate_weights`` (:149-170),
``invert_cdf`` (:195-202) and ``sample_intervals`` (:280-317) with the sample
positions ``u`` given explicitly (the reference draws them in ``sample``,
:216-254; ``sample_u`` restates that formula for given jitter draws).  numpy,
float64.  Pinned by ``tests/test_oracle_stepfun.py`` against ``numpy.interp``
(which the reference's own ``invert_cdf_np`` uses, :205-213) and closed forms.
"""
import numpy as np


def softmax(x):
    """torch.softmax semantics: a row of all -inf gives NaN."""
    m = np.max(x, axis=-1, keepdims=True)
    with np.errstate(invalid='ignore'):
        e = np.exp(x - m)
        return e / e.sum(-1, keepdims=True)


def integrate_weights(w):
    """stepfun.py:149-170: [0, min(cumsum(w[:-1]), 1), 1]."""
    cw = np.minimum(np.cumsum(w[..., :-1], axis=-1), 1.0)
    shape = cw.shape[:-1] + (1,)
    return np.concatenate([np.zeros(shape), cw, np.ones(shape)], -1)


def sorted_interp(x, xp, fp):
    """stepfun.py:5-24, literally: x [..., n], xp / fp [..., m+1]."""
    mask = x[..., None, :] >= xp[..., :, None]  # [..., m+1, n]

    def find_interval(v):
        v0 = np.max(np.where(mask, v[..., None], v[..., :1, None]), -2)
        v1 = np.min(np.where(~mask, v[..., None], v[..., -1:, None]), -2)
        return v0, v1

    fp0, fp1 = find_interval(fp)
    xp0, xp1 = find_interval(xp)
    with np.errstate(divide='ignore', invalid='ignore'):
        offset = np.clip(np.nan_to_num((x - xp0) / (xp1 - xp0), nan=0.0), 0, 1)
    return fp0 + offset * (fp1 - fp0)


def invert_cdf(u, t, w_logits):
    """stepfun.py:195-202."""
    return sorted_interp(u, integrate_weights(softmax(w_logits)), t)


def sample_u(num_samples, rand_draws=None, eps=float(np.finfo(np.float32).eps)):
    """stepfun.py:234-250 with deterministic_center=True: u [..., n] for jitter draws
    [..., 1 or n] in [0, 1) (None: no jitter)."""
    if rand_draws is None:
        pad = 1 / (2 * num_samples)
        return np.linspace(pad, 1. - pad - eps, num_samples)
    u_max = eps + (1 - eps) / num_samples
    max_jitter = (1 - u_max) / (num_samples - 1) - eps
    return np.linspace(0, 1 - u_max, num_samples) + rand_draws * max_jitter


def sample_intervals(u, t, w_logits, domain=(-np.inf, np.inf)):
    """stepfun.py:280-317 for given sample positions u [..., n] (n > 1)."""
    u = np.broadcast_to(u, t.shape[:-1] + (u.shape[-1],))
    centers = invert_cdf(u, t, w_logits)
    mid = (centers[..., 1:] + centers[..., :-1]) / 2
    first = np.maximum(2 * centers[..., :1] - mid[..., :1], domain[0])
    last = np.minimum(2 * centers[..., -1:] - mid[..., -1:], domain[1])
    return np.concatenate([first, mid, last], -1)


F32_EPS = float(np.finfo(np.float32).eps)


def max_dilate(t, p, dilation, domain=(-np.inf, np.inf), edge_dtype=None):
    """stepfun.py:120-133.  edge_dtype=np.float32 rounds the dilated edges t -+ d as the
    reference's float32 tensors do (the sort keys; everything else stays float64)."""
    t0 = t[..., :-1] - dilation
    t1 = t[..., 1:] + dilation
    if edge_dtype is not None:
        t0 = t0.astype(edge_dtype).astype(np.float64)
        t1 = t1.astype(edge_dtype).astype(np.float64)
    td = np.clip(np.sort(np.concatenate([t, t0, t1], -1), -1), *domain)
    cover = (t0[..., None, :] <= td[..., None]) & (t1[..., None, :] > td[..., None])
    pd = np.max(np.where(cover, p[..., None, :], 0.0), -1)[..., :-1]
    return td, pd


def max_dilate_weights(t, w, dilation, domain=(-np.inf, np.inf), renormalize=False, eps=F32_EPS, edge_dtype=None):
    """stepfun.py:136-146 with weight_to_pdf / pdf_to_weight (:109-117)."""
    p = w / np.maximum(t[..., 1:] - t[..., :-1], eps)
    td, pd = max_dilate(t, p, dilation, domain, edge_dtype)
    wd = pd * (td[..., 1:] - td[..., :-1])
    if renormalize:
        wd = wd / np.maximum(wd.sum(-1, keepdims=True), eps)
    return td, wd

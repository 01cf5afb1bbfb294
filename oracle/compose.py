"""Oracle restatement of stroke compositing and the ray march (test infrastructure only).

Follows ``source/strokelib/src/compositing.cu:9-157`` (ordered "over"
compositing and its hand-written adjoint), ``source/strokelib/strokes.py:388-417``
(non-"over" modes) and ``source/utils/render.py:5-116`` (cast_rays,
compute_alpha_weights, volumetric_rendering).

Pinned by ``tests/test_oracle_compose.py``: a single stroke returns its own
colour and ``D = d·α``; all-zero α gives ``D = 0`` and ``C = 0``; the sequential
over-operator equals the closed form with a reversed exclusive ``cumprod``; the
ray march equals ``1 - exp`` / ``cumsum`` closed forms and ``Σw ≤ 1``; the
reference adjoint's density part equals the autograd derivative while its
colour part differs (code-evident, only equal for one stroke).  The eval-time
extras (``volumetric_extras``) are pinned by ``numpy.interp`` on strictly
increasing integrated weights, closed forms for one opaque sample, an empty
ray and uniform weights (``tests/test_oracle_extras.py``).
"""
import torch

# The reference evaluates ``1.0f + 1e-6f - T`` in float32, so the constant it
# actually uses is float32(1 + 1e-6) = 1.00000095367431640625 (compositing.cu:49,106).
_ONE_PLUS_EPS = 1.00000095367431640625


def compose_forward(alphas, colors, density_params, skip_zero=True):
    """compose_forward_kernel, compositing.cu:18-58.

    alphas [N,S], colors [N,S,D], density_params [S].  Strokes are visited from
    index S-1 down to 0 (higher index in front).  Returns (density [N],
    color [N,D], raw color sum [N,D], final transmittance [N]).  ``skip_zero``
    reproduces the kernel's skip of alpha == 0 (identical values; the exact
    adjoint differentiates the unskipped form, whose derivative at alpha = 0 is
    the one-sided limit).
    """
    N, S = alphas.shape
    T = torch.ones(N, dtype=alphas.dtype, device=alphas.device)
    dens = torch.zeros(N, dtype=alphas.dtype, device=alphas.device)
    col = torch.zeros(N, colors.shape[-1], dtype=alphas.dtype, device=alphas.device)
    for k in range(S - 1, -1, -1):
        a = alphas[:, k]
        nz = (a != 0.0) if skip_zero else torch.ones_like(a, dtype=torch.bool)
        w = a * T
        T = torch.where(nz, T * (1.0 - a), T)
        dens = torch.where(nz, dens + density_params[k] * w, dens)
        col = torch.where(nz[:, None], col + colors[:, k, :] * w[:, None], col)
    scale = 1.0 / (_ONE_PLUS_EPS - T)
    out = torch.clamp(col * scale[:, None], 0.0, 1.0)
    return dens, out, col, T


def compose_closed_form(alphas, colors, density_params):
    """Same quantity via exclusive cumprod over the reversed stroke axis (a textbook reduction)."""
    a = torch.flip(alphas, [-1])
    c = torch.flip(colors, [-2])
    d = torch.flip(density_params, [-1])
    T_excl = torch.cumprod(torch.cat([torch.ones_like(a[:, :1]), 1.0 - a[:, :-1]], -1), -1)
    w = a * T_excl
    dens = (w * d).sum(-1)
    col = (w[..., None] * c).sum(-2)
    Tn = torch.prod(1.0 - a, -1)
    out = torch.clamp(col / (_ONE_PLUS_EPS - Tn)[:, None], 0.0, 1.0)
    return dens, out


def compose_backward_reference(grad_density, grad_color, alphas, colors, density_params):
    """Restatement of compose_backward_kernel, compositing.cu:73-156 (colour→α term as written)."""
    N, S = alphas.shape
    Dc = colors.shape[-1]
    dens, _, col, Tn = compose_forward(alphas, colors, density_params)
    if grad_density is None or grad_density.numel() == 0:
        grad_density = torch.zeros(N, dtype=alphas.dtype, device=alphas.device)
    if grad_color is None or grad_color.numel() == 0:
        grad_color = torch.zeros(N, Dc, dtype=alphas.dtype, device=alphas.device)
    O = _ONE_PLUS_EPS - Tn
    scale = 1.0 / O
    scaled = col * scale[:, None]
    in_range = (scaled >= 0.0) & (scaled <= 1.0)
    gC = torch.where(in_range, grad_color, torch.zeros_like(grad_color))
    T = torch.ones(N, dtype=alphas.dtype, device=alphas.device)
    d2 = torch.zeros(N, dtype=alphas.dtype, device=alphas.device)
    c2 = torch.zeros(N, Dc, dtype=alphas.dtype, device=alphas.device)
    g_alpha = torch.zeros_like(alphas)
    g_colors = torch.zeros_like(colors)
    g_dp = torch.zeros_like(density_params)
    for k in range(S - 1, -1, -1):
        a = alphas[:, k]
        pos = a > 0.0
        w = a * T
        T = torch.where(pos, T * (1.0 - a), T)
        d2 = torch.where(pos, d2 + density_params[k] * w, d2)
        c2 = torch.where(pos[:, None], c2 + colors[:, k, :] * w[:, None], c2)
        g_dp[k] = torch.where(pos, grad_density * w, torch.zeros_like(w)).sum()
        g_colors[:, k, :] = torch.where(pos[:, None], gC * scale[:, None] * w[:, None], torch.zeros_like(gC))
        sc = 1.0 / torch.clamp_min(1.0 - a, 1e-6)
        ga = grad_density * (T * density_params[k] - (dens - d2)) * sc
        sc2 = sc * scale * scale
        for i in range(Dc):
            ga = ga + gC[:, i] * (T * colors[:, k, i] * (O - a) - (col[:, i] - c2[:, i])) * sc2
        g_alpha[:, k] = ga
    return g_alpha, g_colors, g_dp


def compose_backward_exact(grad_density, grad_color, alphas, colors, density_params):
    """True derivative of :func:`compose_forward` by torch.autograd (use float64)."""
    with torch.enable_grad():
        a = alphas.detach().clone().requires_grad_(True)
        c = colors.detach().clone().requires_grad_(True)
        d = density_params.detach().clone().requires_grad_(True)
        dens, out, _, _ = compose_forward(a, c, d, skip_zero=False)
        loss = 0.0
        if grad_density is not None and grad_density.numel() > 0:
            loss = loss + (dens * grad_density).sum()
        if grad_color is not None and grad_color.numel() > 0:
            loss = loss + (out * grad_color).sum()
        g = torch.autograd.grad(loss, [a, c, d], allow_unused=True)
    return tuple(gi if gi is not None else torch.zeros_like(x) for gi, x in zip(g, [a, c, d]))


def compose_other(alphas, colors, density_params, composition_type):
    """Non-"over" composition modes, strokes.py:390-417."""
    if composition_type == 'max':
        idx = torch.argmax(alphas, dim=-1, keepdim=True)
        a = torch.take_along_dim(alphas, idx, dim=-1).squeeze(-1)
        return a * density_params[idx].squeeze(-1), torch.take_along_dim(colors, idx[..., None], dim=-2).squeeze(-2)
    if composition_type == 'max_density_weighted':
        idx = torch.argmax(alphas * density_params, dim=-1, keepdim=True)
        a = torch.take_along_dim(alphas, idx, dim=-1).squeeze(-1)
        return a * density_params[idx].squeeze(-1), torch.take_along_dim(colors, idx[..., None], dim=-2).squeeze(-2)
    if composition_type in ('softmax', 'softmax_density_weighted'):
        logits = alphas if composition_type == 'softmax' else alphas * density_params
        w = torch.softmax(logits * (1.0 / 0.05), dim=-1)
        return (w * alphas * density_params).sum(-1), (w[..., None] * colors).sum(-2)
    raise ValueError(composition_type)


# ----------------------------------------------------------------------------
# ray march, render.py:5-116
# ----------------------------------------------------------------------------

def cast_rays(tdist, origins, directions, radii):
    """render.py:5-26."""
    t0 = tdist[..., :-1]
    t1 = tdist[..., 1:]
    t = (t0 + t1) / 2
    coords = directions[..., None, :] * t[..., None] + origins[..., None, :]
    radius = radii * t
    return coords, radius, t


def compute_alpha_weights(density, tdist, dirs, opaque_background=False):
    """render.py:29-48."""
    t_delta = tdist[..., 1:] - tdist[..., :-1]
    delta = t_delta * torch.norm(dirs[..., None, :], dim=-1)
    density_delta = density * delta
    if opaque_background:
        density_delta = torch.cat([density_delta[..., :-1], torch.full_like(density_delta[..., -1:], torch.inf)], -1)
    alpha = 1 - torch.exp(-density_delta)
    trans = torch.exp(-torch.cat([torch.zeros_like(density_delta[..., :1]),
                                  torch.cumsum(density_delta[..., :-1], dim=-1)], dim=-1))
    return alpha * trans, alpha, trans


def volumetric_rendering(rgbs, weights, tdist, bg_rgbs):
    """render.py:73-88 (rgb, depth, acc)."""
    eps = torch.finfo(rgbs.dtype).eps
    acc = weights.sum(dim=-1)
    bg_w = (1 - acc[..., None]).clamp_min(0.)
    rgb = (weights[..., None] * rgbs).sum(dim=-2) + bg_w * bg_rgbs
    t_mids = 0.5 * (tdist[..., :-1] + tdist[..., 1:])
    depth = torch.clip(torch.nan_to_num((weights * t_mids).sum(dim=-1) / acc.clamp_min(eps), torch.inf),
                       tdist[..., 0], tdist[..., -1])
    return rgb, acc, depth


def sorted_interp(x, xp, fp):
    """stepfun.py:5-24.  For each query x (sorted xp): the fenceposts with xp <= x and the rest;
    (x0, f0) = the largest xp and fp of the first set, (x1, f1) = the smallest of the second
    (the first / last entry when a set is empty), then f0 + clip((x - x0) / (x1 - x0)) (f1 - f0)
    with 0 / 0 -> 0."""
    below = xp[..., :, None] <= x[..., None, :]                     # [..., m, k]
    first = lambda v: v[..., :1, None].expand(below.shape)          # noqa: E731
    last = lambda v: v[..., -1:, None].expand(below.shape)          # noqa: E731
    xv = xp[..., :, None].expand(below.shape)
    fv = fp[..., :, None].expand(below.shape)
    x0 = torch.where(below, xv, first(xp)).amax(-2)
    f0 = torch.where(below, fv, first(fp)).amax(-2)
    x1 = torch.where(below, last(xp), xv).amin(-2)
    f1 = torch.where(below, last(fp), fv).amin(-2)
    frac = torch.nan_to_num((x - x0) / (x1 - x0), 0.0).clamp(0.0, 1.0)
    return f0 + frac * (f1 - f0)


def weighted_percentile(t, w, ps):
    """stepfun.py:351-361 over integrate_weights (stepfun.py:149-170): [..., len(ps)]."""
    cw = torch.cumsum(w[..., :-1], dim=-1).clamp_max(1)
    pad = cw.shape[:-1] + (1,)
    cw0 = torch.cat([torch.zeros(pad, dtype=w.dtype, device=w.device), cw,
                     torch.ones(pad, dtype=w.dtype, device=w.device)], -1)
    x = (torch.tensor(ps, dtype=w.dtype, device=w.device) / 100).expand(cw0.shape[:-1] + (len(ps),))
    return sorted_interp(x.contiguous(), cw0, t)


def volumetric_extras(weights, tdist, t_far, eps=None):
    """render.py:98-116 (compute_extras): distance_mean and the 5 / 50 / 95 percentiles of the
    fenceposts extended by t_far [R] with the background weight.  eps: the clamp of acc (the
    reference's is float32's epsilon)."""
    eps = torch.finfo(torch.float32).eps if eps is None else eps
    acc = weights.sum(-1)
    t_mids = 0.5 * (tdist[..., :-1] + tdist[..., 1:])
    mean_log = (weights * torch.log(t_mids)).sum(-1) / acc.clamp_min(eps)
    dmean = torch.clip(torch.nan_to_num(torch.exp(mean_log), torch.inf), tdist[..., 0], tdist[..., -1])
    bg_w = (1 - acc[..., None]).clamp_min(0.)
    t_aug = torch.cat([tdist, t_far.reshape(-1, 1).to(tdist)], -1)
    w_aug = torch.cat([weights, bg_w], -1)
    pc = weighted_percentile(t_aug, w_aug, [5, 50, 95])
    return {'distance_mean': dmean, 'distance_percentile_5': pc[..., 0], 'distance_median': pc[..., 1],
            'distance_percentile_95': pc[..., 2]}

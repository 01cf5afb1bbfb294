"""Seeded inputs for the parity tests (points placed around strokes so many pairs have 0 < alpha < 1)."""
import math

import torch

from neural3dstrokes_amd.strokelib import strokes as dstrokes


def stroke_params(shape_type, color_type, n_strokes, seed):
    """Shape / colour params drawn from the drop-in samplers (the reference's initialisers)."""
    torch.manual_seed(seed)
    _, dim_s, dim_c, _, _, ssamp, csamp = dstrokes.get_stroke(shape_type, color_type, 'recon')
    sp = torch.stack([ssamp(100) for _ in range(n_strokes)]).float()
    cp = torch.stack([csamp() for _ in range(n_strokes)]).float()
    assert sp.shape[1] == dim_s and cp.shape[1] == dim_c
    return sp, cp


def points_near(shape_type, sp, n_rays, n_samples, seed, spread=1.6):
    """Points around random strokes: unit-space U[-spread, spread]^3 mapped to world space."""
    from oracle import strokes as ostrokes
    g = torch.Generator().manual_seed(seed + 1)
    sdf_id = ostrokes.shape_sdf_id(shape_type)
    base = sdf_id >> 4
    N = n_rays * n_samples
    k = torch.randint(0, sp.shape[0], (N,), generator=g)
    if base >= 9:
        ctrl = sp[:, :9 if base == 9 else 12].reshape(sp.shape[0], -1, 3).double()
        t = torch.rand(N, generator=g, dtype=torch.float64)
        j = torch.randint(0, ctrl.shape[1], (N,), generator=g)
        x = ctrl[k, j] * (1 - t)[:, None] + ctrl[k, (j + 1) % ctrl.shape[1]] * t[:, None]
        x = x + (torch.rand(N, 3, generator=g, dtype=torch.float64) - 0.5) * 0.3
    else:
        row = sp[k].double()
        _, scale, angle, trans = ostrokes.split_shape_params(row, sdf_id)
        p = (torch.rand(N, 3, generator=g, dtype=torch.float64) * 2 - 1) * spread
        if scale is not None:
            p = p * scale
        if angle is not None:
            p = torch.einsum('nij,nj->ni', ostrokes.rotation_matrix(angle), p)
        if trans is not None:
            p = p + trans
        x = p
    x = x.float()
    vd = torch.randn(n_rays, 3, generator=g)
    vd = vd / vd.norm(dim=-1, keepdim=True)
    return x.reshape(n_rays, n_samples, 3), vd


def conditioning_mask(shape_type, sp, x, r, vd, laplace, delta=1.0):
    """[N, S] bool: pairs whose SDF gradient is well conditioned in float32.

    A pair is excluded when its float64 dSDF/dx changes by more than 1e-3 under
    a 2e-5 displacement of the point along any axis (a kink of the
    curve-parameter selection, a projection clamp or a face switch lies within
    that distance, far above float32 rounding), or, for the linear transfer,
    when the unclamped alpha lies within 1e-4 of a clamp boundary.  Any float32
    implementation of the reference math (including the reference's own
    kernels) can land on either side of such a kink and then disagrees with
    float64 by O(1) on that pair.
    """
    from oracle import strokes as ostrokes
    sdf_id = ostrokes.shape_sdf_id(shape_type)
    X = x.reshape(-1, 3).double()
    S = sp.shape[0]
    base = sdf_id >> 4
    spd = sp[None].double()

    def pair_grads(xv, prm):
        xx = xv.clone().requires_grad_(True)
        p = ostrokes.inverse_transform(xx[:, None, :], prm, sdf_id)
        s = ostrokes.base_sdf(base, [c.expand(X.shape[0], S) for c in p], prm[..., :ostrokes.NUM_BASE_PARAMS[base]])
        return torch.stack([torch.autograd.grad(s[:, k].sum(), xx, retain_graph=True)[0] for k in range(S)], 1)

    g0 = pair_grads(X, spd)
    bad = torch.zeros(X.shape[0], S, dtype=torch.bool)
    eps = 2e-5
    trials = []
    for i in range(3):
        for sgn in (-1.0, 1.0):
            dx = torch.zeros(3, dtype=torch.float64)
            dx[i] = sgn * eps
            trials.append((X + dx, spd))
    gen = torch.Generator().manual_seed(1234)
    for _ in range(3):  # relative 1e-6 parameter noise: curve geometry rounding amplified by short segments
        trials.append((X, spd * (1 + 1e-6 * torch.randn(spd.shape, generator=gen, dtype=torch.float64))))
    for xv, prm in trials:
        gi = pair_grads(xv, prm)
        bad |= (gi - g0).norm(dim=-1) > 1e-3 * g0.norm(dim=-1).clamp_min(1.0)
    if base >= 9:  # near-ties between curve-parameter candidates (second-order flat kinks)
        p = ostrokes.inverse_transform(X[:, None, :], spd, sdf_id)
        gap = ostrokes.curve_selection_margin(base, [c.expand(X.shape[0], S) for c in p], spd)
        bad |= gap < 1e-5
    if not laplace:
        fw = ostrokes.stroke_forward(X, r.reshape(-1).double(), vd.double(), sp.double(),
                                     torch.zeros(S, 3, dtype=torch.float64), sdf_id, 0, delta, False, False)
        u = -fw['sdf'] * fw['sdf_scale'] + 0.5  # alpha before the clamp
        bad |= (u.abs() < 1e-4) | ((u - 0.9999).abs() < 1e-4)
    return ~bad


def rand_like_seeded(shape, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(*shape, generator=g)


def unit_quat_euler(n, g):
    u1, u2, u3 = (torch.rand(n, generator=g, dtype=torch.float64) for _ in range(3))
    q = torch.stack([torch.sqrt(1 - u1) * torch.sin(2 * math.pi * u2), torch.sqrt(1 - u1) * torch.cos(2 * math.pi * u2),
                     torch.sqrt(u1) * torch.sin(2 * math.pi * u3), torch.sqrt(u1) * torch.cos(2 * math.pi * u3)], -1)
    return q

"""Oracle for the whole stroke-field hot path with a mixed stroke palette (test infrastructure only).

Reference semantics (``source/models.py:660-711`` plus
``source/utils/render.py:5-116``): the reference kernels evaluate one stroke
type per call (``sdf_id`` is a scalar, ``strokes.py:118``), so a mixed palette
is one ``stroke_forward`` per type with the outputs placed into columns by
stroke index (the draw order), followed by one ``compose_forward`` over all
strokes and the torch ray march.  Exact gradients come from torch.autograd of
this composition in float64.

Pinned by ``tests/test_oracle_field.py``: equality with the per-type
``stroke_forward`` + ``compose_forward`` + ray-march chain, invariance of the
result to inserting strokes that cannot touch any sample, and central finite
differences of the float64 composition on random parameters.
"""
import torch

from . import strokes as ostrokes
from . import compose as ocompose


def stroke_columns(x, radius, viewdir, stroke_ids, shape_params, color_params, color_id, sdf_delta,
                   laplace=False, inv_scale_radius=False, exact=True):
    """alpha [N,S], color [N,S,3] for strokes of mixed types (shape_params padded to Pmax)."""
    N = x.shape[0]
    S = len(stroke_ids)
    alpha_cols = [None] * S
    color_cols = [None] * S
    ids = [int(i) for i in stroke_ids]
    for sdf_id in sorted(set(ids)):
        cols = [k for k in range(S) if ids[k] == sdf_id]
        P = ostrokes.num_shape_params(sdf_id)
        idx = torch.tensor(cols, device=x.device)
        sp = shape_params[idx][:, :P]
        cp = color_params[idx]
        fw = ostrokes.stroke_forward(x, radius, viewdir, sp, cp, sdf_id, color_id, sdf_delta, laplace,
                                     inv_scale_radius, exact_radius_grad=exact)
        for j, k in enumerate(cols):
            alpha_cols[k] = fw['alpha'][:, j]
            color_cols[k] = fw['color'][:, j, :]
    return torch.stack(alpha_cols, -1), torch.stack(color_cols, -2)


def _bg(bg_color, like):
    """Background colour(s): a 3-tuple / [3] tensor, or per-ray [R,3] (models.py:213-222)."""
    if torch.is_tensor(bg_color):
        return bg_color.to(dtype=like.dtype, device=like.device)
    return torch.tensor(bg_color, dtype=like.dtype, device=like.device)


def render_field(origins, directions, viewdirs, radii, tdist, stroke_ids, shape_params, color_params,
                 density_params, color_id=0, sdf_delta=1.0, laplace=False, inv_scale_radius=False,
                 bg_color=(1.0, 1.0, 1.0), cone=False, coord_div=1.0, exact=True, opaque_background=False):
    """Forward of the fused op: returns dict(rgb [R,3], acc [R], depth [R], density [R,n], color [R,n,3]).

    ``bg_color``: one colour or per-ray colours [R,3]; ``opaque_background``
    makes the last interval infinitely wide (render.py:35-40).
    """
    R, n1 = tdist.shape
    n = n1 - 1
    coords, rad_cone, t = ocompose.cast_rays(tdist, origins, directions, radii[:, None])
    radius = rad_cone if cone else radii[:, None].expand(R, n)
    x = (coords / coord_div).reshape(-1, 3)
    alphas, colors = stroke_columns(x, radius.reshape(-1), viewdirs, stroke_ids, shape_params, color_params,
                                    color_id, sdf_delta, laplace, inv_scale_radius, exact)
    dens, col, _, _ = ocompose.compose_forward(alphas, colors, density_params)
    dens = dens.reshape(R, n)
    col = col.reshape(R, n, 3)
    w, _, _ = ocompose.compute_alpha_weights(dens, tdist, directions, opaque_background)
    rgb, acc, depth = ocompose.volumetric_rendering(col, w, tdist, _bg(bg_color, tdist))
    return dict(rgb=rgb, acc=acc, depth=depth, density=dens, color=col, weights=w, alphas=alphas)


def render_field_grads_reference(grad_rgb, grad_acc, origins, directions, viewdirs, radii, tdist, stroke_ids,
                                 shape_params, color_params, density_params, color_id=0, sdf_delta=1.0, laplace=False,
                                 inv_scale_radius=False, bg_color=(1.0, 1.0, 1.0), cone=False, coord_div=1.0,
                                 opaque_background=False):
    """Gradients the reference chain produces: torch autograd through the ray march
    (render.py is torch), then the restated hand-written compositing adjoint
    (compositing.cu:61-157) and the per-type stroke adjoints (strokes_backward.cu)."""
    R, n1 = tdist.shape
    n = n1 - 1
    coords, rad_cone, t = ocompose.cast_rays(tdist, origins, directions, radii[:, None])
    radius = (rad_cone if cone else radii[:, None].expand(R, n)).reshape(-1)
    x = (coords / coord_div).reshape(-1, 3)
    alphas, colors = stroke_columns(x, radius, viewdirs, stroke_ids, shape_params, color_params, color_id, sdf_delta,
                                    laplace, inv_scale_radius, exact=False)
    with torch.enable_grad():
        dens, col, _, _ = ocompose.compose_forward(alphas, colors, density_params)
        dens = dens.detach().requires_grad_(True)
        col = col.detach().requires_grad_(True)
        w, _, _ = ocompose.compute_alpha_weights(dens.reshape(R, n), tdist, directions, opaque_background)
        rgb, acc, _ = ocompose.volumetric_rendering(col.reshape(R, n, 3), w, tdist, _bg(bg_color, tdist))
        loss = (rgb * grad_rgb).sum() + (acc * grad_acc).sum()
        g_dens, g_col = torch.autograd.grad(loss, [dens, col])
    g_alpha, g_colors, g_dp = ocompose.compose_backward_reference(g_dens, g_col, alphas, colors, density_params)
    gs = torch.zeros_like(shape_params)
    gc = torch.zeros_like(color_params)
    ids = [int(i) for i in stroke_ids]
    for sdf_id in sorted(set(ids)):
        cols = [k for k in range(len(ids)) if ids[k] == sdf_id]
        P = ostrokes.num_shape_params(sdf_id)
        idx = torch.tensor(cols, device=x.device)
        sp = shape_params[idx][:, :P]
        cp = color_params[idx]
        gsi, gci, _ = ostrokes.stroke_backward_reference(g_alpha[:, idx], g_colors[:, idx], None, x, radius, viewdirs,
                                                         alphas[:, idx], sp, cp, sdf_id, color_id, sdf_delta, laplace,
                                                         inv_scale_radius, need_grad_x=False)
        gs[idx, :P] = gsi
        gc[idx] = gci
    return gs, gc, g_dp


def render_field_grads(grad_rgb, grad_acc, origins, directions, viewdirs, radii, tdist, stroke_ids, shape_params,
                       color_params, density_params, grad_depth=None, **kw):
    """Exact parameter gradients of sum(rgb*grad_rgb) + sum(acc*grad_acc) (+ sum(depth*grad_depth)) by autograd."""
    with torch.enable_grad():
        sp = shape_params.detach().clone().requires_grad_(True)
        cp = color_params.detach().clone().requires_grad_(True)
        dp = density_params.detach().clone().requires_grad_(True)
        out = render_field(origins, directions, viewdirs, radii, tdist, stroke_ids, sp, cp, dp, **kw)
        loss = (out['rgb'] * grad_rgb).sum() + (out['acc'] * grad_acc).sum()
        if grad_depth is not None:
            loss = loss + (out['depth'] * grad_depth).sum()
        g = torch.autograd.grad(loss, [sp, cp, dp], allow_unused=True)
    g = [gi if gi is not None else torch.zeros_like(x) for gi, x in zip(g, [sp, cp, dp])]
    return out, g[0], g[1], g[2]


def render_field_grads_chunked(grad_rgb, grad_acc, origins, directions, viewdirs, radii, tdist, stroke_ids,
                               shape_params, color_params, density_params, chunk=64, grad_depth=None, **kw):
    """``render_field_grads`` over ray chunks (rays are independent, gradients add up).

    Returns (dict(rgb, acc, depth), grad_shape, grad_color, grad_density).
    """
    R = tdist.shape[0]
    outs = {'rgb': [], 'acc': [], 'depth': []}
    gs = torch.zeros_like(shape_params)
    gc = torch.zeros_like(color_params)
    gd = torch.zeros_like(density_params)
    bg = kw.pop('bg_color', (1.0, 1.0, 1.0))
    for r0 in range(0, R, chunk):
        sl = slice(r0, min(R, r0 + chunk))
        bgc = bg[sl] if torch.is_tensor(bg) and bg.dim() == 2 else bg
        o, a, b, c = render_field_grads(grad_rgb[sl], grad_acc[sl], origins[sl], directions[sl], viewdirs[sl],
                                        radii[sl], tdist[sl], stroke_ids, shape_params, color_params, density_params,
                                        grad_depth=None if grad_depth is None else grad_depth[sl], bg_color=bgc, **kw)
        for k in outs:
            outs[k].append(o[k].detach())
        gs += a
        gc += b
        gd += c
    return {k: torch.cat(v) for k, v in outs.items()}, gs, gc, gd

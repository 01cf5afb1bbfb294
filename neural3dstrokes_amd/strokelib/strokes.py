"""Drop-in stroke operators: the reference's ``source/strokelib/strokes.py`` surface.

``get_stroke``, ``compose_strokes``, ``_stroke_fn`` and ``_compositing_fn`` keep
the reference's names, argument meaning, shapes and error behaviour
(``strokes.py:109-419``); the math runs in the gfx950 kernels of
``libn3ds.so`` through :mod:`.backend`.

The type registry below has to match the reference's table entry for entry
(``strokes.py:12-106``): the integer ids and parameter layouts are the
interface contract of the kernels.

Adjoint selection (no counterpart in the reference): ``set_adjoint('exact')``
(default) differentiates the forward exactly; ``set_adjoint('reference')``
reproduces the reference's hand-written CUDA adjoints term for term.  The
environment variable ``N3DS_ADJOINT`` sets the initial value.
"""
import math
import os

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import backend as _backend

_ADJOINT = os.environ.get('N3DS_ADJOINT', 'exact')


def set_adjoint(mode):
    """Select 'exact' or 'reference' adjoints for the stroke and compositing operators."""
    global _ADJOINT
    if mode not in ('exact', 'reference'):
        raise ValueError(f"adjoint must be 'exact' or 'reference', got {mode!r}")
    _ADJOINT = mode


def get_adjoint():
    return _ADJOINT


_base_sdf_id = {
    'unit_sphere': 0,
    'unit_cube': 1,
    'unit_round_cube': 2,
    'unit_capped_torus': 3,
    'unit_capsule': 4,
    'unit_line': 5,
    'unit_triprism': 6,
    'unit_octahedron': 7,
    'unit_tetrahedron': 8,
    'quadratic_bezier': 9,
    'cubic_bezier': 10,
    'catmull_rom': 11,
}


def _spline_sampler(n_points, n_radius):
    """Control points around a centre with radii proportional to the sampled scale (strokes.py:27-38)."""
    def sample(args):
        scale_min, scale_max, centre = args
        scale = torch.rand(1) * (scale_max - scale_min) + scale_min
        pts = [centre + (torch.rand(3) - 0.5) * scale for _ in range(n_points)]
        radii = (0.2 + 0.2 * torch.rand(n_radius)) * scale
        return torch.cat(pts + [radii])
    return sample


_UNIT = (-1, 1)
_RADIUS = (0.001, 0.2)

# shape type -> (base sdf, base param ranges, base sampler, translation, rotation, single scale, multi scale)
_sdf_dict = {
    'sphere': ('unit_sphere', [], None, True, False, True, False),
    'ellipsoid': ('unit_sphere', [], None, True, True, False, True),
    'aacube': ('unit_cube', [], None, True, False, True, False),
    'cube': ('unit_cube', [], None, True, True, True, False),
    'aabb': ('unit_cube', [], None, True, False, False, True),
    'obb': ('unit_cube', [], None, True, True, False, True),
    'roundcube': ('unit_round_cube', [(0, 1)], lambda _: 0.8 * torch.rand(1), True, True, True, False),
    'roundbox': ('unit_round_cube', [(0, 1)], lambda _: 0.8 * torch.rand(1), True, True, False, True),
    'cappedtorus': ('unit_capped_torus', [(0, math.pi), (0, None)], lambda _: torch.rand(2),
                    True, True, True, False),
    'capsule': ('unit_capsule', [(0.25, None)], lambda _: torch.rand(1) + 0.25, True, True, True, False),
    'scapsule': ('unit_capsule', [(0.25, None)], lambda _: torch.rand(1) + 0.25, True, True, False, True),
    'line': ('unit_line', [(0.25, None), (-0.8, 0.8)],
             lambda _: torch.cat([torch.rand(1) + 0.25, torch.rand(1) - 0.5]), True, True, True, False),
    'triprism': ('unit_triprism', [(0, None)], lambda _: torch.rand(1), True, True, True, False),
    'octahedron': ('unit_octahedron', [], None, True, True, True, False),
    'tetrahedron': ('unit_tetrahedron', [], None, True, True, True, False),
    'quadratic_bezier': ('quadratic_bezier', [_UNIT] * 9 + [_RADIUS] * 2, _spline_sampler(3, 2),
                         False, False, False, False),
    'cubic_bezier': ('cubic_bezier', [_UNIT] * 12 + [_RADIUS] * 2, _spline_sampler(4, 2),
                     False, False, False, False),
    'catmull_rom': ('catmull_rom', [_UNIT] * 12 + [_RADIUS] * 2, _spline_sampler(4, 2),
                    False, False, False, False),
}

# colour type -> (colour id, param ranges, sampler)
_color_dict = {
    'constant_rgb': (0, [(0, 1)] * 3, lambda: torch.rand(3)),
    'gradient_rgb': (1, [(-1, 1)] * 6 + [(0, 1)] * 6, lambda: torch.cat([torch.rand(6) - 0.5, torch.rand(6)], dim=-1)),
    'constant_sh2': (2, [(None, None)] * 12, lambda: torch.randn(12)),
    'constant_sh3': (3, [(None, None)] * 27, lambda: torch.randn(27)),
}

_color_dim = [3, 3, 3, 3]


def _make_sdf_id(base_sdf_name, enable_translation, enable_rotation, enable_singlescale, enable_multiscale):
    """``base << 4`` with transform flag bits T=1, R=2, SS=4, MS=8 (strokes.py:95-106)."""
    flags = (1 if enable_translation else 0) | (2 if enable_rotation else 0) | \
        (4 if enable_singlescale else 0) | (8 if enable_multiscale else 0)
    return (_base_sdf_id[base_sdf_name] << 4) | flags


def _as_f32(t, shape):
    return t.contiguous().reshape(shape).float()


class _stroke_fn(Function):
    """Per-(point, stroke) alpha / colour (/ sdf / texcoord) of one stroke type (strokes.py:109-224).

    forward(x [..., S_pts, 3], radius [..., S_pts], viewdir [..., 3], shape_params [K, P],
    color_params [K, Q], sdf_id, color_id, sdf_delta, use_laplace_transform=False,
    inv_scale_radius=False, no_sdf=True, return_texcoord=False)
    -> (alpha [..., S_pts, K], color [..., S_pts, K, 3], sdf or None, texcoord or None)
    """

    @staticmethod
    @torch.amp.custom_fwd(device_type='cuda')
    def forward(ctx, x, radius, viewdir, shape_params, color_params, sdf_id, color_id, sdf_delta,
                use_laplace_transform=False, inv_scale_radius=False, no_sdf=True, return_texcoord=False):
        assert x.shape[-1] == 3, 'x must have shape [..., num_samples, 3]'
        assert viewdir.shape[:-1] == x.shape[:-2] and viewdir.shape[-1] == 3, 'viewdir must have shape [..., 3]'
        assert shape_params.ndim == 2, 'params must have shape [num_strokes, num_shape_params]'
        assert color_params.ndim == 2, 'color_params must have shape [num_strokes, num_color_params]'
        assert shape_params.shape[0] == color_params.shape[0], 'num_strokes must be the same'
        pre_shape = x.shape[:-1]
        x = _as_f32(x, (-1, 3))
        radius = _as_f32(radius, (-1,))
        viewdir = _as_f32(viewdir, (-1, 3))
        shape_params = shape_params.contiguous().float()
        color_params = color_params.contiguous().float()
        n_pts, n_str = x.shape[0], shape_params.shape[0]
        dev = x.device
        alpha = torch.empty((n_pts, n_str), dtype=torch.float32, device=dev)
        color = torch.empty((n_pts, n_str, _color_dim[color_id]), dtype=torch.float32, device=dev)
        sdf = torch.empty(0 if no_sdf else (n_pts, n_str), dtype=torch.float32, device=dev)
        tex = torch.empty((n_pts, n_str, 2) if return_texcoord else 0, dtype=torch.float32, device=dev)
        _backend.stroke_forward(alpha, color, sdf, tex, x, radius, viewdir, shape_params, color_params, sdf_id,
                                color_id, sdf_delta, use_laplace_transform, inv_scale_radius)
        if any(ctx.needs_input_grad[i] for i in (0, 3, 4)):
            ctx.save_for_backward(x, radius, viewdir, alpha, shape_params, color_params)
            ctx.cfg = (sdf_id, color_id, sdf_delta, use_laplace_transform, inv_scale_radius, pre_shape,
                       get_adjoint())
        texo = None
        if return_texcoord:
            # the reference's backward discards grad_texcoord (strokes.py:188-224): no gradient
            # flows through texture coordinates, so consumers need not compute one
            texo = tex.reshape(*pre_shape, n_str, 2)
            ctx.mark_non_differentiable(texo)
        return (alpha.reshape(*pre_shape, n_str), color.reshape(*pre_shape, n_str, -1),
                None if no_sdf else sdf.reshape(*pre_shape, n_str), texo)

    @staticmethod
    @once_differentiable
    @torch.amp.custom_bwd(device_type='cuda')
    def backward(ctx, grad_alpha, grad_color, grad_sdf=None, grad_texcoord=None):
        nones = (None,) * 12
        if not any(ctx.needs_input_grad[i] for i in (0, 3, 4)):
            return nones
        x, radius, viewdir, alpha, shape_params, color_params = ctx.saved_tensors
        sdf_id, color_id, sdf_delta, laplace, isr, pre_shape, adjoint = ctx.cfg
        n_str = shape_params.shape[0]
        dev = x.device
        grad_alpha = torch.zeros_like(alpha) if grad_alpha is None else _as_f32(grad_alpha, (-1, n_str))
        grad_color = torch.zeros((x.shape[0], n_str, _color_dim[color_id]), device=dev) if grad_color is None \
            else _as_f32(grad_color, (-1, n_str, _color_dim[color_id]))
        grad_sdf = torch.zeros(0, device=dev) if grad_sdf is None else _as_f32(grad_sdf, (-1, n_str))
        g_shape = torch.zeros_like(shape_params)
        g_color = torch.zeros_like(color_params)
        g_x = torch.zeros(x.shape if ctx.needs_input_grad[0] else 0, dtype=torch.float32, device=dev)
        _backend.stroke_backward(g_shape, g_color, g_x, grad_alpha, grad_color, grad_sdf, x, radius, viewdir, alpha,
                                 shape_params, color_params, sdf_id, color_id, sdf_delta, laplace, isr, adjoint)
        g_x = g_x.reshape(*pre_shape, 3) if ctx.needs_input_grad[0] else None
        return (g_x, None, None, g_shape, g_color) + (None,) * 7


def get_stroke(shape_type, color_type, init_type):
    """Stroke function and its parameter bookkeeping (strokes.py:227-313).

    Returns (stroke_fn, dim_shape, dim_color, shape_param_ranges, color_param_ranges,
    shape_sampler, color_sampler).
    """
    base_name, base_ranges, base_sampler, use_t, use_r, use_ss, use_ms = _sdf_dict[shape_type]
    color_id, color_ranges, color_sampler = _color_dict[color_type]
    sdf_id = _make_sdf_id(base_name, use_t, use_r, use_ss, use_ms)
    shape_ranges = list(base_ranges)
    if use_ss:
        shape_ranges += [(0.01, 0.5)]
    elif use_ms:
        shape_ranges += [(0.01, 0.5)] * 3
    if use_r:
        shape_ranges += [(-math.pi, math.pi)] * 3
    if use_t:
        shape_ranges += [(-1.0, 1.0)] * 3

    def shape_sampler(stroke_step, error_coord=None):
        decay = math.exp(-stroke_step / 200)
        if init_type == 'recon':
            lo, hi = torch.full((3,), -0.5), torch.full((3,), 0.5)
            extent = torch.sqrt(torch.square(hi - lo).sum())
            s_min, s_max = (0.02 + 0.12 * decay) * extent, (0.04 + 0.18 * decay) * extent
            centre = error_coord if error_coord is not None else lo + (hi - lo) * torch.rand(3)
        elif init_type == 'gen_box':
            lo, hi = torch.full((3,), -0.6), torch.full((3,), 0.6)
            extent = torch.sqrt(torch.square(hi - lo).sum())
            centre = lo + (hi - lo) * torch.rand(3)
            dist = 1 - 2. * torch.sqrt(torch.square(centre).sum()) / extent
            s_min, s_max = 0.05 * extent * (4 ** -dist), 0.10 * extent * (4 ** -dist)
        elif init_type == 'gen_sphere':
            s_min, s_max = 0.04 + 0.06 * decay, 0.06 + 0.09 * decay
            theta = torch.rand(1) * 2 * math.pi
            phi = torch.rand(1) * 2 * math.pi
            rad = 1.0 * (1 - decay)
            centre = torch.cat([rad * torch.sin(phi) * torch.cos(theta), rad * torch.sin(phi) * torch.sin(theta),
                                rad * torch.cos(phi)])
        else:
            raise KeyError(init_type)
        parts = []
        if base_sampler is not None:
            parts.append(base_sampler((s_min, s_max, centre)))
        if use_ss:
            parts.append(torch.rand(1) * (s_max - s_min) + s_min)
        elif use_ms:
            parts.append(torch.rand(3) * (s_max - s_min) + s_min)
        if use_r:
            parts.append((torch.rand(3) * 2 - 1) * math.pi)
        if use_t:
            parts.append(centre)
        return torch.cat(parts, dim=-1) if parts else torch.empty(0)

    def stroke_fn(x, radius, viewdir, shape_params, color_params, *args):
        return _stroke_fn.apply(x, radius, viewdir, shape_params, color_params, sdf_id, color_id, *args)

    return stroke_fn, len(shape_ranges), len(color_ranges), shape_ranges, list(color_ranges), shape_sampler, \
        color_sampler


class _compositing_fn(Function):
    """Ordered "over" compositing (strokes.py:316-373): (density [...], color [..., D])."""

    @staticmethod
    @torch.amp.custom_fwd(device_type='cuda')
    def forward(ctx, alphas, colors, density_params):
        assert alphas.ndim >= 2, 'alphas must have shape [..., num_strokes]'
        assert colors.ndim >= 3, 'colors must have shape [..., num_strokes, color_dim]'
        assert density_params.ndim == 1, 'density_params must have shape [num_strokes]'
        assert alphas.shape[:-1] == colors.shape[:-2], \
            'alphas and colors must have the same shape except the last two dimensions'
        assert alphas.shape[-1] == colors.shape[-2] == density_params.shape[0], \
            'alphas, colors and density_params must have the same number of strokes'
        pre_shape = alphas.shape[:-1]
        n_str = density_params.shape[0]
        alphas = _as_f32(alphas, (-1, n_str))
        colors = _as_f32(colors, (-1, n_str, colors.shape[-1]))
        density_params = density_params.contiguous().float()
        dens = torch.empty(alphas.shape[0], dtype=torch.float32, device=alphas.device)
        col = torch.empty((colors.shape[0], colors.shape[-1]), dtype=torch.float32, device=alphas.device)
        _backend.compose_forward(dens, col, alphas, colors, density_params)
        if any(ctx.needs_input_grad):
            ctx.save_for_backward(alphas, colors, density_params)
            ctx.cfg = (pre_shape, get_adjoint())
        return dens.reshape(*pre_shape), col.reshape(*pre_shape, -1)

    @staticmethod
    @once_differentiable
    @torch.amp.custom_bwd(device_type='cuda')
    def backward(ctx, grad_density, grad_color):
        alphas, colors, density_params = ctx.saved_tensors
        pre_shape, adjoint = ctx.cfg
        n_str = density_params.shape[0]
        dev = alphas.device
        gd = torch.zeros(0, device=dev) if grad_density is None else _as_f32(grad_density, (-1,))
        gc = torch.zeros(0, device=dev) if grad_color is None else _as_f32(grad_color, (-1, colors.shape[-1]))
        g_alpha = torch.zeros_like(alphas)
        g_colors = torch.zeros_like(colors)
        g_dp = torch.zeros_like(density_params)
        _backend.compose_backward(g_alpha, g_colors, g_dp, gd, gc, alphas, colors, density_params, adjoint)
        return g_alpha.reshape(*pre_shape, n_str), g_colors.reshape(*pre_shape, n_str, colors.shape[-1]), g_dp


def compose_strokes(alphas, colors, density_params, composition_type):
    """Composite strokes (strokes.py:376-419): 'over' runs the gfx950 kernels, the
    selection / softmax modes are differentiable tensor expressions as in the reference."""
    if composition_type == 'over':
        return _compositing_fn.apply(alphas, colors, density_params)
    if composition_type in ('max', 'max_density_weighted'):
        key = alphas if composition_type == 'max' else alphas * torch.broadcast_to(density_params, alphas.shape)
        idx = torch.argmax(key, dim=-1, keepdim=True)
        a = torch.take_along_dim(alphas, idx, dim=-1).squeeze(-1)
        return a * density_params[idx].squeeze(-1), torch.take_along_dim(colors, idx[..., None], dim=-2).squeeze(-2)
    if composition_type in ('softmax', 'softmax_density_weighted'):
        weighted = alphas * torch.broadcast_to(density_params, alphas.shape)
        key = alphas if composition_type == 'softmax' else weighted
        w = torch.softmax(key * (1.0 / 0.05), dim=-1)
        return (w * weighted).sum(-1), (w[..., None] * colors).sum(-2)
    assert 0, f'Unknown composition type {composition_type}'

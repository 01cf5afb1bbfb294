"""Fused stroke-field rendering operator (no single counterpart in the reference).

``render_strokes`` computes, for a batch of rays and a mixed-type stroke set,
exactly what the reference computes with the chain
``StrokeField.predict_density`` (``source/models.py:660-711``: one
``stroke_fn`` call per stroke type, outputs placed by stroke index,
``compose_strokes(..., 'over')``) followed by ``render.compute_alpha_weights``
and ``render.volumetric_rendering`` (``source/utils/render.py:29-88``):
per-ray ``rgb``, ``acc`` and ``depth``.  It runs as one gfx950 forward kernel
and one adjoint kernel (plus a tiny stroke-table preparation and a gradient
reduction); no per-(sample, stroke) tensor is materialised.

Stroke sets are passed as ``stroke_ids`` (the reference's ``sdf_id`` per
stroke, ``strokes.py:95-106``), ``shape_params`` padded to a common width,
``color_params`` of one colour type and ``density_params`` (already scaled by
``density_scale``).
"""
import weakref

import torch
from torch.autograd import Function

from . import _lib
from .scenes import n_shape_params
from .strokelib.strokes import get_adjoint

COLOR_WIDTH = [3, 12, 12, 27]
N_BASE = 12


def prepare_table(stroke_ids, shape_params):
    """Per-stroke record table consumed by the render kernels (size from the library)."""
    S, Pmax = shape_params.shape
    nt = int(_lib.lib().n3ds_render_table_floats(S))
    table = torch.empty(max(nt, 1), dtype=torch.float32, device=shape_params.device)
    status = torch.zeros(1, dtype=torch.int32, device=shape_params.device)
    _lib.check(_lib.lib().n3ds_render_prepare(_lib.ptr(table), _lib.ptr(stroke_ids), _lib.ptr(shape_params), S, Pmax,
                                              _lib.ptr(status), _lib.stream()))
    return table, status


# Stroke-id tensors already checked on the device, by id(): (weak reference, version, shape width).
# render_prep_kernel flags ids it cannot decode (and strokes wider than shape_params); reading that
# flag costs a host sync, so it is read once per ids tensor and again only after it changes.
_CHECKED_IDS = {}


def _check_device_ids(stroke_ids, Pmax, status):
    if getattr(stroke_ids, '_n3ds_host_checked', False) or torch.cuda.is_current_stream_capturing():
        return
    rec = _CHECKED_IDS.get(id(stroke_ids))
    if rec is not None and rec[0]() is stroke_ids and rec[1:] == (stroke_ids._version, Pmax):
        return
    if int(status.item()):
        raise RuntimeError('stroke_ids holds an invalid stroke id, or a stroke with more shape parameters than '
                           f'shape_params has columns ({Pmax})')
    if len(_CHECKED_IDS) > 256:
        _CHECKED_IDS.clear()
    _CHECKED_IDS[id(stroke_ids)] = (weakref.ref(stroke_ids), stroke_ids._version, Pmax)


# Kept activations ("store"): per ray, the forward keeps its culled hits and the
# alpha of each evaluated (sample, stroke) pair so the backward neither culls nor
# evaluates alpha again; a ray that outgrows its capacity falls back to
# recomputation chunk by chunk.  Capacities per ray: hits min(S, 2n), pairs 8n.
STORE_BUDGET_BYTES = 8 << 30


def keep_capacities(R, n, S, color_id):
    cap_h = max(1, min(S, 2 * n))
    cap_p = 8 * n
    npf = 4 if color_id == 1 else 1
    nch = (S + 63) // 64
    nbytes = R * (cap_h * 16 + cap_p * npf * 4 + nch * 12)
    return cap_h, cap_p, npf, nch, nbytes


def _keep_buffers(store, R, n, S, color_id, dev):
    if store is False:
        return None
    if isinstance(store, tuple):
        cap_h, cap_p = store
        npf, nch = (4 if color_id == 1 else 1), (S + 63) // 64
    else:
        cap_h, cap_p, npf, nch, nbytes = keep_capacities(R, n, S, color_id)
        if store == 'auto' and nbytes > STORE_BUDGET_BYTES:
            return None
    return (torch.empty(R * nch * 3, dtype=torch.int32, device=dev),
            torch.empty(R * cap_h * 4, dtype=torch.int32, device=dev),
            torch.empty(R * npf * cap_p, dtype=torch.float32, device=dev), cap_h, cap_p)


def _launch_forward(origins, directions, viewdirs, radii, tdist, stroke_ids, shape_params, color_params,
                    density_params, cfg, need_grad):
    """The forward launch: (rgb, acc, depth, per-sample state [5, R*n], stroke table, per-ray
    backgrounds or None, host background colour, kept buffers or None)."""
    (color_id, sdf_delta, laplace, isr, cone, coord_div, bg, cull_eps, adjoint, counters, store, scenes,
     opaque) = cfg
    R, n1 = tdist.shape
    S, Pmax = shape_params.shape
    if scenes is not None:
        scenes.check(R, n1 - 1, S, Pmax, color_params.shape[1], color_id)
    for name, t in (('origins', origins), ('directions', directions), ('viewdirs', viewdirs), ('radii', radii),
                    ('tdist', tdist), ('shape_params', shape_params), ('color_params', color_params),
                    ('density_params', density_params)):
        _lib.require_cuda(name, t)
    _lib.require_cuda('stroke_ids', stroke_ids, dtype=torch.int32)
    dev = origins.device
    table, status = prepare_table(stroke_ids, shape_params)
    _check_device_ids(stroke_ids, Pmax, status)
    bg_rays = None
    if torch.is_tensor(bg) and bg.dim() == 2:  # per-ray background colours [R, 3]
        bg_rays = bg.to(device=origins.device, dtype=torch.float32).contiguous()
        if bg_rays.shape != (R, 3):
            raise RuntimeError(f'bg_color has shape {tuple(bg_rays.shape)}, per-ray colours need ({R}, 3)')
        bgt = torch.ones(3)
    else:
        bgt = torch.as_tensor(bg, dtype=torch.float32).cpu()
    bgv = (_lib.ctypes.c_float * 3)(*[float(v) for v in bgt.reshape(-1).expand(3)])
    rgb = torch.empty(R, 3, dtype=torch.float32, device=dev)
    acc = torch.empty(R, dtype=torch.float32, device=dev)
    depth = torch.empty(R, dtype=torch.float32, device=dev)
    state = torch.empty(5, R * (n1 - 1), dtype=torch.float32, device=dev)
    s_keep = S if scenes is None else scenes.s_max  # kept-hit layout follows the largest scene
    keep = _keep_buffers(store, R, n1 - 1, s_keep, color_id, dev) if need_grad else None
    hd, hb, pb, cap_h, cap_p = keep if keep is not None else (None, None, None, 0, 0)
    if scenes is None:
        fwd, batch_args = _lib.lib().n3ds_render_forward, ()
    else:
        fwd, batch_args = _lib.lib().n3ds_render_batch_forward, scenes.c_args()
    # culling as its own pass (260 B per (ray, chunk of 64 strokes)) unless the caller
    # asked for the low-memory mode
    cws = None
    if store is not False:
        nbytes = int(_lib.lib().n3ds_render_forward_workspace_bytes(R, s_keep))
        cws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
    _lib.check(fwd(
        _lib.ptr(origins), _lib.ptr(directions), _lib.ptr(viewdirs), _lib.ptr(radii), _lib.ptr(tdist),
        _lib.ptr(table), _lib.ptr(color_params), _lib.ptr(density_params), *batch_args, R, n1 - 1, S, Pmax,
        color_params.shape[1], color_id, float(sdf_delta), int(laplace), int(isr), int(cone), float(coord_div),
        _lib.ctypes.cast(bgv, _lib.ctypes.c_void_p), _lib.ptr(bg_rays), int(opaque), float(cull_eps),
        _lib.ptr(rgb), _lib.ptr(acc),
        _lib.ptr(depth), _lib.ptr(state), _lib.ptr(counters), _lib.ptr(hd), _lib.ptr(hb), _lib.ptr(pb), cap_h,
        cap_p, _lib.ptr(cws), 0 if cws is None else int(cws.numel()), _lib.stream()))
    del cws
    return rgb, acc, depth, state, table, bg_rays, bgv, keep


class _render_fn(Function):

    @staticmethod
    def forward(ctx, origins, directions, viewdirs, radii, tdist, stroke_ids, shape_params, color_params,
                density_params, cfg):
        need_grad = any(ctx.needs_input_grad[6:9])
        (rgb, acc, depth, state, table, bg_rays, bgv, keep) = _launch_forward(
            origins, directions, viewdirs, radii, tdist, stroke_ids, shape_params, color_params, density_params, cfg,
            need_grad)
        hd, hb, pb, cap_h, cap_p = keep if keep is not None else (None, None, None, 0, 0)
        empty = torch.empty(0, device=origins.device)
        ctx.save_for_backward(origins, directions, viewdirs, radii, tdist, table, shape_params, color_params,
                              density_params, state, hd if hd is not None else empty,
                              hb if hb is not None else empty, pb if pb is not None else empty,
                              bg_rays if bg_rays is not None else empty)
        ctx.caps = (cap_h, cap_p)
        ctx.cfg = cfg
        ctx.bgv = bgv
        return rgb, acc, depth

    @staticmethod
    def backward(ctx, g_rgb, g_acc, g_depth):
        (origins, directions, viewdirs, radii, tdist, table, shape_params, color_params, density_params,
         state, hd, hb, pb, bg_rays) = ctx.saved_tensors
        (color_id, sdf_delta, laplace, isr, cone, coord_div, bg, cull_eps, adjoint, counters, store,
         scenes, opaque) = ctx.cfg
        cap_h, cap_p = ctx.caps
        R, n1 = tdist.shape
        S, Pmax = shape_params.shape
        Q = color_params.shape[1]
        dev = origins.device
        g_rgb = torch.zeros(R, 3, device=dev) if g_rgb is None else g_rgb.float().contiguous()
        g_acc = None if g_acc is None else g_acc.float().contiguous()
        g_depth = None if g_depth is None else g_depth.float().contiguous()
        if scenes is None:
            nws = _lib.lib().n3ds_render_backward_workspace_floats(R, S, Pmax, Q)
            if hd.numel() > 0 and nws >= 0:  # kept hits: room for the split adjoint
                nsp = _lib.lib().n3ds_render_backward_split_floats(R, S, Pmax, Q)
                nws = -1 if nsp < 0 else max(nws, nsp)
            bwd, batch_args = _lib.lib().n3ds_render_backward, ()
        else:
            nws = _lib.lib().n3ds_render_batch_backward_workspace_floats(scenes.host_ptr(), scenes.n_scenes, n1 - 1,
                                                                        Pmax, Q, color_id)
            bwd, batch_args = _lib.lib().n3ds_render_batch_backward, scenes.c_args()
        if nws < 0:
            _lib.check(2)
        ws = torch.empty(max(int(nws), 1), dtype=torch.float32, device=dev)
        gs = torch.zeros_like(shape_params)
        gc = torch.zeros_like(color_params)
        gd = torch.zeros_like(density_params)
        _lib.check(bwd(
            _lib.ptr(origins), _lib.ptr(directions), _lib.ptr(viewdirs), _lib.ptr(radii), _lib.ptr(tdist),
            _lib.ptr(table), _lib.ptr(color_params), _lib.ptr(density_params), *batch_args, R, n1 - 1, S, Pmax, Q,
            color_id,
            float(sdf_delta), int(laplace), int(isr), int(cone), float(coord_div),
            _lib.ctypes.cast(ctx.bgv, _lib.ctypes.c_void_p), _lib.ptr(bg_rays), int(opaque), float(cull_eps),
            1 if adjoint == 'exact' else 0,
            _lib.ptr(g_rgb), _lib.ptr(g_acc), _lib.ptr(g_depth), _lib.ptr(state), _lib.ptr(hd), _lib.ptr(hb),
            _lib.ptr(pb), cap_h, cap_p, _lib.ptr(ws), int(ws.numel()), _lib.ptr(gs), _lib.ptr(gc), _lib.ptr(gd),
            _lib.stream()))
        return (None, None, None, None, None, None, gs if ctx.needs_input_grad[6] else None,
                gc if ctx.needs_input_grad[7] else None, gd if ctx.needs_input_grad[8] else None, None)


def render_strokes(origins, directions, viewdirs, radii, tdist, stroke_ids, shape_params, color_params,
                   density_params, color_id=0, sdf_delta=1.0, use_laplace_transform=False, inv_scale_radius=False,
                   cone=False, coord_div=1.0, bg_color=(1.0, 1.0, 1.0), cull_eps=1e-7, adjoint=None,
                   counters=None, store='auto', scenes=None, opaque_background=False):
    """Render rays through a stroke field; returns (rgb [R,3], acc [R], depth [R]).

    origins/directions/viewdirs [R,3], radii [R] (sample radius = radii·t if ``cone``
    else radii, as in ``render.cast_rays``), tdist [R, n+1] fenceposts,
    stroke_ids int32 [S], shape_params [S, Pmax], color_params [S, Q],
    density_params [S].  Sample coordinates are ``(o + d·t) / coord_div`` (the
    ``_warp_coords`` bound of ``models.py:19-33``).  Gradients flow to
    shape_params, color_params and density_params.

    ``store``: what the forward keeps for the backward.  ``False``: only the
    per-sample composite state (20 B per ray-sample; the backward recomputes
    culling and alpha); ``True``: also the culled hits and per-pair alpha
    (``keep_capacities``); ``'auto'``: ``True`` when that fits
    ``STORE_BUDGET_BYTES``; a ``(cap_hits, cap_pairs)`` tuple sets the per-ray
    capacities.  Results are identical in every mode.

    ``bg_color``: one background colour (3 floats), or per-ray colours
    ``[R, 3]`` as ``StrokeField`` draws them when ``bg_intensity_range`` is a
    range (``models.py:213-222``); no gradient flows to it.
    ``opaque_background``: the last interval is infinitely wide
    (``render.compute_alpha_weights``, ``utils/render.py:35-40``).
    ``inv_scale_radius`` defaults to False here (the 500-stroke benchmark's
    setting); ``StrokeField`` itself defaults to True (``models.py:540``).

    ``scenes``: a ``multiscene.SceneBatch`` when the arrays hold many
    independent scenes packed one after another (each ray composites only its
    own scene's strokes); every stage then runs as one launch for all scenes.
    """
    return _render_fn.apply(*_render_args(origins, directions, viewdirs, radii, tdist, stroke_ids, shape_params,
                                          color_params, density_params, color_id, sdf_delta, use_laplace_transform,
                                          inv_scale_radius, cone, coord_div, bg_color, cull_eps, adjoint, counters,
                                          store, scenes, opaque_background))


def _render_args(origins, directions, viewdirs, radii, tdist, stroke_ids, shape_params, color_params,
                 density_params, color_id=0, sdf_delta=1.0, use_laplace_transform=False, inv_scale_radius=False,
                 cone=False, coord_div=1.0, bg_color=(1.0, 1.0, 1.0), cull_eps=1e-7, adjoint=None,
                 counters=None, store='auto', scenes=None, opaque_background=False):
    """Checked, contiguous float32 arguments and the configuration tuple of the render op."""
    if color_params.shape[1] != COLOR_WIDTH[color_id]:
        raise RuntimeError(f'color_params has width {color_params.shape[1]}, color_id {color_id} needs '
                           f'{COLOR_WIDTH[color_id]}')
    if stroke_ids.device.type == 'cpu':
        ids = stroke_ids.tolist()
        for k, i in enumerate(ids):
            base, flags = i >> 4, i & 15
            if base >= N_BASE or (base >= 9 and flags) or ((flags & 4) and (flags & 8)):
                raise RuntimeError(f'invalid stroke id {i}')
            if n_shape_params(i) > shape_params.shape[1]:
                raise RuntimeError(f'stroke {k} (id {i}) has {n_shape_params(i)} shape parameters, shape_params '
                                   f'has width {shape_params.shape[1]}')
        stroke_ids = stroke_ids.to(device=shape_params.device, dtype=torch.int32)
        stroke_ids._n3ds_host_checked = True
    cfg = (int(color_id), float(sdf_delta), bool(use_laplace_transform), bool(inv_scale_radius), bool(cone),
           float(coord_div), bg_color, float(cull_eps), adjoint or get_adjoint(), counters, store, scenes,
           bool(opaque_background))
    f = lambda t: t.contiguous().float()  # noqa: E731
    args = (f(origins), f(directions), f(viewdirs), f(radii), f(tdist), stroke_ids.contiguous(), f(shape_params),
            f(color_params), f(density_params), cfg)
    return args


def render_rays(origins, directions, viewdirs, radii, tdist, stroke_ids, shape_params, color_params,
                density_params, t_far=None, compute_extras=False, **kw):
    """Eval-time rendering (no gradients) with the reference's rendering dict
    (``volumetric_rendering``, ``utils/render.py:51-116``): 'rgb', 'depth', 'acc' and, with
    ``compute_extras``, 'distance_mean', 'distance_median', 'distance_percentile_5' and
    'distance_percentile_95' (fenceposts extended by the far distance ``t_far``, [R] or a
    number, ``batch['far']``) from the fused forward's per-sample state
    (``n3ds_render_extras``).  Other arguments as ``render_strokes``."""
    R = tdist.shape[0]
    if compute_extras and t_far is None:
        raise ValueError('compute_extras needs t_far (the far distance of each ray)')
    args = _render_args(origins, directions, viewdirs, radii, tdist, stroke_ids, shape_params, color_params,
                        density_params, **kw)
    with torch.no_grad():
        rgb, acc, depth, state, _, _, _, _ = _launch_forward(*args, need_grad=False)
        out = {'rgb': rgb, 'depth': depth, 'acc': acc}
        if compute_extras:
            far = torch.as_tensor(t_far, dtype=torch.float32, device=rgb.device).reshape(-1)
            far = far.expand(R).contiguous() if far.numel() == 1 else far.contiguous()
            if far.numel() != R:
                raise RuntimeError(f't_far has {far.numel()} values for {R} rays')
            ex = torch.empty(R, 4, dtype=torch.float32, device=rgb.device)
            _lib.check(_lib.lib().n3ds_render_extras(_lib.ptr(state), _lib.ptr(args[4]), _lib.ptr(args[1]),
                                                     _lib.ptr(far), R, tdist.shape[1] - 1, int(args[9][12]),
                                                     _lib.ptr(ex), _lib.stream()))
            for i, k in enumerate(('distance_mean', 'distance_percentile_5', 'distance_median',
                                   'distance_percentile_95')):
                out[k] = ex[:, i]
    return out

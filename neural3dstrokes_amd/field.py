"""Stroke field and error field with the reference's module layout.

``StrokeField`` keeps the configuration attributes, parameter and buffer names
of the reference's ``StrokeField`` (``source/models.py:526-575``:
``shape_params``, ``color_params``, ``density_params``, ``stroke_step``,
``shape_params_grad``), so the ``nerf.*`` entries of a reference checkpoint
load into it, and its parameter bookkeeping: ``clip_params``
(``models.py:577-584``) and ``step_update`` (``models.py:586-658``: strokes
added on the ``step_power`` schedule, near-transparent strokes reset, new
strokes placed at the error field's largest sampled errors or split from the
strokes with the largest shape-gradient moving average).  ``render`` is the
fused operator in place of ``predict_density`` + ``compute_alpha_weights`` +
``volumetric_rendering`` (``models.py:660-711``, ``utils/render.py``): the
same stroke range (strokes before ``stroke_step - max_opt_strokes`` take no
gradient), the same training-time ``sdf_delta`` schedule, the cone radii of
``render.cast_rays`` and the ``_warp_coords`` divisor.

``ErrorField`` is ``ErrorMLP`` (``models.py:507-523``) on the fused proposal
kernels: grid level dim 2, 2^20 rows per level, resolution 256, the density
head, and ``sample_error``.

``load_model_state`` splits a reference checkpoint (``accelerator.save_state``,
``train.py``) into the state dicts of its submodules (``nerf``,
``error_field``, ``prop_mlp_<i>``) for ``load_state_dict``.
"""
import math

import torch
import torch.nn as nn

from . import render as _render
from .checkpoint import load_state, stroke_sdf_id
from .propfield import PropDensityField, prop_field
from .strokelib.strokes import _color_dict, get_stroke


def inv_contract(z):
    """Inverse of the 'contract' warp (``utils/coord.py:14-20``)."""
    eps = torch.finfo(z.dtype).eps
    m2 = torch.sum(z ** 2, dim=-1, keepdim=True).clamp_min(eps)
    return torch.where(m2 <= 1, z, z / (2 * torch.sqrt(m2) - m2).clamp_min(eps))


def unwarp_coords(warp_fn, coords, bbox_size=2.0, no_warp=False):
    """``_unwarp_coords`` (``models.py:36-49``): [-1, 1] coordinates back to world space."""
    if no_warp:
        return coords
    if warp_fn is None:
        return coords * (bbox_size / 2)
    if warp_fn == 'contract':
        return inv_contract(coords * 2.0)
    raise NotImplementedError(f'Unknown warp function {warp_fn}')


class ErrorField(PropDensityField):
    """The reference's ``ErrorMLP`` with ``disable_rgb=True`` (no directional error)."""
    grid_level_dim: int = 2
    grid_log2_hashmap_size: int = 20
    grid_disired_resolution: int = 256

    def sample_error(self, coords, no_warp=False):
        """Error density at world-space points [..., 3] (``models.py:512-523``)."""
        prefix = coords.shape[:-1]
        lin0, lin2 = self.density_layer[0], self.density_layer[2]
        density, _ = prop_field(coords.reshape(-1, 3), self.encoder, lin0.weight, lin0.bias, lin2.weight, lin2.bias,
                                self.density_bias, self.warp_fn, self.bbox_size, no_warp)
        return density.view(prefix)


class StrokeField(nn.Module):
    """A vector stroke field (``models.py:526-575``) rendered by the fused operator."""
    shape_type: str = 'sphere'
    color_type: str = 'constant_rgb'
    init_type: str = 'recon'
    composition_type: str = 'over'
    init_num_strokes: int = 10
    max_num_strokes: int = 500
    max_opt_strokes: int = 500
    density_scale: float = 20.
    sdf_delta: float = 7.0
    sdf_delta_eval: float = 1.0
    use_laplace_transform: bool = False
    inv_scale_radius: bool = True
    no_adaptive_delta: bool = False
    use_error_field: bool = True
    reset_density: float = 0.01
    use_shape_grads: bool = False
    shape_grad_ema: float = 0.99
    shape_split_update_rate: float = 0.05
    disable_density_normals: bool = True
    bbox_size: float = 4.
    warp_fn = None
    min_update_interval: int = 2
    min_reset_interval: int = 50
    max_reset_frac: float = 0.95
    error_point_samples: int = 30000
    step_power: float = 0.5

    def __init__(self, fix_shape_params=False, fix_color_params=False, fix_density_params=False, **kwargs):
        super().__init__()
        for k, v in kwargs.items():
            setattr(self, k, v)
        if self.composition_type != 'over':
            raise NotImplementedError("the fused renderer composites with 'over' only")
        if not self.disable_density_normals:
            raise NotImplementedError('density normals are not part of the fused renderer')
        self.stroke_fn, self.d_shape, self.d_color, self.shape_param_ranges, self.color_param_ranges, \
            self.shape_param_sampler, self.color_param_sampler = get_stroke(self.shape_type, self.color_type,
                                                                            self.init_type)
        self.sdf_id = stroke_sdf_id(self.shape_type)
        self.color_id = _color_dict[self.color_type][0]
        self.shape_params = nn.Parameter(torch.zeros(self.max_num_strokes, self.d_shape),
                                         requires_grad=not fix_shape_params)
        self.color_params = nn.Parameter(torch.zeros(self.max_num_strokes, self.d_color),
                                         requires_grad=not fix_color_params)
        self.density_params = nn.Parameter(torch.ones(self.max_num_strokes), requires_grad=not fix_density_params)
        self.register_buffer('stroke_step', torch.tensor(0, dtype=torch.int32))
        self.register_buffer('shape_params_grad', torch.zeros(self.max_num_strokes, self.d_shape))
        self.stroke_step_limit = None
        self.last_update_step = 0
        self._host_step = None  # host copy of stroke_step (None: read it from the device once)
        self._ids = None        # device stroke ids, built once per device

    def _load_from_state_dict(self, *args, **kwargs):
        self._host_step = None
        super()._load_from_state_dict(*args, **kwargs)

    def n_strokes(self):
        """stroke_step as a host int, read from the device only after it changes outside step_update."""
        if self._host_step is None:
            self._host_step = int(self.stroke_step.item())
        return self._host_step

    def clip_params(self):
        """Clamp the parameters to their ranges (``models.py:577-584``)."""
        for i, (lo, hi) in enumerate(self.shape_param_ranges):
            if lo is not None or hi is not None:
                self.shape_params.data[:, i].clamp_(lo, hi)
        for i, (lo, hi) in enumerate(self.color_param_ranges):
            if lo is not None or hi is not None:
                self.color_params.data[:, i].clamp_(lo, hi)
        self.density_params.data.clamp_(min=0.0)

    @torch.no_grad()
    def step_update(self, cur_step, max_step, error_field=None):
        """Add strokes on the schedule and reset near-transparent ones (``models.py:586-658``).

        Returns (previous stroke count, new stroke count, number of strokes reset) when the
        field was updated, else None.
        """
        steps_per_stroke = max_step // self.max_num_strokes
        if steps_per_stroke <= 1:
            raise AssertionError(f'Too few steps per stroke: {steps_per_stroke}')
        train_frac = cur_step / (max_step - steps_per_stroke)
        prev = self.n_strokes()
        nxt = int(self.max_num_strokes * min(max(train_frac, 0.0), 1.0) ** self.step_power)
        nxt = min(max(nxt + self.init_num_strokes, prev), self.max_num_strokes)
        if self.shape_params.grad is not None:
            self.shape_params_grad.copy_(self.shape_params_grad * self.shape_grad_ema +
                                         self.shape_params.grad * (1 - self.shape_grad_ema))
        reset_density = self.reset_density if self.density_params.requires_grad else -math.inf
        reset = torch.nonzero(self.density_params[:prev] < reset_density).squeeze(1)
        n_reset = reset.numel()
        updated = None
        if nxt - prev >= self.min_update_interval or (
                n_reset > 0 and cur_step - self.last_update_step >= self.min_reset_interval and
                train_frac < self.max_reset_frac):
            dev = self.shape_params.device
            coords_top = None
            split_params = lambda _: None  # noqa: E731
            if prev > self.init_num_strokes and self.use_error_field:
                if error_field is None:
                    raise RuntimeError('use_error_field needs the error field')
                pts = torch.rand((self.error_point_samples, 3), device=dev) * 2 - 1
                errors = error_field.sample_error(unwarp_coords(self.warp_fn, pts, self.bbox_size))
                _, top = torch.topk(errors, k=nxt - prev + n_reset, dim=-1)
                coords_top = pts[top].cpu()
            elif prev > self.init_num_strokes and self.use_shape_grads:
                g = self.shape_params_grad[:prev].sum(-1)
                g[reset] = -math.inf
                _, top = torch.topk(g, k=min(nxt - prev + n_reset, prev - n_reset))
                params_top = self.shape_params[top]

                def split_params(i):
                    if i < params_top.shape[0]:
                        return params_top[i] + self.shape_params_grad[i] * self.shape_split_update_rate
                    return None
            for i in range(prev, nxt):
                sp = split_params(i - prev)
                if sp is None:
                    sp = self.shape_param_sampler(i, coords_top[i - prev] if coords_top is not None else None)
                self.shape_params.data[i] = sp.to(dev)
                self.color_params.data[i] = self.color_param_sampler().to(dev)
            off = nxt - prev
            for i in reset.tolist():
                sp = split_params(off)
                if sp is None:
                    sp = self.shape_param_sampler(nxt, coords_top[off] if coords_top is not None else None)
                self.shape_params.data[i] = sp.to(dev)
                self.color_params.data[i] = self.color_param_sampler().to(dev)
                self.density_params.data[i] = 1.0
                off += 1
            self.stroke_step.fill_(nxt)
            self._host_step = nxt
            self.last_update_step = cur_step
            updated = (prev, nxt, n_reset)
        self.clip_params()
        return updated

    def stroke_range(self):
        """(first optimised stroke, stroke count) as predict_density truncates them (``models.py:666-673``)."""
        step = self.n_strokes()
        fixed = max(0, step - self.max_opt_strokes)
        if self.stroke_step_limit is not None:
            step = min(step, self.stroke_step_limit)
            fixed = min(fixed, self.stroke_step_limit)
        return fixed, step

    def current_sdf_delta(self):
        if self.training:
            step = self.stroke_range()[1]
            frac = min(max(step / self.max_num_strokes, 0), 1)
            return self.sdf_delta * (1 - frac) + self.sdf_delta_eval * frac
        return self.sdf_delta_eval

    def render(self, origins, directions, viewdirs, radii, tdist, bg_color=(1.0, 1.0, 1.0),
               opaque_background=False, **kw):
        """rgb [R,3], acc [R], depth [R] of rays through the field (radii [R] or [R,1], tdist [R, n+1])."""
        return _render.render_strokes(*self._render_args(origins, directions, viewdirs, radii, tdist),
                                      bg_color=bg_color, opaque_background=opaque_background, **self._render_kw(),
                                      **kw)

    def render_eval(self, origins, directions, viewdirs, radii, tdist, t_far, compute_extras=True,
                    bg_color=(1.0, 1.0, 1.0), opaque_background=False, **kw):
        """The reference's rendering dict of an eval pass (no gradients): rgb, depth, acc and the
        compute_extras distances (``utils/render.py:88-116``; t_far = batch['far'])."""
        return _render.render_rays(*self._render_args(origins, directions, viewdirs, radii, tdist), t_far=t_far,
                                   compute_extras=compute_extras, bg_color=bg_color,
                                   opaque_background=opaque_background, **self._render_kw(), **kw)

    def _render_kw(self):
        return dict(color_id=self.color_id, sdf_delta=self.current_sdf_delta(),
                    use_laplace_transform=self.use_laplace_transform, inv_scale_radius=self.inv_scale_radius,
                    coord_div=self.bbox_size / 2, cone=not self.no_adaptive_delta)

    def _render_args(self, origins, directions, viewdirs, radii, tdist):
        if self.warp_fn is not None:
            raise NotImplementedError("the fused renderer takes the linear warp (warp_fn=None) only")
        fixed, step = self.stroke_range()
        radii = radii.reshape(-1).float()
        if self.no_adaptive_delta:  # models.py:728-729: every sample gets the mean cone radius (cone off)
            t = (tdist[:, 1:] + tdist[:, :-1]) / 2
            radii = torch.full_like(radii, float((radii[:, None] * t).mean()))
        sp, cp = self.shape_params[:step], self.color_params[:step]
        dp = self.density_params[:step] * self.density_scale
        if fixed > 0:  # strokes behind the optimised window take no gradient (models.py:686-697)
            sp = torch.cat([sp[:fixed].detach(), sp[fixed:]])
            cp = torch.cat([cp[:fixed].detach(), cp[fixed:]])
            dp = torch.cat([dp[:fixed].detach(), dp[fixed:]])
        dev = self.shape_params.device
        if self._ids is None or self._ids.device != dev:
            self._ids = torch.full((self.max_num_strokes,), self.sdf_id, dtype=torch.int32, device=dev)
        ids = self._ids[:step]
        ids._n3ds_host_checked = True  # one known stroke type whose width is the parameters' width
        return origins, directions, viewdirs, radii, tdist, ids, sp, cp, dp


def load_model_state(path):
    """{submodule name: state dict} of a reference checkpoint file (``nerf``, ``error_field``,
    ``prop_mlp_<i>``, ...), read with loaders that execute nothing from the file."""
    state = load_state(path)
    out = {}
    for k, v in state.items():
        head, _, rest = k.partition('.')
        if rest:
            out.setdefault(head, {})[rest] = v
    return out

"""Fused proposal density field: a drop-in for the reference's ``PropMLP``.

The reference's proposal levels evaluate ``PropMLP`` (``source/models.py``
:286-523 with ``disable_rgb=True``, level_dim 1): ``_warp_coords`` (``:19-33``,
the 'contract' warp of ``utils/coord.py:5-11``), the hash-grid encoder with
bound 1, ``density_layer = Linear(F, 64) -> ReLU -> Linear(64, 1)`` and
``softplus(raw + density_bias)``.  Here the whole chain from sample positions
to densities is one HIP kernel (``n3ds_prop_field_forward``) and its adjoint
another (``n3ds_prop_field_backward``, recomputing features and hidden units
instead of keeping the [N, 64] activations), followed by the encoder's table
scatter (``n3ds_grid_encode_backward``).

``PropDensityField`` keeps ``PropMLP``'s configuration attributes, submodule
and parameter names (``encoder.embeddings``, ``density_layer.0.weight`` ...),
so a reference checkpoint's ``prop_mlp_*`` state dict loads into it, and its
``forward(coords, radius, viewdirs=None, no_warp=False)`` returns the same
dict.  Density normals (``disable_density_normals=False``) and the RGB head are
not part of the proposal path and raise.
"""
import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F
from torch.autograd import Function

from . import _lib
from .gridencoder import GridEncoder
from .gridencoder import grid as _grid

HIDDEN = 64
_WARP = {None: 2, 'contract': 1}


class _prop_field_fn(Function):

    @staticmethod
    def forward(ctx, coords, embeddings, w1, b1, w2, b2, cfg):
        (warp, box, offsets, L, C, S, H, gridtype, align, interp, dbias) = cfg
        for name, t in (('coords', coords), ('w1', w1), ('b1', b1), ('w2', w2), ('b2', b2)):
            _lib.require_cuda(name, t)
        _lib.require_cuda('embeddings', embeddings, dtype=None)
        N = coords.shape[0]
        dev = coords.device
        density = torch.empty(N, dtype=torch.float32, device=dev)
        coord = torch.empty(N, 3, dtype=torch.float32, device=dev)
        keep = any(ctx.needs_input_grad[1:6])
        # the encoder's outputs (F floats per point) and inputs, kept so the adjoint gathers nothing
        feats = torch.empty(N, L * C, dtype=torch.float32, device=dev) if keep else None
        x01 = torch.empty(N, 3, dtype=torch.float32, device=dev) if keep else None
        _lib.check(_lib.lib().n3ds_prop_field_forward(
            _lib.ptr(coords), N, warp, float(box), _lib.ptr(embeddings), _lib.ptr(offsets), L, C, S, H, gridtype,
            align, interp, _grid._dtype_flag(embeddings), _lib.ptr(w1), _lib.ptr(b1), _lib.ptr(w2), _lib.ptr(b2),
            float(dbias), _lib.ptr(density), _lib.ptr(coord), _lib.ptr(feats), _lib.ptr(x01), _lib.stream()))
        empty = torch.empty(0, device=dev)
        ctx.save_for_backward(coords, embeddings, w1, b1, w2, b2, feats if keep else empty, x01 if keep else empty)
        ctx.cfg = cfg
        ctx.mark_non_differentiable(coord)
        return density, coord

    @staticmethod
    def backward(ctx, g_density, g_coord):
        coords, embeddings, w1, b1, w2, b2, feats, x01 = ctx.saved_tensors
        (warp, box, offsets, L, C, S, H, gridtype, align, interp, dbias) = ctx.cfg
        N = coords.shape[0]
        dev = coords.device
        Fd = L * C
        g_density = torch.zeros(N, device=dev) if g_density is None else g_density.float().contiguous()
        gfeat = torch.empty(N, Fd, dtype=torch.float32, device=dev)
        gw1, gb1, gw2, gb2 = (torch.zeros_like(t) for t in (w1, b1, w2, b2))
        _lib.check(_lib.lib().n3ds_prop_field_backward(
            _lib.ptr(coords), N, warp, float(box), _lib.ptr(embeddings), _lib.ptr(offsets), L, C, S, H, gridtype,
            align, interp, _grid._dtype_flag(embeddings), _lib.ptr(w1), _lib.ptr(b1), _lib.ptr(w2), _lib.ptr(b2),
            float(dbias), _lib.ptr(g_density), _lib.ptr(feats), _lib.ptr(gfeat), _lib.ptr(x01), _lib.ptr(gw1),
            _lib.ptr(gb1), _lib.ptr(gw2), _lib.ptr(gb2), _lib.stream()))
        g_emb = None
        if ctx.needs_input_grad[1]:
            gfeat_t = gfeat.to(embeddings.dtype).contiguous()
            g32 = torch.zeros(embeddings.shape, dtype=torch.float32, device=dev)
            g_emb = g32 if embeddings.dtype == torch.float32 else torch.empty_like(embeddings)
            rows = embeddings.shape[0]
            nws = int(_lib.lib().n3ds_grid_encode_backward_workspace_bytes(N, 3, C, L, rows))
            ws = torch.empty(max(nws, 1), dtype=torch.uint8, device=dev) if _grid.SCATTER == 'bucketed' else None
            _lib.check(_lib.lib().n3ds_grid_encode_backward(
                _lib.ptr(gfeat_t), _lib.ptr(x01), _lib.ptr(offsets), _lib.ptr(g32), _lib.ptr(g_emb), rows, N, 3, C,
                L, S, H, None, None, gridtype, align, interp, _grid._dtype_flag(embeddings), _lib.ptr(ws),
                nws if ws is not None else 0, _lib.stream()))
        return None, g_emb, gw1, gb1, gw2, gb2, None


def prop_field(coords, encoder, w1, b1, w2, b2, density_bias, warp_fn='contract', bbox_size=4.0, no_warp=False):
    """density [N], warped coords [N, 3] of points coords [N, 3] (see PropDensityField)."""
    warp = 0 if no_warp else _WARP[warp_fn]
    cfg = (warp, float(bbox_size), encoder.offsets, encoder.num_levels, encoder.level_dim,
           float(np.log2(encoder.per_level_scale)), int(encoder.base_resolution), encoder.gridtype_id,
           int(bool(encoder.align_corners)), encoder.interp_id, float(density_bias))
    f = lambda t: t.contiguous().float()
    return _prop_field_fn.apply(f(coords), encoder.embeddings.contiguous(), f(w1), f(b1), f(w2), f(b2), cfg)


class PropDensityField(nn.Module):
    """Proposal density field with the reference ``PropMLP``'s attributes and parameter names."""
    density_bias: float = -1.
    disable_density_normals: bool = True
    disable_rgb: bool = True
    bbox_size: float = 4.
    warp_fn = 'contract'
    use_grid_encoder: bool = True
    grid_level_interval: int = 2
    grid_level_dim: int = 1
    grid_base_resolution: int = 16
    grid_disired_resolution: int = 8192
    grid_log2_hashmap_size: int = 21

    def __init__(self, **kwargs):
        super().__init__()
        for k, v in kwargs.items():
            setattr(self, k, v)
        if not self.use_grid_encoder or not self.disable_rgb or not self.disable_density_normals:
            raise RuntimeError('the fused proposal field covers the grid-encoder density head only '
                               '(use_grid_encoder=True, disable_rgb=True, disable_density_normals=True)')
        self.grid_num_levels = int(np.log(self.grid_disired_resolution / self.grid_base_resolution) /
                                   np.log(self.grid_level_interval)) + 1
        self.encoder = GridEncoder(input_dim=3, num_levels=self.grid_num_levels, level_dim=self.grid_level_dim,
                                   base_resolution=self.grid_base_resolution,
                                   desired_resolution=self.grid_disired_resolution,
                                   log2_hashmap_size=self.grid_log2_hashmap_size, gridtype='hash',
                                   align_corners=False)
        self.density_layer = nn.Sequential(nn.Linear(self.encoder.output_dim, HIDDEN), nn.ReLU(),
                                           nn.Linear(HIDDEN, 1))

    def forward(self, coords, radius=None, viewdirs=None, no_warp=False):
        prefix = coords.shape[:-1]
        lin0, lin2 = self.density_layer[0], self.density_layer[2]
        density, coord = prop_field(coords.reshape(-1, 3), self.encoder, lin0.weight, lin0.bias, lin2.weight,
                                    lin2.bias, self.density_bias, self.warp_fn, self.bbox_size, no_warp)
        density = density.view(prefix)
        hash_levelwise_mean = None
        if self.training:
            # hash decay statistics of models.py:481-493: the mean of the squared table entries of
            # each level; the rows of a level are contiguous (offsets), so a slice per level
            # replaces the scatter_reduce over all rows
            param_sq = torch.square(self.encoder.embeddings.float())
            offs = self.encoder.offsets.tolist() if not hasattr(self, '_offs') else self._offs
            self._offs = offs
            hash_levelwise_mean = torch.stack([param_sq[offs[l]:offs[l + 1]].mean(0)
                                               for l in range(self.encoder.num_levels)])
        return dict(coord=coord.view(*prefix, 3), density=density,
                    rgb=torch.zeros(density.shape + (3,), device=density.device), normals=None,
                    hash_levelwise_mean=hash_levelwise_mean)


def reference_forward(module, coords, no_warp=False):
    """The same field as plain torch ops over the module's parameters (the unfused formulation:
    the warp as torch ops, GridEncoder, nn.Linear layers, softplus); timing comparator."""
    prefix = coords.shape[:-1]
    x = coords.reshape(-1, 3).float()
    if not no_warp:
        if module.warp_fn == 'contract':
            eps = torch.finfo(x.dtype).eps
            m2 = torch.sum(x ** 2, dim=-1, keepdim=True).clamp_min(eps)
            x = torch.where(m2 <= 1, x, ((2 * torch.sqrt(m2) - 1) / m2) * x) / 2.0
        else:
            x = x / (module.bbox_size / 2)
    feats = module.encoder(x, bound=1.0)
    raw = module.density_layer(feats)[..., 0]
    return F.softplus(raw + module.density_bias).view(prefix)

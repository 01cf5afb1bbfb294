"""Drop-in hash-grid encoder: the reference's ``source/gridencoder/grid.py`` surface.

``grid_encode`` / ``_grid_encode`` / ``GridEncoder`` keep the reference's
arguments, level bookkeeping and initialisation (``grid.py:24-194``); the
gathers and scatters run in ``libn3ds.so``.  Differences in mechanism, not in
results: outputs are written directly in ``[B, L*C]`` order (no permute copy),
fp16 tables are interpolated and accumulated in fp32, and the adjoint
accumulates fp16 tables' gradients in an fp32 buffer before rounding once.
"""
import numpy as np
import torch
import torch.nn as nn
from torch.autograd import Function

from .. import _lib

_gridtype_to_id = {'hash': 0, 'tiled': 1}
# table-gradient scatter of the backward: 'bucketed' (records grouped by row
# bucket and summed in LDS; needs a workspace of about (4 + 4C) bytes per
# (point, level, corner)) or 'atomic' (one float atomic per contribution)
SCATTER = 'bucketed'
_interp_to_id = {'linear': 0, 'smoothstep': 1}


def _dtype_flag(t):
    if t.dtype == torch.float32:
        return 0
    if t.dtype == torch.float16:
        return 1
    raise RuntimeError(f'embeddings must be float32 or float16, got {t.dtype}')


class _grid_encode(Function):
    """grid.py:24-89: inputs [B, D] in [0, 1] -> [B, L*C]."""

    @staticmethod
    @torch.amp.custom_fwd(device_type='cuda')
    def forward(ctx, inputs, embeddings, offsets, per_level_scale, base_resolution, calc_grad_inputs=False,
                gridtype=0, align_corners=False, interpolation=0):
        inputs = inputs.contiguous()
        B, D = inputs.shape
        L = offsets.shape[0] - 1
        C = embeddings.shape[1]
        S = float(np.log2(per_level_scale))
        H = int(base_resolution)
        if torch.is_autocast_enabled() and C % 2 == 0:
            embeddings = embeddings.to(torch.half)
        embeddings = embeddings.contiguous()
        _lib.require_cuda('inputs', inputs)
        _lib.require_cuda('embeddings', embeddings, dtype=None)
        _lib.require_cuda('offsets', offsets, dtype=torch.int32)
        outputs = torch.empty(B, L * C, device=inputs.device, dtype=embeddings.dtype)
        dy_dx = torch.empty(B, L * D * C, device=inputs.device, dtype=embeddings.dtype) if calc_grad_inputs else None
        _lib.check(_lib.lib().n3ds_grid_encode_forward(
            _lib.ptr(inputs), _lib.ptr(embeddings), _lib.ptr(offsets), _lib.ptr(outputs), B, D, C, L, S, H,
            _lib.ptr(dy_dx), int(gridtype), int(bool(align_corners)), int(interpolation), _dtype_flag(embeddings),
            _lib.stream()))
        ctx.save_for_backward(inputs, embeddings, offsets, dy_dx)
        ctx.dims = (B, D, C, L, S, H, gridtype, interpolation, align_corners)
        return outputs

    @staticmethod
    @torch.amp.custom_bwd(device_type='cuda')
    def backward(ctx, grad):
        inputs, embeddings, offsets, dy_dx = ctx.saved_tensors
        B, D, C, L, S, H, gridtype, interpolation, align_corners = ctx.dims
        grad = grad.to(embeddings.dtype).contiguous()
        g32 = torch.zeros(embeddings.shape, dtype=torch.float32, device=embeddings.device)
        g_out = g32 if embeddings.dtype == torch.float32 else torch.empty_like(embeddings)
        grad_inputs = torch.zeros_like(inputs, dtype=embeddings.dtype) if dy_dx is not None else None
        rows = embeddings.shape[0]
        nws = int(_lib.lib().n3ds_grid_encode_backward_workspace_bytes(B, D, C, L, rows))
        ws = torch.empty(max(nws, 1), dtype=torch.uint8, device=embeddings.device) if SCATTER == 'bucketed' else None
        _lib.check(_lib.lib().n3ds_grid_encode_backward(
            _lib.ptr(grad), _lib.ptr(inputs), _lib.ptr(offsets), _lib.ptr(g32), _lib.ptr(g_out), rows,
            B, D, C, L, S, H, _lib.ptr(dy_dx), _lib.ptr(grad_inputs), int(gridtype), int(bool(align_corners)),
            int(interpolation), _dtype_flag(embeddings), _lib.ptr(ws), nws if ws is not None else 0, _lib.stream()))
        if grad_inputs is not None:
            grad_inputs = grad_inputs.to(inputs.dtype)
        return grad_inputs, g_out, None, None, None, None, None, None, None


grid_encode = _grid_encode.apply


class GridEncoder(nn.Module):
    """Multiresolution hash-grid encoder module (grid.py:96-194)."""

    def __init__(self, input_dim=3, num_levels=16, level_dim=2, per_level_scale=2, base_resolution=16,
                 log2_hashmap_size=19, desired_resolution=None, gridtype='hash', align_corners=False,
                 interpolation='linear', init_std=1e-4):
        super().__init__()
        if desired_resolution is not None:
            per_level_scale = np.exp2(np.log2(desired_resolution / base_resolution) / (num_levels - 1))
        self.input_dim = input_dim
        self.num_levels = num_levels
        self.level_dim = level_dim
        self.per_level_scale = per_level_scale
        self.log2_hashmap_size = log2_hashmap_size
        self.base_resolution = base_resolution
        self.output_dim = num_levels * level_dim
        self.gridtype = gridtype
        self.gridtype_id = _gridtype_to_id[gridtype]
        self.interpolation = interpolation
        self.interp_id = _interp_to_id[interpolation]
        self.align_corners = align_corners
        self.init_std = init_std
        self.max_params = 2 ** log2_hashmap_size
        offsets, resolutions = [], []
        total = 0
        for level in range(num_levels):
            res = int(np.ceil(base_resolution * per_level_scale ** level))
            res = res if align_corners else res + 1
            rows = int(np.ceil(min(self.max_params, res ** input_dim) / 8) * 8)
            resolutions.append(res)
            offsets.append(total)
            total += rows
        offsets.append(total)
        self.register_buffer('offsets', torch.from_numpy(np.array(offsets, dtype=np.int32)))
        idx = torch.empty(total, dtype=torch.long)
        for level in range(num_levels):
            idx[offsets[level]:offsets[level + 1]] = level
        self.register_buffer('idx', idx)
        self.register_buffer('grid_sizes', torch.from_numpy(np.array(resolutions, dtype=np.int32)))
        self.n_params = total * level_dim
        self.embeddings = nn.Parameter(torch.empty(total, level_dim))
        self.reset_parameters()

    def reset_parameters(self):
        self.embeddings.data.uniform_(-self.init_std, self.init_std)

    def __repr__(self):
        top = int(round(self.base_resolution * self.per_level_scale ** (self.num_levels - 1)))
        return (f'GridEncoder: input_dim={self.input_dim} num_levels={self.num_levels} level_dim={self.level_dim} '
                f'resolution={self.base_resolution} -> {top} per_level_scale={self.per_level_scale:.4f} '
                f'params={tuple(self.embeddings.shape)} gridtype={self.gridtype} align_corners={self.align_corners} '
                f'interpolation={self.interpolation}')

    def forward(self, inputs, bound=1):
        inputs = (inputs + bound) / (2 * bound)
        prefix = list(inputs.shape[:-1])
        inputs = inputs.view(-1, self.input_dim)
        out = grid_encode(inputs, self.embeddings, self.offsets, self.per_level_scale, self.base_resolution,
                          inputs.requires_grad, self.gridtype_id, self.align_corners, self.interp_id)
        return out.view(prefix + [self.output_dim])

    @torch.amp.autocast('cuda', enabled=False)
    def grad_total_variation(self, weight=1e-7, inputs=None, bound=1, B=1000000):
        """Adds the total-variation gradient into ``embeddings.grad`` (grid.py:172-194)."""
        D = self.input_dim
        C = self.embeddings.shape[1]
        L = self.offsets.shape[0] - 1
        S = float(np.log2(self.per_level_scale))
        H = int(self.base_resolution)
        if inputs is None:
            inputs = torch.rand(B, self.input_dim, device=self.embeddings.device)
        else:
            inputs = ((inputs + bound) / (2 * bound)).view(-1, self.input_dim)
            B = inputs.shape[0]
        if self.embeddings.grad is None:
            raise ValueError('grad is None, should be called after loss.backward() and before optimizer.step()!')
        inputs = inputs.float().contiguous()
        emb = self.embeddings.detach().contiguous()
        g32 = self.embeddings.grad.float().contiguous()
        _lib.check(_lib.lib().n3ds_grid_grad_tv(
            _lib.ptr(inputs), _lib.ptr(emb), _lib.ptr(g32), _lib.ptr(self.offsets), float(weight), B, D, C, L, S, H,
            self.gridtype_id, int(bool(self.align_corners)), _dtype_flag(emb), _lib.stream()))
        self.embeddings.grad.copy_(g32.to(self.embeddings.grad.dtype))

"""Stroke textures: drop-in for the reference's ``source/textures.py``.

``ImageTexture`` tints stroke colours (and optionally alphas) with a bilinear
sample of an image at each (sample, stroke) texture coordinate that the stroke
operator returns with ``return_texcoord=True``; the reference applies it between
the stroke and compositing operators (``models.py:704-706``).  The image is set up
as in the reference (loaded as float32, divided by 255, ``F.interpolate(mode='area')``
to ``texture_size``; ``textures.py:45-49``); on first use on a device it is packed
by ``n3ds_texture_pack`` into the layout of the gfx950 kernels ``n3ds_texture_forward``
/ ``n3ds_texture_backward`` (one entry per bilinear cell with its four corners).  As in
the reference, the image has no gradient.
"""
import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import _lib


def get_stroke_texture(config):
    """textures.py:10-16: ``config.stroke_texture`` in {'none', 'image'}."""
    if config.stroke_texture == 'none':
        return None
    if config.stroke_texture == 'image':
        return ImageTexture(config.texture_image_path, config.texture_image_size)
    raise AssertionError(f'Unknown stroke texture {config.stroke_texture}')


class StrokeTexture(nn.Module):
    """textures.py:19-33: ``forward(texcoords [..., 2], colors [..., 3], alphas [...])``
    returns the textured ``(colors, alphas)``."""

    def forward(self, texcoords, colors, alphas):
        raise NotImplementedError


def _load_image(path):
    """misc.load_img (utils/misc.py:72-75): the image file as a float32 array."""
    from PIL import Image
    return np.array(Image.open(path), dtype=np.float32)


def _aligned(t, align):
    t = t.contiguous()
    return t if t.data_ptr() % align == 0 else t.clone()


def _texture_forward(quads, hw, channels, texcoords, colors, alphas, tint_ratio, modulate_alpha):
    for name, t in (('texcoords', texcoords), ('colors', colors), ('alphas', alphas), ('texture', quads)):
        _lib.require_cuda(name, t)
    colors_out = torch.empty_like(colors)
    alphas_out = torch.empty_like(alphas) if modulate_alpha else None
    _lib.check(_lib.lib().n3ds_texture_forward(
        _lib.ptr(colors_out), _lib.ptr(alphas_out), _lib.ptr(texcoords), _lib.ptr(colors),
        _lib.ptr(alphas) if modulate_alpha else None, _lib.ptr(quads), hw[0], hw[1], channels,
        colors.shape[0], float(tint_ratio), int(bool(modulate_alpha)), _lib.stream()))
    return colors_out, alphas_out


class _texture_fn(Function):
    """(texcoords [N,2], colors [N,3], alphas [N]) -> (colors [N,3], alphas [N] or None)."""

    @staticmethod
    @torch.amp.custom_fwd(device_type='cuda')
    def forward(ctx, texcoords, colors, alphas, quads, hw, channels, tint_ratio, modulate_alpha):
        texcoords = _aligned(texcoords.float(), 8)
        colors = colors.contiguous().float()
        alphas = alphas.contiguous().float()
        colors_out, alphas_out = _texture_forward(quads, hw, channels, texcoords, colors, alphas, tint_ratio,
                                                  modulate_alpha)
        ctx.save_for_backward(texcoords, colors, alphas, quads)
        ctx.cfg = (hw, channels, tint_ratio, modulate_alpha)
        if alphas_out is None:
            # alphas pass through unchanged (textures.py:79-80): the caller keeps its own tensor
            alphas_out = torch.zeros(0, dtype=torch.float32, device=colors.device)
            ctx.mark_non_differentiable(alphas_out)
        return colors_out, alphas_out

    @staticmethod
    @once_differentiable
    @torch.amp.custom_bwd(device_type='cuda')
    def backward(ctx, grad_colors_out, grad_alphas_out):
        texcoords, colors, alphas, quads = ctx.saved_tensors
        hw, channels, tint_ratio, modulate_alpha = ctx.cfg
        need_uv = ctx.needs_input_grad[0]
        gc_out = None if grad_colors_out is None else grad_colors_out.contiguous().float()
        ga_out = None if (grad_alphas_out is None or not modulate_alpha) else grad_alphas_out.contiguous().float()
        g_colors = torch.empty_like(colors)
        g_alphas = torch.empty_like(alphas) if modulate_alpha else None
        g_uv = torch.empty_like(texcoords) if need_uv else None
        _lib.check(_lib.lib().n3ds_texture_backward(
            _lib.ptr(g_colors), _lib.ptr(g_alphas), _lib.ptr(g_uv), _lib.ptr(gc_out), _lib.ptr(ga_out),
            _lib.ptr(texcoords), _lib.ptr(colors), _lib.ptr(alphas), _lib.ptr(quads), hw[0], hw[1], channels,
            colors.shape[0], float(tint_ratio), int(bool(modulate_alpha)), _lib.stream()))
        return g_uv, g_colors, g_alphas, None, None, None, None, None


class ImageTexture(StrokeTexture):
    """textures.py:36-82.

    ``tint_image_path``: an image file (as in the reference) or an image array
    ``[H, W]`` / ``[H, W, C]`` with values in 0..255; ``texture_size = (H', W')``.
    Keyword arguments set ``modulate_alpha`` (default False) and ``tint_ratio``
    (default 1.0).
    """
    modulate_alpha: bool = False
    tint_ratio: float = 1.0

    def __init__(self, tint_image_path, texture_size, **kwargs):
        super().__init__()
        for k, v in kwargs.items():
            setattr(self, k, v)
        if isinstance(tint_image_path, torch.Tensor):
            image = tint_image_path.detach().to('cpu', torch.float32)
        elif isinstance(tint_image_path, np.ndarray):
            image = torch.from_numpy(np.asarray(tint_image_path, dtype=np.float32))
        else:
            image = torch.from_numpy(_load_image(tint_image_path))
        if image.ndim == 2:
            image = image.unsqueeze(-1)
        image = (image / 255.).permute(2, 0, 1).unsqueeze(0)  # [1, C, H, W]
        image = F.interpolate(image, tuple(texture_size), mode='area')
        self.register_buffer('image_texture', image, persistent=False)
        self._packed_key, self._packed_quads = None, None

    @property
    def channels(self):
        return int(self.image_texture.shape[1])

    def _check_channels(self):
        if self.channels not in (1, 3):
            raise AssertionError('Texture should be RGB or grayscale')

    @property
    def size(self):
        return tuple(int(v) for v in self.image_texture.shape[2:])

    def packed(self):
        """The image in the kernels' quad layout on the image's device (packed again
        after the image buffer moves or changes)."""
        img = self.image_texture
        key = (img.device, img.data_ptr(), img._version)
        if self._packed_key != key:
            _lib.require_cuda('image_texture', img, contiguous=False)
            (h, w), c = self.size, self.channels
            lib = _lib.lib()
            n = int(lib.n3ds_texture_pack_floats(h, w, c))
            if n < 0:
                _lib.check(1)
            quads = torch.empty(n, dtype=torch.float32, device=img.device)
            src = img[0].contiguous()
            _lib.check(lib.n3ds_texture_pack(_lib.ptr(quads), _lib.ptr(src), h, w, c, _lib.stream()))
            self._packed_key, self._packed_quads = key, quads
        return self._packed_quads

    def sample_texture(self, texcoords):
        """textures.py:51-61: texture values [N, C] at texcoords [N, 2]."""
        self._check_channels()
        tc = texcoords.reshape(-1, 2)
        ones = torch.ones(tc.shape[0], 3, dtype=torch.float32, device=tc.device)
        alphas = torch.zeros(tc.shape[0], dtype=torch.float32, device=tc.device)
        tint, _ = _texture_fn.apply(tc, ones, alphas, self.packed(), self.size, self.channels, 1.0, False)
        return tint[:, :self.channels]

    def forward(self, texcoords, colors, alphas):
        assert colors.shape[-1] == 3, 'Colors should be RGB for ImageTexture'
        self._check_channels()
        pre_shape = colors.shape[:-1]
        modulate = bool(self.modulate_alpha)
        colors_out, alphas_out = _texture_fn.apply(texcoords.reshape(-1, 2), colors.reshape(-1, 3),
                                                   alphas.reshape(-1), self.packed(), self.size, self.channels,
                                                   float(self.tint_ratio), modulate)
        return colors_out.reshape(*pre_shape, 3), (alphas_out.reshape(*pre_shape) if modulate else alphas)

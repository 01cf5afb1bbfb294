"""Oracle restatement of the image texture on stroke colours (test infrastructure only).

Follows ``source/textures.py:51-82``: the tint is a bilinear sample of the image
at ``texcoords * 2 - 1`` with ``grid_sample(align_corners=True)`` and zero
padding, i.e. at pixel ``(ix, iy) = (u * (W - 1), v * (H - 1))`` with corners
outside the image contributing zero; colours become
``(c * tint_ratio + 1 - tint_ratio) * tint`` and, with ``modulate_alpha``,
alphas ``alpha * mean(tint)``.  The adjoint is what autograd gives those ops
(the image is a buffer without gradient); the texcoord gradient differentiates
the bilinear weights with the out-of-range corners held at zero.

Written as explicit per-corner numpy loops in float64 over the texture laid out
``[C, H, W]`` (the reference's ``image_texture[0]``).  Pinned by
``tests/test_oracle_texture.py`` against ``torch.nn.functional.grid_sample`` (the
library routine the reference calls) and its autograd, and against closed
forms (texel centres return the texel, a constant image returns the constant).
"""
import numpy as np


def _taps(image, texcoords):
    """Corner weights, their derivatives w.r.t. (ix, iy), and corner values.

    image [C, H, W]; texcoords [N, 2].  Returns lists over the four corners of
    (w [N], dw/dix [N], dw/diy [N], val [N, C])."""
    C, H, W = image.shape
    u = texcoords[:, 0].astype(np.float64)
    v = texcoords[:, 1].astype(np.float64)
    ix = ((u * 2 - 1) + 1) / 2 * (W - 1)
    iy = ((v * 2 - 1) + 1) / 2 * (H - 1)
    with np.errstate(invalid='ignore'):
        x0 = np.floor(ix)
        y0 = np.floor(iy)
    taps = []
    for dx, dy in ((0, 0), (1, 0), (0, 1), (1, 1)):
        cx, cy = x0 + dx, y0 + dy
        # weight of a corner: (1 - |ix - cx|)(1 - |iy - cy|) inside the cell
        wx = (ix - x0) if dx else (x0 + 1 - ix)
        wy = (iy - y0) if dy else (y0 + 1 - iy)
        dwx = 1.0 if dx else -1.0
        dwy = 1.0 if dy else -1.0
        inb = (cx >= 0) & (cx <= W - 1) & (cy >= 0) & (cy <= H - 1)
        inb &= np.isfinite(ix) & np.isfinite(iy)
        val = np.zeros((len(u), C))
        xi = np.where(inb, cx, 0).astype(np.int64)
        yi = np.where(inb, cy, 0).astype(np.int64)
        val[inb] = image[:, yi[inb], xi[inb]].T
        w = np.where(inb, wx * wy, 0.0)
        taps.append((w, np.where(inb, dwx * wy, 0.0), np.where(inb, wx * dwy, 0.0), val))
    return taps, W, H


def sample_texture(image, texcoords):
    """textures.py:51-61: tint [N, C]."""
    taps, _, _ = _taps(image, texcoords)
    return sum(w[:, None] * val for w, _, _, val in taps)


def texture_forward(image, texcoords, colors, alphas, tint_ratio=1.0, modulate_alpha=False):
    """textures.py:63-82 on flattened pairs: (colors [N, 3], alphas [N])."""
    tint = sample_texture(image, texcoords)
    cmod = colors * tint_ratio + (1.0 - tint_ratio)
    out_c = cmod * tint  # tint [N, 1] broadcasts for a grey image
    out_a = alphas * tint.mean(-1) if modulate_alpha else alphas
    return out_c, out_a


def texture_backward(image, texcoords, colors, alphas, g_colors, g_alphas, tint_ratio=1.0, modulate_alpha=False):
    """Adjoint of texture_forward: (d colors [N, 3], d alphas [N], d texcoords [N, 2])."""
    taps, W, H = _taps(image, texcoords)
    C = image.shape[0]
    tint = sum(w[:, None] * val for w, _, _, val in taps)
    cmod = colors * tint_ratio + (1.0 - tint_ratio)
    d_cmod = g_colors * tint
    d_colors = d_cmod * tint_ratio
    d_tint = g_colors * cmod
    if C == 1:
        d_tint = d_tint.sum(-1, keepdims=True)
    if modulate_alpha:
        d_alphas = g_alphas * tint.mean(-1)
        d_tint = d_tint + (g_alphas * alphas)[:, None] / C
    else:
        d_alphas = g_alphas
    d_ix = sum((d_tint * val).sum(-1) * dwx for _, dwx, _, val in taps)
    d_iy = sum((d_tint * val).sum(-1) * dwy for _, _, dwy, val in taps)
    d_uv = np.stack([d_ix * (W - 1), d_iy * (H - 1)], -1)
    return d_colors, d_alphas, d_uv

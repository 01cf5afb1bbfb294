"""Oracle of the proposal density field (test infrastructure only).

Restates the reference's PropMLP density head (``source/models.py:286-523``
with ``disable_rgb=True``): ``_warp_coords`` (``models.py:19-33``) with
``coord.contract`` (``utils/coord.py:5-11``), the encoder's bound-1 mapping
(``gridencoder/grid.py:158``), the grid features from :func:`oracle.grid.grid_forward`,
``Linear(F, 64) -> ReLU -> Linear(64, 1)`` and ``softplus(raw + density_bias)``
in float64.  Gradients: torch.autograd of the MLP part; the table gradient
from :func:`oracle.grid.grid_backward` of the feature gradient.

Pinned by ``tests/test_oracle_propfield.py``: the contraction against its
closed form (identity in the unit ball, ``|z| = 2 - 1/|x|`` outside, radial),
the head against ``torch.nn.Linear`` / ``F.softplus`` modules, and central
finite differences of the whole float64 chain for the weight and table
gradients.
"""
import numpy as np
import torch

from . import grid as ogrid


def contract(x):
    """coord.contract, utils/coord.py:5-11 (float64 eps of the input dtype)."""
    eps = torch.finfo(x.dtype).eps
    m2 = torch.sum(x ** 2, dim=-1, keepdim=True).clamp_min(eps)
    return torch.where(m2 <= 1, x, ((2 * torch.sqrt(m2) - 1) / m2) * x)


def warp_coords(coords, warp_fn='contract', bbox_size=4.0, no_warp=False):
    """models.py:19-33: into [-1, 1]."""
    if no_warp:
        return coords
    if warp_fn is None:
        return coords / (bbox_size / 2)
    return contract(coords) / 2.0


def features(coords, embeddings, offsets, per_level_scale, base_resolution, warp_fn='contract', bbox_size=4.0,
             no_warp=False, warped=None, f32_grid=False):
    """Encoder outputs [N, L*C] (float64 numpy) and the encoder inputs in [0, 1].

    ``warped``: use these warped coordinates (in [-1, 1]) instead of warping ``coords``;
    ``f32_grid``: the grid's lattice positions in float32 as the reference kernel computes
    them (``oracle.grid.grid_forward(f32=True)``), with the inputs mapped to [0, 1] in float32."""
    if warped is None:
        cw = warp_coords(torch.as_tensor(coords, dtype=torch.float64), warp_fn, bbox_size, no_warp)
    else:
        cw = torch.as_tensor(warped, dtype=torch.float64)
    x01 = ((cw.float() + 1.0) / 2.0).double().numpy() if f32_grid else ((cw + 1.0) / 2.0).numpy()
    out, _ = ogrid.grid_forward(x01, np.asarray(embeddings, dtype=np.float64), np.asarray(offsets),
                                per_level_scale, base_resolution, f32=f32_grid)
    return np.asarray(out, dtype=np.float64), x01, cw


def density_and_grads(coords, embeddings, offsets, per_level_scale, base_resolution, w1, b1, w2, b2, density_bias,
                      grad_density, warp_fn='contract', bbox_size=4.0, no_warp=False, warped=None, f32_grid=False):
    """Forward densities and the gradients of sum(density * grad_density).

    Returns dict(density, coord, grad_w1, grad_b1, grad_w2, grad_b2, grad_embeddings)."""
    feats, x01, cw = features(coords, embeddings, offsets, per_level_scale, base_resolution, warp_fn, bbox_size,
                              no_warp, warped, f32_grid)
    f = torch.tensor(feats, dtype=torch.float64, requires_grad=True)
    W1, B1, W2, B2 = (torch.as_tensor(t, dtype=torch.float64).clone().requires_grad_(True) for t in (w1, b1, w2, b2))
    h = torch.relu(f @ W1.T + B1)
    raw = (h @ W2.T + B2)[:, 0]
    density = torch.nn.functional.softplus(raw + density_bias)
    gd = torch.as_tensor(grad_density, dtype=torch.float64)
    gW1, gB1, gW2, gB2, gf = torch.autograd.grad((density * gd).sum(), [W1, B1, W2, B2, f])
    g_emb, _ = ogrid.grid_backward(gf.numpy(), x01, np.asarray(embeddings, dtype=np.float64), np.asarray(offsets),
                                   per_level_scale, base_resolution, f32=f32_grid)
    return dict(density=density.detach(), coord=cw, grad_w1=gW1, grad_b1=gB1, grad_w2=gW2, grad_b2=gB2,
                grad_embeddings=torch.from_numpy(g_emb))

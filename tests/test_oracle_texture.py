"""CPU pins of the texture oracle (oracle/texture.py) against torch's grid_sample,
the library routine the reference's ImageTexture calls (textures.py:51-61), its
autograd, and closed forms."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import texture as O


def _case(seed, n, c, h=9, w=13, lo=-0.15, hi=1.15):
    g = np.random.default_rng(seed)
    image = g.uniform(0, 1, (c, h, w))
    tc = g.uniform(lo, hi, (n, 2))
    # exact edges and texel centres
    tc[:4] = [[0, 0], [1, 1], [0, 1], [1, 0]]
    tc[4:8, 0] = np.arange(4) / (w - 1)
    tc[4:8, 1] = np.arange(4) / (h - 1)
    colors = g.uniform(0, 1, (n, 3))
    alphas = g.uniform(0, 1, n)
    return image, tc, colors, alphas


def _torch_reference(image, tc, colors, alphas, tint_ratio, modulate_alpha):
    """The ops of textures.py:51-82 in float64 torch."""
    img = torch.from_numpy(image)[None]
    t = torch.from_numpy(tc).requires_grad_(True)
    c = torch.from_numpy(colors).requires_grad_(True)
    a = torch.from_numpy(alphas).requires_grad_(True)
    uvs = t.reshape(1, 1, -1, 2) * 2 - 1
    tint = F.grid_sample(img, uvs, align_corners=True)[0, :, 0, :].permute(1, 0)
    cc = (c * tint_ratio + (1.0 - tint_ratio)) * tint
    aa = a * tint.mean(-1) if modulate_alpha else a
    return (t, c, a), tint, cc, aa


@pytest.mark.parametrize('channels', [3, 1])
def test_sample_matches_grid_sample(channels):
    image, tc, _, _ = _case(0, 500, channels)
    _, tint, _, _ = _torch_reference(image, tc, np.zeros((500, 3)), np.zeros(500), 1.0, False)
    np.testing.assert_allclose(O.sample_texture(image, tc), tint.detach().numpy(), rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize('channels,tint_ratio,modulate', [(3, 1.0, False), (3, 0.6, True), (1, 0.25, True),
                                                          (1, 1.0, False)])
def test_forward_backward_match_torch_autograd(channels, tint_ratio, modulate):
    n = 400
    image, tc, colors, alphas = _case(1 + channels, n, channels)
    (t, c, a), _, cc, aa = _torch_reference(image, tc, colors, alphas, tint_ratio, modulate)
    oc, oa = O.texture_forward(image, tc, colors, alphas, tint_ratio, modulate)
    np.testing.assert_allclose(oc, cc.detach().numpy(), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(oa, aa.detach().numpy(), rtol=1e-13, atol=1e-15)
    g = np.random.default_rng(7)
    gc, ga = g.standard_normal((n, 3)), g.standard_normal(n)
    want = torch.autograd.grad([cc, aa], [c, a, t], [torch.from_numpy(gc), torch.from_numpy(ga)], allow_unused=True)
    got = O.texture_backward(image, tc, colors, alphas, gc, ga, tint_ratio, modulate)
    for x, y in zip(got, want):
        np.testing.assert_allclose(x, y.numpy(), rtol=1e-11, atol=1e-12)


def test_closed_forms():
    h, w = 9, 5  # power-of-two spans: texel-centre texcoords are exact
    image = np.arange(3 * h * w, dtype=np.float64).reshape(3, h, w)
    ys, xs = np.meshgrid(np.arange(h), np.arange(w), indexing='ij')
    tc = np.stack([xs.ravel() / (w - 1), ys.ravel() / (h - 1)], -1)
    # texel centres return the texel
    np.testing.assert_array_equal(O.sample_texture(image, tc), image.reshape(3, -1).T)
    # the midpoint of four texels is their mean
    mid = np.array([[0.5 / (w - 1), 0.5 / (h - 1)]])
    np.testing.assert_allclose(O.sample_texture(image, mid)[0], image[:, :2, :2].mean((1, 2)))
    # a constant image is constant inside [0, 1]^2 and fades to zero one texel outside
    const = np.full((1, h, w), 0.7)
    inside = np.random.default_rng(3).uniform(0, 1, (50, 2))
    np.testing.assert_allclose(O.sample_texture(const, inside), 0.7)
    far = np.array([[-1.0 / (w - 1), 0.5], [0.5, 1 + 1.0 / (h - 1)], [2.0, 2.0], [np.nan, 0.5]])
    np.testing.assert_array_equal(O.sample_texture(const, far), 0.0)
    # tint_ratio 0 gives the tint itself; tint_ratio 1 gives colour x tint
    colors = np.random.default_rng(4).uniform(0, 1, (len(tc), 3))
    c0, _ = O.texture_forward(image, tc, colors, np.ones(len(tc)), 0.0, False)
    np.testing.assert_array_equal(c0, image.reshape(3, -1).T)
    c1, _ = O.texture_forward(image, tc, colors, np.ones(len(tc)), 1.0, False)
    np.testing.assert_allclose(c1, colors * image.reshape(3, -1).T)

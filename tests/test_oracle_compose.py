"""CPU pins of the compositing / ray-march oracle (oracle/compose.py).

Checks: the sequential "over" loop against the reversed exclusive-cumprod
closed form and brute-force products, single-stroke and all-zero special
cases, float32 constant semantics, the exact adjoint against central finite
differences, the reference adjoint's exact density part and its inexact
colour part (equal only for a single stroke), the ray march against the
telescoping identity acc = 1 - exp(-sum sigma delta), and the non-"over"
composition modes against their definitions.
"""
import numpy as np
import pytest
import torch

from oracle import compose as C

D64 = torch.float64


def _inputs(N=40, S=7, Dc=3, seed=0, zero_frac=0.3):
    g = torch.Generator().manual_seed(seed)
    a = torch.rand(N, S, generator=g, dtype=D64) * 0.95
    a = torch.where(torch.rand(N, S, generator=g, dtype=D64) < zero_frac, torch.zeros_like(a), a)
    c = torch.rand(N, S, Dc, generator=g, dtype=D64)
    d = torch.rand(S, generator=g, dtype=D64) * 10
    return a, c, d


def test_over_matches_closed_form_and_brute_force():
    a, c, d = _inputs()
    dens, col, raw, Tn = C.compose_forward(a, c, d)
    dens2, col2 = C.compose_closed_form(a, c, d)
    torch.testing.assert_close(dens, dens2, atol=1e-12, rtol=1e-12)
    torch.testing.assert_close(col, col2, atol=1e-12, rtol=1e-12)
    # brute force: weight of stroke k = alpha_k * prod over strokes in front (index > k) of (1 - alpha)
    N, S = a.shape
    w = torch.zeros_like(a)
    for k in range(S):
        front = torch.ones(N, dtype=D64)
        for j in range(k + 1, S):
            front = front * (1 - a[:, j])
        w[:, k] = a[:, k] * front
    torch.testing.assert_close(dens, (w * d).sum(1), atol=1e-12, rtol=1e-12)
    torch.testing.assert_close(raw, (w[..., None] * c).sum(1), atol=1e-12, rtol=1e-12)
    torch.testing.assert_close(Tn, torch.prod(1 - a, 1), atol=1e-12, rtol=1e-12)
    assert (w.sum(1) <= 1 + 1e-12).all()


def test_single_stroke_and_empty():
    a = torch.tensor([[0.5], [0.0], [0.9999]], dtype=D64)
    c = torch.tensor([[[0.2, 0.4, 0.6]]] * 3, dtype=D64)
    d = torch.tensor([3.0], dtype=D64)
    dens, col, _, Tn = C.compose_forward(a, c, d)
    torch.testing.assert_close(dens, 3.0 * a[:, 0])
    # C = c alpha / (1 + eps - (1 - alpha)): the stroke's own colour up to the epsilon
    expect = c[:, 0] * (a / (a + (C._ONE_PLUS_EPS - 1.0)))
    torch.testing.assert_close(col, expect)
    torch.testing.assert_close(col[0], c[0, 0], atol=3e-6, rtol=0)
    assert col[1].abs().max() == 0 and dens[1] == 0  # all-zero alpha: D = 0, C = 0
    assert Tn[1] == 1.0


def test_epsilon_is_float32_of_one_plus_1e6():
    assert C._ONE_PLUS_EPS == float(np.float32(1.0) + np.float32(1e-6))


def test_exact_adjoint_matches_finite_differences():
    a, c, d = _inputs(N=12, S=5, zero_frac=0.0, seed=1)
    a = a.clamp(0.05, 0.9)
    g = torch.Generator().manual_seed(2)
    gd = torch.randn(12, generator=g, dtype=D64)
    gc = torch.randn(12, 3, generator=g, dtype=D64) * 0.2
    c = c * 0.5  # keep the clamp to [0, 1] inactive

    def loss(av, cv, dv):
        dens, col, _, _ = C.compose_forward(av, cv, dv, skip_zero=False)
        return (dens * gd).sum() + (col * gc).sum()

    ga, gcol, gdp = C.compose_backward_exact(gd, gc, a, c, d)
    h = 1e-6
    for x, gx, idx in ((a, ga, 0), (c, gcol, 1), (d, gdp, 2)):
        fd = torch.zeros_like(x)
        for i in range(x.numel()):
            e = torch.zeros_like(x).view(-1)
            e[i] = h
            args_p = [a, c, d]
            args_m = [a, c, d]
            args_p[idx] = x + e.view_as(x)
            args_m[idx] = x - e.view_as(x)
            fd.view(-1)[i] = (loss(*args_p) - loss(*args_m)) / (2 * h)
        torch.testing.assert_close(gx, fd, atol=1e-7, rtol=1e-6)


def test_reference_adjoint_density_exact_colour_inexact():
    a, c, d = _inputs(N=30, S=5, zero_frac=0.0, seed=3)
    c = c * 0.5
    g = torch.Generator().manual_seed(4)
    gd = torch.randn(30, generator=g, dtype=D64)
    gc = torch.randn(30, 3, generator=g, dtype=D64)
    # density-only loss: the reference is the true derivative
    ra, rc, rd = C.compose_backward_reference(gd, None, a, c, d)
    ea, ec, ed = C.compose_backward_exact(gd, None, a, c, d)
    torch.testing.assert_close(ra, ea, atol=1e-12, rtol=1e-10)
    torch.testing.assert_close(rd, ed, atol=1e-12, rtol=1e-10)
    # colour loss: dL/dcolour and dL/ddensity exact, dL/dalpha not (several strokes)
    ra, rc, rd = C.compose_backward_reference(None, gc, a, c, d)
    ea, ec, ed = C.compose_backward_exact(None, gc, a, c, d)
    torch.testing.assert_close(rc, ec, atol=1e-12, rtol=1e-10)
    assert (ra - ea).norm() > 1e-2 * ea.norm()
    # ... but exact for a single stroke
    ra1, _, _ = C.compose_backward_reference(None, gc, a[:, :1], c[:, :1], d[:1])
    ea1, _, _ = C.compose_backward_exact(None, gc, a[:, :1], c[:, :1], d[:1])
    torch.testing.assert_close(ra1, ea1, atol=1e-10, rtol=1e-8)


def test_reference_adjoint_treats_empty_grads_as_zero():
    a, c, d = _inputs(N=8, S=4, seed=5)
    z = torch.empty(0, dtype=D64)
    ra, rc, rd = C.compose_backward_reference(z, z, a, c, d)
    assert ra.abs().max() == 0 and rc.abs().max() == 0 and rd.abs().max() == 0


def test_ray_march_identities():
    g = torch.Generator().manual_seed(6)
    R, n = 16, 32
    t = torch.sort(torch.rand(R, n + 1, generator=g, dtype=D64) * 4 + 2, -1).values
    dirs = torch.randn(R, 3, generator=g, dtype=D64)
    sigma = torch.rand(R, n, generator=g, dtype=D64) * 3
    w, alpha, trans = C.compute_alpha_weights(sigma, t, dirs)
    sd = sigma * (t[:, 1:] - t[:, :-1]) * dirs.norm(dim=-1, keepdim=True)
    # sequential transmittance
    T = torch.ones(R, dtype=D64)
    for i in range(n):
        torch.testing.assert_close(trans[:, i], T, atol=1e-13, rtol=1e-12)
        T = T * torch.exp(-sd[:, i])
    torch.testing.assert_close(w.sum(-1), 1 - torch.exp(-sd.sum(-1)), atol=1e-13, rtol=1e-12)  # telescoping sum
    assert (trans[:, 1:] <= trans[:, :-1]).all()
    col = torch.rand(R, n, 3, generator=g, dtype=D64)
    bg = torch.tensor([0.3, 0.5, 0.7], dtype=D64)
    rgb, acc, depth = C.volumetric_rendering(col, w, t, bg)
    torch.testing.assert_close(rgb, (w[..., None] * col).sum(1) + (1 - acc)[:, None] * bg)
    assert ((depth >= t[:, 0]) & (depth <= t[:, -1])).all()
    # constant colour c: rgb = acc c + (1 - acc) bg
    cc = torch.full_like(col, 0.25)
    rgb_c, acc_c, _ = C.volumetric_rendering(cc, w, t, bg)
    torch.testing.assert_close(rgb_c, acc_c[:, None] * 0.25 + (1 - acc_c)[:, None] * bg)
    # empty ray: 0 / eps = 0 is clipped to the near bound
    _, acc0, depth0 = C.volumetric_rendering(col, torch.zeros_like(w), t, bg)
    assert (acc0 == 0).all() and torch.equal(depth0, t[:, 0])


def test_cast_rays_midpoints():
    t = torch.tensor([[2.0, 3.0, 5.0]], dtype=D64)
    o = torch.tensor([[1.0, 0.0, 0.0]], dtype=D64)
    d = torch.tensor([[0.0, 1.0, 0.0]], dtype=D64)
    x, r, tm = C.cast_rays(t, o, d, torch.tensor([[0.5]], dtype=D64))
    torch.testing.assert_close(tm, torch.tensor([[2.5, 4.0]], dtype=D64))
    torch.testing.assert_close(x, torch.tensor([[[1.0, 2.5, 0.0], [1.0, 4.0, 0.0]]], dtype=D64))
    torch.testing.assert_close(r, torch.tensor([[1.25, 2.0]], dtype=D64))


@pytest.mark.parametrize('mode', ['max', 'max_density_weighted', 'softmax', 'softmax_density_weighted'])
def test_other_modes(mode):
    a, c, d = _inputs(N=20, S=6, seed=7, zero_frac=0.0)
    dens, col = C.compose_other(a, c, d, mode)
    logits = a * d if 'density' in mode else a
    if mode.startswith('max'):
        k = logits.argmax(1)
        r = torch.arange(20)
        torch.testing.assert_close(dens, a[r, k] * d[k])
        torch.testing.assert_close(col, c[r, k])
    else:
        w = torch.exp(logits / 0.05)
        w = w / w.sum(1, keepdim=True)
        torch.testing.assert_close(dens, (w * a * d).sum(1))
        torch.testing.assert_close(col, (w[..., None] * c).sum(1))

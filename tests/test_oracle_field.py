"""CPU pins of the whole-path oracle (oracle/field.py).

Checks: central finite differences of the float64 composition (stroke field ->
over compositing -> ray march) for shape, colour and density parameters;
invariance to inserting a stroke that cannot touch any sample; equality of
the reference-adjoint chain with the exact gradient where every reference
adjoint on the path is exact (density-only loss, ellipsoids and capsules); and
the committed golden fixture (tests/golden/field_small.npz).
"""
import os

import numpy as np
import pytest
import torch

from neural3dstrokes_amd import scenes
from oracle import field as F

D64 = torch.float64
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'field_small.npz')


def _scene(types, seed, n_rays=16, n_samples=32, radius=0.05):
    gen = torch.Generator().manual_seed(seed)
    st = scenes.make_strokes(types, gen, scale_range=(0.3, 0.6))
    rays = scenes.make_rays(n_rays, n_samples, gen, n_cams=4, radius=radius)
    st.update(rays)
    return {k: (v.double() if torch.is_tensor(v) and v.is_floating_point() else v) for k, v in st.items()}


def _args(sc):
    return (sc['origins'], sc['directions'], sc['viewdirs'], sc['radii'], sc['tdist'], sc['stroke_ids'])


@pytest.mark.parametrize('laplace', [True])
def test_exact_gradients_match_finite_differences(laplace):
    sc = _scene([0, 4, 0, 1, 4, 0], seed=3)
    g = torch.Generator().manual_seed(4)
    gr = torch.randn(16, 3, generator=g, dtype=D64)
    ga = torch.randn(16, generator=g, dtype=D64)
    kw = dict(laplace=laplace)
    out, gs, gc, gd = F.render_field_grads(gr, ga, *_args(sc), sc['shape_params'], sc['color_params'],
                                           sc['density_params'], **kw)
    assert (out['acc'] > 1e-3).sum() >= 4  # the rays see strokes

    def loss(sp, cp, dp):
        o = F.render_field(*_args(sc), sp, cp, dp, **kw)
        return (o['rgb'] * gr).sum() + (o['acc'] * ga).sum()

    h = 1e-6
    sel = torch.randperm(sc['shape_params'].numel(), generator=g)[:16]
    valid = torch.zeros_like(sc['shape_params'], dtype=torch.bool)
    for k, sid in enumerate(sc['stroke_ids'].tolist()):
        valid[k, :scenes.n_shape_params(sid)] = True
    for i in sel.tolist():
        if not valid.view(-1)[i]:
            continue
        e = torch.zeros_like(sc['shape_params']).view(-1)
        e[i] = h
        e = e.view_as(sc['shape_params'])
        fd = (loss(sc['shape_params'] + e, sc['color_params'], sc['density_params']) -
              loss(sc['shape_params'] - e, sc['color_params'], sc['density_params'])) / (2 * h)
        assert abs(fd - gs.view(-1)[i]) <= 1e-6 * (1 + abs(fd)), (i, float(fd), float(gs.view(-1)[i]))
    for k in range(6):
        e = torch.zeros_like(sc['density_params'])
        e[k] = h
        fd = (loss(sc['shape_params'], sc['color_params'], sc['density_params'] + e) -
              loss(sc['shape_params'], sc['color_params'], sc['density_params'] - e)) / (2 * h)
        assert abs(fd - gd[k]) <= 1e-6 * (1 + abs(fd))
        e = torch.zeros_like(sc['color_params'])
        e[k, 1] = h
        fd = (loss(sc['shape_params'], sc['color_params'] + e, sc['density_params']) -
              loss(sc['shape_params'], sc['color_params'] - e, sc['density_params'])) / (2 * h)
        assert abs(fd - gc[k, 1]) <= 1e-6 * (1 + abs(fd))


def test_far_stroke_changes_nothing():
    sc = _scene([0, 4, 0, 1], seed=5)
    base = F.render_field(*_args(sc), sc['shape_params'], sc['color_params'], sc['density_params'])
    sp = torch.cat([sc['shape_params'][:2], sc['shape_params'][:1], sc['shape_params'][2:]])
    sp[2, -3:] = 50.0  # translation far outside every ray segment
    ids = torch.cat([sc['stroke_ids'][:2], sc['stroke_ids'][:1], sc['stroke_ids'][2:]])
    cp = torch.cat([sc['color_params'][:2], sc['color_params'][:1], sc['color_params'][2:]])
    dp = torch.cat([sc['density_params'][:2], sc['density_params'][:1], sc['density_params'][2:]])
    args = (sc['origins'], sc['directions'], sc['viewdirs'], sc['radii'], sc['tdist'], ids)
    out = F.render_field(*args, sp, cp, dp)
    assert torch.equal(out['rgb'], base['rgb']) and torch.equal(out['acc'], base['acc'])
    g = torch.Generator().manual_seed(6)
    _, gs, gc, gd = F.render_field_grads(torch.randn(16, 3, generator=g, dtype=D64),
                                         torch.randn(16, generator=g, dtype=D64), *args, sp, cp, dp)
    assert gs[2].abs().max() == 0 and gc[2].abs().max() == 0 and gd[2] == 0


def test_reference_chain_exact_for_density_loss_on_ellipsoids_and_capsules():
    sc = _scene([0, 4, 0, 4, 0], seed=7)
    g = torch.Generator().manual_seed(8)
    ga = torch.randn(16, generator=g, dtype=D64)
    zero_rgb = torch.zeros(16, 3, dtype=D64)
    _, es, ec, ed = F.render_field_grads(zero_rgb, ga, *_args(sc), sc['shape_params'], sc['color_params'],
                                         sc['density_params'])
    rs, rc, rd = F.render_field_grads_reference(zero_rgb, ga, *_args(sc), sc['shape_params'], sc['color_params'],
                                                sc['density_params'])
    assert es.abs().max() > 0
    torch.testing.assert_close(rs, es, atol=1e-9, rtol=1e-8)
    torch.testing.assert_close(rd, ed, atol=1e-9, rtol=1e-8)


def test_golden_fixture_reproduces():
    """The committed float64 fixture (written by tests/golden/make_golden.py) matches the oracle."""
    z = np.load(GOLDEN)
    t = {k: torch.from_numpy(z[k]) for k in z.files}
    for case in ('solids50_4096x64', 'palette500_65536x128'):
        args = [t[f'{case}_{k}'].double() for k in ('origins', 'directions', 'viewdirs', 'radii', 'tdist')]
        out, gs, gc, gd = F.render_field_grads(t[f'{case}_g_rgb'], t[f'{case}_g_acc'], *args,
                                               t[f'{case}_stroke_ids'], t[f'{case}_shape_params'].double(),
                                               t[f'{case}_color_params'].double(),
                                               t[f'{case}_density_params'].double())
        torch.testing.assert_close(out['rgb'], t[f'{case}_rgb'], atol=1e-12, rtol=1e-10)
        torch.testing.assert_close(out['acc'], t[f'{case}_acc'], atol=1e-12, rtol=1e-10)
        torch.testing.assert_close(gs, t[f'{case}_grad_shape'], atol=1e-10, rtol=1e-8)
        torch.testing.assert_close(gc, t[f'{case}_grad_color'], atol=1e-10, rtol=1e-8)
        torch.testing.assert_close(gd, t[f'{case}_grad_density'], atol=1e-10, rtol=1e-8)

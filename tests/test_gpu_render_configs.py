"""Fused renderer on the settings the reference's configurations use, against the float64 oracle.

* ``inv_scale_radius=True`` (the ``StrokeField`` default, ``models.py:540``)
  with the Laplace transfer and 32 stroke-field samples per ray
  (``configs/*.gin``), single- and per-axis-scale strokes, cone radii and the
  coordinate warp, with a nonzero upstream depth gradient;
* spherical-harmonic colours (``constant_sh2`` / ``constant_sh3``);
* the bases outside the benchmark palette (capped torus, line, triangular
  prism, Catmull-Rom, capsule) through the fused kernels;
* per-ray random backgrounds (``models.py:213-222``, ``bg_intensity_range``
  of ``configs/clip.gin``) and ``opaque_background`` (``utils/render.py:35-40``,
  ``configs/llff.gin``).

The oracle's torch ops run in float64 on the same device, a few rays at a time.
"""
import pytest
import torch

from conftest import rel_err
from neural3dstrokes_amd import render, scenes
from oracle import field as ofield

D64 = torch.float64


def _scene(types, n_rays, n_samples, seed, color_id=0, singlescale=False, radius=0.01):
    gen = torch.Generator().manual_seed(seed)
    sc = scenes.make_strokes(types, gen, color_id=color_id, singlescale=singlescale)
    sc.update(scenes.make_rays(n_rays, n_samples, gen, radius=radius))
    return sc


def _check(sc, dev, seed, depth_grad=True, rgb_tol=1e-5, grad_tol=1e-4, fp32_bar=False, **kw):
    """Fused render + adjoint against the float64 oracle.  fp32_bar: a gradient may also be up to
    twice as far from float64 as the oracle's own arithmetic run in float32 (the reference's
    formulation in single precision) is."""
    g = torch.Generator().manual_seed(seed + 100)
    R = sc['tdist'].shape[0]
    g_rgb = torch.randn(R, 3, generator=g)
    g_acc = torch.randn(R, generator=g)
    g_dep = torch.randn(R, generator=g) if depth_grad else None
    f = lambda k: sc[k].to(dev)
    # depth = sum(w t) / acc: on a ray whose weights are only Laplace tails (acc ~ 1e-8, every
    # pair below cull_eps) it is a ratio of culled terms and its gradient grows as 1 / acc, so
    # depth and its upstream gradient are compared on rays with acc >= 1e-2 only
    with torch.no_grad():
        _, acc0, _ = render.render_strokes(f('origins'), f('directions'), f('viewdirs'), f('radii'), f('tdist'),
                                           sc['stroke_ids'], f('shape_params'), f('color_params'),
                                           f('density_params'), **{k: v for k, v in kw.items() if k != 'bg_color'})
    dmask = (acc0 >= 1e-2).cpu()
    if g_dep is not None:
        g_dep = g_dep * dmask
    sp = f('shape_params').requires_grad_(True)
    cp = f('color_params').requires_grad_(True)
    dp = f('density_params').requires_grad_(True)
    bg = kw.pop('bg_color', (1.0, 1.0, 1.0))
    rkw = dict(kw)
    okw = dict(kw)
    okw.pop('cull_eps', None)  # the oracle evaluates every pair
    for a, b in (('use_laplace_transform', 'laplace'),):
        if a in okw:
            okw[b] = okw.pop(a)
    rgb, acc, depth = render.render_strokes(f('origins'), f('directions'), f('viewdirs'), f('radii'), f('tdist'),
                                            sc['stroke_ids'], sp, cp, dp,
                                            bg_color=bg.to(dev) if torch.is_tensor(bg) else bg, **rkw)
    loss = (rgb * g_rgb.to(dev)).sum() + (acc * g_acc.to(dev)).sum()
    if g_dep is not None:
        loss = loss + (depth * g_dep.to(dev)).sum()
    loss.backward()
    args = tuple(sc[k].to(dev, D64) for k in ('origins', 'directions', 'viewdirs', 'radii', 'tdist'))
    out, rs, rc, rd = ofield.render_field_grads_chunked(
        g_rgb.to(dev, D64), g_acc.to(dev, D64), *args, sc['stroke_ids'], sc['shape_params'].to(dev, D64),
        sc['color_params'].to(dev, D64), sc['density_params'].to(dev, D64), chunk=32,
        grad_depth=None if g_dep is None else g_dep.to(dev, D64),
        bg_color=bg.to(dev, D64) if torch.is_tensor(bg) else bg, **okw)
    tol = dict(shape=grad_tol, color=grad_tol, density=grad_tol)
    if fp32_bar:
        args32 = tuple(sc[k].to(dev) for k in ('origins', 'directions', 'viewdirs', 'radii', 'tdist'))
        _, s32, c32, d32 = ofield.render_field_grads_chunked(
            g_rgb.to(dev), g_acc.to(dev), *args32, sc['stroke_ids'], sc['shape_params'].to(dev),
            sc['color_params'].to(dev), sc['density_params'].to(dev), chunk=32,
            grad_depth=None if g_dep is None else g_dep.to(dev),
            bg_color=bg.to(dev) if torch.is_tensor(bg) else bg, **okw)
        for k, a in (('shape', s32), ('color', c32), ('density', d32)):
            tol[k] = max(grad_tol, 2.0 * rel_err(a, {'shape': rs, 'color': rc, 'density': rd}[k]))
        print('float32 formulation vs float64:', tol)
    dm = dmask.to(dev)
    errs = dict(rgb=rel_err(rgb, out['rgb']), acc=rel_err(acc, out['acc']),
                depth=rel_err(depth[dm], out['depth'][dm]),
                shape=rel_err(sp.grad, rs), color=rel_err(cp.grad, rc), density=rel_err(dp.grad, rd))
    print(kw, errs)
    assert errs['rgb'] < rgb_tol and errs['acc'] < rgb_tol and errs['depth'] < rgb_tol, errs
    assert errs['shape'] < tol['shape'] and errs['color'] < tol['color'] and errs['density'] < tol['density'], \
        (errs, tol)
    assert float(out['acc'].max()) > 0.3  # the rays do cross strokes
    assert int(dmask.sum()) >= R // 8
    return errs


MIXED = [10, 9, 0, 1, 2, 7, 8, 4]


@pytest.mark.gpu
@pytest.mark.parametrize('singlescale', [False, True])
def test_inv_scale_radius_laplace_32_samples(singlescale, dev):
    """StrokeField's defaults: inverse-scaled radius, Laplace transfer, 32 samples per ray, cone radii,
    the warp's coordinate divisor and a training-time sdf_delta."""
    types = [MIXED[i % len(MIXED)] for i in range(96)]
    sc = _scene(types, 96, 32, seed=31 + singlescale, singlescale=singlescale, radius=0.0015)
    _check(sc, dev, 31, use_laplace_transform=True, inv_scale_radius=True, cone=True, coord_div=1.25,
           sdf_delta=3.0)


@pytest.mark.gpu
def test_linear_transfer_inv_scale_radius(dev):
    types = [MIXED[i % len(MIXED)] for i in range(80)]
    sc = _scene(types, 64, 64, seed=35, radius=0.004)
    _check(sc, dev, 35, inv_scale_radius=True, cone=True)


@pytest.mark.gpu
@pytest.mark.parametrize('color_id', [2, 3])
def test_spherical_harmonic_colours(color_id, dev):
    types = [MIXED[i % len(MIXED)] for i in range(72)]
    sc = _scene(types, 64, 64, seed=40 + color_id, color_id=color_id)
    _check(sc, dev, 40 + color_id, color_id=color_id)


@pytest.mark.gpu
@pytest.mark.parametrize('laplace', [False, True])
def test_bases_outside_the_benchmark_palette(laplace, dev):
    """Capped torus, line, triangular prism, Catmull-Rom and capsule strokes."""
    types = [[3, 5, 6, 11, 4][i % 5] for i in range(60)]
    sc = _scene(types, 64, 64, seed=50 + laplace, radius=0.006)
    _check(sc, dev, 50, use_laplace_transform=laplace, sdf_delta=1.5 if laplace else 1.0)


@pytest.mark.gpu
def test_per_ray_random_background(dev):
    types = [MIXED[i % len(MIXED)] for i in range(64)]
    sc = _scene(types, 64, 64, seed=60)
    g = torch.Generator().manual_seed(61)
    bg = torch.rand(64, 3, generator=g)  # bg_intensity_range (0, 1)
    _check(sc, dev, 60, bg_color=bg)


@pytest.mark.gpu
@pytest.mark.parametrize('laplace', [False, True])
def test_opaque_background(laplace, dev):
    """The last interval infinitely wide.  With the Laplace transfer a ray that misses every
    stroke gives its last sample weight ~1, and that sample's normalised colour
    C / (1e-6 + 1 - T) is a ratio of Laplace tails amplified ~1e6: the tail cull is lowered
    to 1e-20 so the culled pairs stay far below the 1e-6 floor, and the shape gradient,
    whose tail terms carry the float32 rounding of the sample positions (relative ~5e-5 in
    alpha at the cull boundary), is held to twice what the same formulation reaches in float32
    (the oracle in float32: 1.6e-4 with the device's math, 4.2e-4 with the host's; the kernels
    2.0e-4)."""
    types = [MIXED[i % len(MIXED)] for i in range(64)]
    sc = _scene(types, 64, 48, seed=70 + laplace)
    kw = dict(cull_eps=1e-20, fp32_bar=True) if laplace else {}
    errs = _check(sc, dev, 70, opaque_background=True, use_laplace_transform=laplace,
                  bg_color=(0.2, 0.5, 0.9), **kw)
    assert errs['acc'] < 1e-5


@pytest.mark.gpu
def test_opaque_background_sums_weights_to_one(dev):
    """With the last interval infinitely wide every ray's weights sum to one (acc = 1)."""
    types = [MIXED[i % len(MIXED)] for i in range(32)]
    sc = _scene(types, 128, 32, seed=80)
    f = lambda k: sc[k].to(dev)
    _, acc, _ = render.render_strokes(f('origins'), f('directions'), f('viewdirs'), f('radii'), f('tdist'),
                                      sc['stroke_ids'], f('shape_params'), f('color_params'), f('density_params'),
                                      opaque_background=True)
    assert torch.allclose(acc, torch.ones_like(acc), atol=2e-6)


@pytest.mark.gpu
def test_device_stroke_ids_are_checked(dev):
    """An undecodable id given on the device, or a stroke wider than shape_params, raises."""
    types = [0, 1, 10]
    sc = _scene(types, 8, 16, seed=90)
    f = lambda k: sc[k].to(dev)
    bad = sc['stroke_ids'].clone()
    bad[1] = 15 << 4  # no such base
    with pytest.raises(RuntimeError, match='invalid stroke id'):
        render.render_strokes(f('origins'), f('directions'), f('viewdirs'), f('radii'), f('tdist'), bad.to(dev),
                              f('shape_params'), f('color_params'), f('density_params'))
    narrow = f('shape_params')[:, :9].contiguous()  # the cubic Bezier needs 14 columns
    with pytest.raises(RuntimeError, match='invalid stroke id'):
        render.render_strokes(f('origins'), f('directions'), f('viewdirs'), f('radii'), f('tdist'),
                              sc['stroke_ids'].to(dev), narrow, f('color_params'), f('density_params'))
    with pytest.raises(RuntimeError, match='shape parameters'):
        render.render_strokes(f('origins'), f('directions'), f('viewdirs'), f('radii'), f('tdist'),
                              sc['stroke_ids'], narrow, f('color_params'), f('density_params'))

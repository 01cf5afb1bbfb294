"""Eval-time rendering with the reference's extras (volumetric_rendering with compute_extras,
``utils/render.py:88-116``) against the float64 oracle."""
import pytest
import torch

from conftest import rel_err
from neural3dstrokes_amd import render, scenes
from oracle import compose as ocompose
from oracle import field as ofield

D64 = torch.float64
KEYS = ('origins', 'directions', 'viewdirs', 'radii', 'tdist')
EXTRAS = ('distance_mean', 'distance_percentile_5', 'distance_median', 'distance_percentile_95')


@pytest.mark.gpu
@pytest.mark.parametrize('opaque', [False, True])
def test_render_rays_extras_against_oracle(opaque, dev):
    gen = torch.Generator().manual_seed(21 + opaque)
    types = [[10, 9, 0, 1, 2, 7, 8, 4][i % 8] for i in range(80)]
    sc = scenes.make_strokes(types, gen)
    sc.update(scenes.make_rays(200, 64, gen))
    far = sc['tdist'][:, -1] + 0.25
    f = lambda k: sc[k].to(dev)  # noqa: E731
    args = [f(k) for k in KEYS] + [sc['stroke_ids'], f('shape_params'), f('color_params'), f('density_params')]
    out = render.render_rays(*args, t_far=far.to(dev), compute_extras=True, opaque_background=opaque)
    rgb, acc, depth = render.render_strokes(*args, opaque_background=opaque)
    assert torch.equal(out['rgb'], rgb) and torch.equal(out['acc'], acc) and torch.equal(out['depth'], depth)
    ref = ofield.render_field(*[sc[k].to(dev, D64) for k in KEYS], sc['stroke_ids'], f('shape_params').double(),
                              f('color_params').double(), f('density_params').double(), opaque_background=opaque)
    ex = ocompose.volumetric_extras(ref['weights'], sc['tdist'].to(dev, D64), far.to(dev, D64))
    if not opaque:  # rays that hit strokes and rays that mostly miss them (opaque: every acc is 1)
        assert float(ref['acc'].max()) > 0.5 and float(ref['acc'].min()) < 0.5
    for k in EXTRAS:
        assert rel_err(out[k], ex[k]) < 1e-5, (k, rel_err(out[k], ex[k]))
    # every percentile lies between the first fencepost and the far distance, in order
    assert bool((out['distance_percentile_5'] <= out['distance_median'] + 1e-6).all())
    assert bool((out['distance_median'] <= out['distance_percentile_95'] + 1e-6).all())
    assert bool((out['distance_percentile_95'] <= far.to(dev) + 1e-5).all())


@pytest.mark.gpu
def test_render_rays_needs_the_far_distance(dev):
    gen = torch.Generator().manual_seed(3)
    sc = scenes.make_strokes([0, 1], gen)
    sc.update(scenes.make_rays(8, 16, gen))
    f = lambda k: sc[k].to(dev)  # noqa: E731
    args = [f(k) for k in KEYS] + [sc['stroke_ids'], f('shape_params'), f('color_params'), f('density_params')]
    with pytest.raises(ValueError):
        render.render_rays(*args, compute_extras=True)
    out = render.render_rays(*args, t_far=7.0, compute_extras=True)
    assert out['distance_median'].shape == (8,)

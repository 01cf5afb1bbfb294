"""The two large workloads on the GPU.

* 32 independent scenes (stroke counts log-uniform in [20, 1000], ray counts
  log-uniform in [1024, 65536], 128 samples) rendered as one packed batch:
  every scene's outputs and gradients against the float64 oracle of that
  scene alone, on a subset of its rays (the upstream gradient is zero on the
  other rays, so the batch gradient of the scene is the subset's gradient).
* 1000 strokes x 800x800 rays x 192 samples: a ray subset against the float64
  oracle with kept activations and with recomputation, and the full image
  with both backward modes agreeing (outputs bit for bit, gradients 1e-5).

The oracle's torch ops run in float64 on the same device, a few rays at a time.
"""
import pytest
import torch

from conftest import rel_err
from neural3dstrokes_amd import render, scenes
from oracle import field as ofield

D64 = torch.float64
KEYS = ('origins', 'directions', 'viewdirs', 'radii', 'tdist')


def _oracle(sc, dev, g_rgb, g_acc, rows=slice(None)):
    args = tuple(sc[k][rows].to(dev, D64) for k in KEYS)
    return ofield.render_field_grads_chunked(g_rgb.to(dev, D64), g_acc.to(dev, D64), *args, sc['stroke_ids'],
                                             sc['shape_params'].to(dev, D64), sc['color_params'].to(dev, D64),
                                             sc['density_params'].to(dev, D64), chunk=16)


@pytest.mark.gpu
def test_multiscene_32_scenes_against_oracle(dev):
    from neural3dstrokes_amd.multiscene import SceneBatch, pack
    sc, rc, kc = scenes.make_multiscene(n_scenes=32, seed=0)
    packed = {k: v.to(dev) for k, v in pack(sc).items()}
    batch = SceneBatch(rc, kc, 128, 14, 3, 0, device=dev)
    print(f'32 scenes: {sum(rc)} rays, {sum(kc)} strokes, strokes {min(kc)}..{max(kc)}, rays {min(rc)}..{max(rc)}',
          flush=True)
    n_sub = 24
    g = torch.Generator().manual_seed(17)
    g_rgb = torch.zeros(sum(rc), 3)
    g_acc = torch.zeros(sum(rc))
    for i in range(32):
        rs = batch.rays(i)
        g_rgb[rs.start:rs.start + n_sub] = torch.randn(n_sub, 3, generator=g)
        g_acc[rs.start:rs.start + n_sub] = torch.randn(n_sub, generator=g)
    leaves = [packed[k].clone().requires_grad_(True) for k in ('shape_params', 'color_params', 'density_params')]
    rgb, acc, _ = render.render_strokes(*[packed[k] for k in KEYS], packed['stroke_ids'], *leaves, scenes=batch)
    ((rgb * g_rgb.to(dev)).sum() + (acc * g_acc.to(dev)).sum()).backward()
    worst = {}
    for i in range(32):
        rs, ks = batch.rays(i), batch.strokes(i)
        sub = slice(rs.start, rs.start + n_sub)
        out, gs, gc, gd = _oracle(sc[i], dev, g_rgb[sub], g_acc[sub], rows=slice(0, n_sub))
        S = ks.stop - ks.start
        P = sc[i]['shape_params'].shape[1]
        e = dict(rgb=rel_err(rgb[sub], out['rgb']), acc=rel_err(acc[sub], out['acc']),
                 shape=rel_err(leaves[0].grad[ks][:, :P], gs), color=rel_err(leaves[1].grad[ks], gc),
                 density=rel_err(leaves[2].grad[ks], gd))
        for k, v in e.items():
            worst[k] = max(worst.get(k, 0.0), v)
        assert e['rgb'] < 1e-5 and e['acc'] < 1e-5, (i, S, e)
        assert e['shape'] < 1e-4 and e['color'] < 1e-4 and e['density'] < 1e-4, (i, S, e)
    print('worst per-scene rel errors vs float64 oracle:', worst)


@pytest.fixture(scope='module')
def palette1000():
    return scenes.make_workload('palette1000_640000x192', seed=0)


@pytest.mark.gpu
@pytest.mark.parametrize('store', [True, False])
def test_palette1000_ray_subset_against_oracle(palette1000, store, dev):
    sc = palette1000
    rows = torch.randperm(sc['tdist'].shape[0], generator=torch.Generator().manual_seed(3))[:160].sort().values
    sub = {k: sc[k][rows] for k in KEYS}
    sub.update({k: sc[k] for k in ('stroke_ids', 'shape_params', 'color_params', 'density_params')})
    g = torch.Generator().manual_seed(4)
    g_rgb, g_acc = torch.randn(160, 3, generator=g), torch.randn(160, generator=g)
    f = lambda k: sub[k].to(dev)
    leaves = [f(k).requires_grad_(True) for k in ('shape_params', 'color_params', 'density_params')]
    rgb, acc, depth = render.render_strokes(*[f(k) for k in KEYS], sub['stroke_ids'], *leaves, store=store)
    ((rgb * g_rgb.to(dev)).sum() + (acc * g_acc.to(dev)).sum()).backward()
    out, gs, gc, gd = _oracle(sub, dev, g_rgb, g_acc)
    e = dict(rgb=rel_err(rgb, out['rgb']), acc=rel_err(acc, out['acc']), depth=rel_err(depth, out['depth']),
             shape=rel_err(leaves[0].grad, gs), color=rel_err(leaves[1].grad, gc),
             density=rel_err(leaves[2].grad, gd))
    print('1000 strokes x 192 samples, 160 rays, store', store, e)
    assert e['rgb'] < 1e-5 and e['acc'] < 1e-5 and e['depth'] < 1e-5, e
    assert e['shape'] < 1e-4 and e['color'] < 1e-4 and e['density'] < 1e-4, e
    assert float(out['acc'].max()) > 0.5


@pytest.mark.gpu
def test_palette1000_full_image_kept_equals_recompute(palette1000, dev):
    """All 640,000 rays: kept activations and recomputation give the same outputs bit for bit
    and the same gradients up to float summation order."""
    sc = palette1000
    f = lambda k: sc[k].to(dev)
    g = torch.Generator().manual_seed(5)
    R = sc['tdist'].shape[0]
    g_rgb, g_acc = torch.randn(R, 3, generator=g).to(dev), torch.randn(R, generator=g).to(dev)
    res = []
    for store in (True, False):
        leaves = [f(k).requires_grad_(True) for k in ('shape_params', 'color_params', 'density_params')]
        out = render.render_strokes(*[f(k) for k in KEYS], sc['stroke_ids'], *leaves, store=store)
        ((out[0] * g_rgb).sum() + (out[1] * g_acc).sum()).backward()
        torch.cuda.synchronize()
        res.append(([o.detach() for o in out], [l.grad for l in leaves]))
        print('640000 rays x 192 samples x 1000 strokes, store', store, 'done', flush=True)
    (o0, g0), (o1, g1) = res
    for a, b in zip(o0, o1):
        assert torch.equal(a, b)
    for a, b in zip(g0, g1):
        assert rel_err(a, b) < 1e-5, rel_err(a, b)
    assert float(o0[1].max()) > 0.5

"""Fused proposal density field (n3ds_prop_field_forward / _backward + the grid scatter) against
the float64 oracle, and the drop-in module against the plain torch formulation."""
import numpy as np
import pytest
import torch

from conftest import rel_err
from neural3dstrokes_amd import propfield
from oracle import propfield as oprop

D64 = torch.float64


def _module(dev, seed, **kw):
    torch.manual_seed(seed)
    m = propfield.PropDensityField(**kw).to(dev)
    with torch.no_grad():  # a trained-looking table rather than the U(-1e-4, 1e-4) initialisation
        m.encoder.embeddings.uniform_(-0.3, 0.3)
    return m


@pytest.mark.gpu
@pytest.mark.parametrize('kw,warp', [
    (dict(grid_log2_hashmap_size=16, grid_disired_resolution=1024), 'contract'),  # 7 levels x 1
    (dict(grid_level_dim=2, grid_log2_hashmap_size=15, grid_disired_resolution=512), 'contract'),
    (dict(grid_log2_hashmap_size=21), 'contract'),                               # PropMLP defaults, 10 levels
    (dict(grid_level_dim=4, grid_log2_hashmap_size=14, grid_disired_resolution=256), None),
    (dict(grid_log2_hashmap_size=14, grid_disired_resolution=128, grid_level_interval=4), 'none')])
@pytest.mark.parametrize('n', [3000, 1])
def test_prop_field_against_oracle(kw, warp, n, dev):
    m = _module(dev, 0, warp_fn=None if warp == 'none' else warp, **kw)
    g = torch.Generator().manual_seed(1)
    coords = torch.randn(n, 3, generator=g) * (0.6 if warp == 'none' else 2.5)
    gd = torch.randn(n, generator=g)
    out = m(coords.to(dev), no_warp=warp == 'none')
    out['density'].backward(gd.to(dev))
    enc = m.encoder
    lin0, lin2 = m.density_layer[0], m.density_layer[2]
    args = (coords.double(), enc.embeddings.detach().double().cpu().numpy(), enc.offsets.cpu().numpy(),
            enc.per_level_scale, enc.base_resolution, lin0.weight.detach().double().cpu(),
            lin0.bias.detach().double().cpu(), lin2.weight.detach().double().cpu(), lin2.bias.detach().double().cpu(),
            -1.0, gd.double())
    wkw = dict(warp_fn=None if warp == 'none' else warp, no_warp=warp == 'none')
    # the warp on its own against float64
    coord_err = rel_err(out['coord'], oprop.warp_coords(coords.double(), **wkw))
    # the rest from the kernel's warped coordinates with float32 lattice positions, as the reference's
    # float32 grid kernel computes them (at resolution 8192 a float32 position has an ulp of ~5e-4 cell)
    ref = oprop.density_and_grads(*args, **wkw, warped=out['coord'].detach().double().cpu(), f32_grid=True)
    errs = dict(density=rel_err(out['density'], ref['density']), coord=coord_err,
                w1=rel_err(lin0.weight.grad, ref['grad_w1']), b1=rel_err(lin0.bias.grad, ref['grad_b1']),
                w2=rel_err(lin2.weight.grad, ref['grad_w2']), b2=rel_err(lin2.bias.grad, ref['grad_b2']),
                emb=rel_err(enc.embeddings.grad, ref['grad_embeddings']))
    print(kw, warp, n, errs)
    assert errs['density'] < 1e-5 and errs['coord'] < 1e-6, errs
    assert max(errs[k] for k in ('w1', 'b1', 'w2', 'b2', 'emb')) < 1e-4, errs


@pytest.mark.gpu
def test_prop_field_fp16_table(dev):
    """float16 table: features interpolated in float32; error against float64 stated."""
    m = _module(dev, 2, grid_log2_hashmap_size=16, grid_disired_resolution=1024)
    m.encoder.embeddings.data = m.encoder.embeddings.data.half()
    g = torch.Generator().manual_seed(3)
    coords = torch.randn(2000, 3, generator=g) * 2
    out = m(coords.to(dev))
    enc = m.encoder
    lin0, lin2 = m.density_layer[0], m.density_layer[2]
    ref = oprop.density_and_grads(coords.double(), enc.embeddings.detach().double().cpu().numpy(),
                                  enc.offsets.cpu().numpy(), enc.per_level_scale, enc.base_resolution,
                                  lin0.weight.detach().double().cpu(), lin0.bias.detach().double().cpu(),
                                  lin2.weight.detach().double().cpu(), lin2.bias.detach().double().cpu(), -1.0,
                                  torch.ones(2000, dtype=D64))
    assert rel_err(out['density'], ref['density']) < 1e-5


@pytest.mark.gpu
def test_prop_field_module_matches_plain_torch_and_loads_reference_keys(dev):
    m = _module(dev, 4, grid_log2_hashmap_size=16, grid_disired_resolution=1024)
    keys = set(m.state_dict())
    assert {'encoder.embeddings', 'encoder.offsets', 'density_layer.0.weight', 'density_layer.0.bias',
            'density_layer.2.weight', 'density_layer.2.bias'} <= keys
    g = torch.Generator().manual_seed(5)
    coords = (torch.randn(4096, 3, generator=g) * 2).to(dev)
    fused = m(coords)['density']
    plain = propfield.reference_forward(m, coords)
    assert rel_err(fused, plain) < 1e-5
    m.train()
    stats = m(coords)['hash_levelwise_mean']
    assert stats.shape == (m.encoder.num_levels, 1)

"""Host-side checks that need no GPU.

* the C-ABI library loads and exports every symbol declared in include/n3ds.h,
  and the ctypes signature table covers exactly those symbols (only the two
  size/version queries, which touch no device, are called);
* the drop-in operator surface (stroke registry, samplers, adjoint switch,
  grid-encoder bookkeeping) matches the reference's tables;
* the product path refuses CPU tensors (there is no CPU fallback);
* the seeded workload generator is deterministic and draws the stated
  distributions (Euler angles checked against scipy's quaternion rotation);
* bench.py refuses multi-GPU runs.
"""
import ctypes
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from neural3dstrokes_amd import _lib, scenes
from neural3dstrokes_amd.gridencoder import GridEncoder
from neural3dstrokes_amd.strokelib import strokes as dstrokes
from oracle import grid as ogrid
from oracle import strokes as ostrokes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'n3ds.h')


def _header_symbols():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(n3ds_\w+)\s*\(', text)))


def test_header_and_signature_table_agree():
    syms = _header_symbols()
    assert len(syms) == 36
    assert syms == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('libn3ds.so is not built (run __graft_entry__.build())')
    handle = ctypes.CDLL(_lib.LIB_PATH)
    for name in _header_symbols():
        assert hasattr(handle, name), name
    lib = _lib.lib()
    assert lib.n3ds_version() == 1
    assert lib.n3ds_render_table_floats(10) == 1280  # 128 floats per stroke record
    assert isinstance(lib.n3ds_last_error(), bytes)  # message of the last failed call, if any


@pytest.mark.parametrize('shape', sorted(ostrokes.SHAPE_TYPES))
@pytest.mark.parametrize('color', sorted(ostrokes.COLOR_TYPES))
def test_get_stroke_registry(shape, color):
    fn, dim_s, dim_c, s_ranges, c_ranges, ssamp, csamp = dstrokes.get_stroke(shape, color, 'recon')
    assert dim_s == ostrokes.num_shape_params(ostrokes.shape_sdf_id(shape)) == len(s_ranges)
    assert dim_c == ostrokes.NUM_COLOR_PARAMS[ostrokes.COLOR_TYPES[color]] == len(c_ranges)
    assert all(lo is None or hi is None or lo < hi for lo, hi in s_ranges + c_ranges)  # None: unbounded
    torch.manual_seed(0)
    assert ssamp(100).shape == (dim_s,)
    assert csamp().shape == (dim_c,)


def test_adjoint_switch():
    assert dstrokes.get_adjoint() == 'exact'
    with pytest.raises(ValueError):
        dstrokes.set_adjoint('fast')
    dstrokes.set_adjoint('reference')
    try:
        assert dstrokes.get_adjoint() == 'reference'
    finally:
        dstrokes.set_adjoint('exact')


def test_grid_encoder_bookkeeping_matches_reference_layout():
    enc = GridEncoder(input_dim=3, num_levels=16, level_dim=2, base_resolution=16, log2_hashmap_size=19,
                      desired_resolution=2048)
    offsets, res, pls = ogrid.grid_layout(3, 16, 2, base_resolution=16, log2_hashmap_size=19, desired_resolution=2048)
    np.testing.assert_array_equal(enc.offsets.cpu().numpy(), offsets)
    assert enc.embeddings.shape == (6_119_864, 2)
    assert abs(enc.per_level_scale - pls) < 1e-12


def test_product_path_rejects_cpu_tensors():
    from neural3dstrokes_amd import render
    sc = scenes.make_workload('solids50_4096x64', seed=0, n_rays=4, n_samples=8, n_strokes=3)
    with pytest.raises(RuntimeError, match='CUDA'):
        render.render_strokes(sc['origins'], sc['directions'], sc['viewdirs'], sc['radii'], sc['tdist'],
                              sc['stroke_ids'], sc['shape_params'], sc['color_params'], sc['density_params'])
    fn = dstrokes.get_stroke('ellipsoid', 'constant_rgb', 'recon')[0]
    sp = torch.rand(2, 9)
    with pytest.raises(RuntimeError, match='CUDA'):
        fn(torch.rand(4, 2, 3), torch.rand(4, 2), torch.rand(4, 3), sp, torch.rand(2, 3), 1.0)
    with pytest.raises(RuntimeError, match='CUDA'):
        dstrokes.compose_strokes(torch.rand(8, 2), torch.rand(8, 2, 3), torch.rand(2), 'over')


def test_workload_is_deterministic_and_has_the_stated_palette():
    a = scenes.make_workload('palette500_65536x128', seed=3, n_rays=64, n_samples=16)
    b = scenes.make_workload('palette500_65536x128', seed=3, n_rays=64, n_samples=16)
    for k in ('shape_params', 'color_params', 'density_params', 'origins', 'directions', 'tdist'):
        assert torch.equal(a[k], b[k]), k
    types = a['types']
    assert types.count(10) == 200 and types.count(9) == 100
    assert sorted(set(t for t in types if t < 9)) == [0, 1, 2, 7, 8]
    assert ((a['density_params'] >= 1) & (a['density_params'] <= 100)).all()
    # rays start on the radius-4 sphere, unit directions, sorted fenceposts in [near, far]
    torch.testing.assert_close(a['origins'].norm(dim=-1), torch.full((64,), 4.0))
    torch.testing.assert_close(a['directions'].norm(dim=-1), torch.ones(64))
    t = a['tdist']
    assert (t[:, 1:] >= t[:, :-1]).all() and t.min() >= 2.0 and t.max() <= 6.0
    # every ray points into the 40-degree cone around the camera axis
    cosang = -(a['origins'] * a['directions']).sum(-1) / 4.0
    assert cosang.min() >= math.cos(math.radians(20)) - 1e-6
    solids = scenes.make_workload('solids50_4096x64', seed=0)
    assert set(solids['types']) <= {0, 1, 4} and len(solids['types']) == 50


def test_random_euler_is_the_quaternion_rotation():
    g1 = torch.Generator().manual_seed(9)
    eul = scenes.random_euler(200, g1)
    g2 = torch.Generator().manual_seed(9)
    u = torch.rand(3, 200, generator=g2, dtype=torch.float64).numpy()
    q = np.stack([np.sqrt(1 - u[0]) * np.sin(2 * np.pi * u[1]), np.sqrt(1 - u[0]) * np.cos(2 * np.pi * u[1]),
                  np.sqrt(u[0]) * np.sin(2 * np.pi * u[2]), np.sqrt(u[0]) * np.cos(2 * np.pi * u[2])], -1)
    np.testing.assert_allclose(ostrokes.rotation_matrix(eul).numpy(), Rotation.from_quat(q).as_matrix(), atol=1e-12)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert 'one GPU' in r.stdout


def test_flop_model_covers_the_palettes():
    sys.path.insert(0, ROOT)
    import bench
    for b in (0, 1, 2, 4, 7, 8, 9, 10):
        fwd, tot = bench.FLOPS[b]
        assert 0 < fwd < tot


def test_render_tuning_parameters():
    """n3ds_render_tuning reads / sets the launch-shape parameters and rejects bad ones (no device needed)."""
    from neural3dstrokes_amd import autotune
    old = autotune.current()
    assert old['flush_pairs'] == 128 and old['bwd_oversub'] >= 1 and old['batch_oversub'] >= 1
    assert autotune.set('flush_pairs', 192) == 128
    assert autotune.get('flush_pairs') == 192
    lib = _lib.lib()
    assert lib.n3ds_render_tuning(0, 5000) == -1 and b'outside' in lib.n3ds_last_error()
    assert lib.n3ds_render_tuning(8, 1) == -1
    assert old['bwd_queue'] == 1 and lib.n3ds_render_tuning(4, 2) == -1
    assert old['cull_grid'] == 1 and lib.n3ds_render_tuning(5, 2) == -1
    assert old['bwd_split'] == 1 and old['split_cap'] == 32 and lib.n3ds_render_tuning(7, 0) == -1
    assert autotune.get('flush_pairs') == 192
    for k, v in old.items():
        autotune.set(k, v)
    assert autotune.current() == old

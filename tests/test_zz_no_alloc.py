"""Allocation-free render backward (n3ds_render_backward_noalloc) against n3ds_render_backward.

Both entry points are called through the C-ABI on the same forward state; the
allocation-free one keeps its partial gradient rows in the library's fixed
device scratch.  Two calls in a row, the second with fewer rays and strokes,
check that nothing left in the scratch by one call leaks into the next.
"""
import ctypes

import pytest
import torch

from conftest import rel_err
from neural3dstrokes_amd import _lib, render, scenes


def _forward(sc, dev):
    """Forward through the C-ABI with no kept activations; returns the backward's arguments."""
    L = _lib.lib()
    f = lambda k: sc[k].to(dev).float().contiguous()
    o, d, vd, rad, td = f('origins'), f('directions'), f('viewdirs'), f('radii'), f('tdist')
    sp, cp, dp = f('shape_params'), f('color_params'), f('density_params')
    ids = sc['stroke_ids'].to(dev, torch.int32).contiguous()
    R, n1 = td.shape
    S, P = sp.shape
    Q = cp.shape[1]
    table, status = render.prepare_table(ids, sp)
    assert int(status.item()) == 0
    bg = (ctypes.c_float * 3)(1.0, 1.0, 1.0)
    rgb = torch.empty(R, 3, device=dev)
    acc = torch.empty(R, device=dev)
    depth = torch.empty(R, device=dev)
    state = torch.empty(5, R * (n1 - 1), device=dev)
    _lib.check(L.n3ds_render_forward(
        _lib.ptr(o), _lib.ptr(d), _lib.ptr(vd), _lib.ptr(rad), _lib.ptr(td), _lib.ptr(table), _lib.ptr(cp),
        _lib.ptr(dp), R, n1 - 1, S, P, Q, 0, 1.0, 0, 0, 0, 1.0, ctypes.cast(bg, ctypes.c_void_p), None, 0, 1e-7,
        _lib.ptr(rgb), _lib.ptr(acc), _lib.ptr(depth), _lib.ptr(state), None, None, None, None, 0, 0, None, 0,
        _lib.stream()))
    g = torch.Generator().manual_seed(S)
    g_rgb = torch.randn(R, 3, generator=g).to(dev)
    g_acc = torch.randn(R, generator=g).to(dev)
    common = (_lib.ptr(o), _lib.ptr(d), _lib.ptr(vd), _lib.ptr(rad), _lib.ptr(td), _lib.ptr(table), _lib.ptr(cp),
              _lib.ptr(dp), R, n1 - 1, S, P, Q, 0, 1.0, 0, 0, 0, 1.0, ctypes.cast(bg, ctypes.c_void_p), None, 0, 1e-7, 1,
              _lib.ptr(g_rgb), _lib.ptr(g_acc), None, _lib.ptr(state), None, None, None, 0, 0)
    keep = (o, d, vd, rad, td, sp, cp, dp, table, state, g_rgb, g_acc, bg)
    return common, (S, P, Q, R), keep


def _grads(dev, S, P, Q):
    return (torch.zeros(S, P, device=dev), torch.zeros(S, Q, device=dev), torch.zeros(S, device=dev))


def _with_workspace(common, dims, dev):
    S, P, Q, R = dims
    L = _lib.lib()
    nws = int(L.n3ds_render_backward_workspace_floats(R, S, P, Q))
    ws = torch.empty(max(nws, 1), device=dev)
    gs, gc, gd = _grads(dev, S, P, Q)
    _lib.check(L.n3ds_render_backward(*common, _lib.ptr(ws), int(ws.numel()), _lib.ptr(gs), _lib.ptr(gc),
                                      _lib.ptr(gd), _lib.stream()))
    torch.cuda.synchronize()
    return gs, gc, gd


def _no_alloc(common, dims, dev):
    S, P, Q, R = dims
    gs, gc, gd = _grads(dev, S, P, Q)
    _lib.check(_lib.lib().n3ds_render_backward_noalloc(*common, _lib.ptr(gs), _lib.ptr(gc), _lib.ptr(gd),
                                                       _lib.stream()))
    torch.cuda.synchronize()
    return gs, gc, gd


@pytest.mark.gpu
def test_no_alloc_backward_matches_workspace_backward(dev):
    assert _lib.lib().n3ds_render_backward_noalloc_scratch_floats() == 8 << 20
    big = scenes.make_workload('palette500_65536x128', seed=2, n_rays=2048, n_samples=128, n_strokes=500)
    small = scenes.make_workload('palette500_65536x128', seed=3, n_rays=300, n_samples=64, n_strokes=120)
    for sc in (big, small):
        common, dims, keep = _forward(sc, dev)
        print('no-alloc backward case: rays', dims[3], 'strokes', dims[0], flush=True)
        ref = _with_workspace(common, dims, dev)
        got = _no_alloc(common, dims, dev)
        assert float(ref[0].abs().sum()) > 0
        for a, b in zip(got, ref):
            assert torch.isfinite(a).all()
            assert rel_err(a, b) < 1e-5, rel_err(a, b)


@pytest.mark.gpu
def test_no_alloc_backward_rejects_oversized_stroke_sets(dev):
    """A stroke set whose gradient rows exceed the scratch is refused before any launch."""
    S = (8 << 20) // 18 + 1  # 18 gradient columns per stroke (14 shape + 3 colour + 1 density)
    L = _lib.lib()
    bg = (ctypes.c_float * 3)(1.0, 1.0, 1.0)
    z = torch.zeros(4, device=dev)
    st = L.n3ds_render_backward_noalloc(*([_lib.ptr(z)] * 8), 1, 64, S, 14, 3, 0, 1.0, 0, 0, 0, 1.0,
                                        ctypes.cast(bg, ctypes.c_void_p), None, 0, 1e-7, 1, _lib.ptr(z), None, None,
                                        _lib.ptr(z), None, None, None, 0, 0, _lib.ptr(z), _lib.ptr(z), _lib.ptr(z),
                                        _lib.stream())
    assert st == 1
    assert b'scratch' in L.n3ds_last_error()

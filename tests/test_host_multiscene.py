"""Host-side checks of the scene-batch layout (no GPU calls)."""
import numpy as np
import pytest
import torch

from neural3dstrokes_amd import scenes as W


def test_multiscene_generator_counts_and_packing():
    sc, rc, kc = W.make_multiscene(n_scenes=6, seed=3, stroke_range=(20, 60), ray_range=(16, 64), n_samples=8)
    assert len(sc) == 6 and all(20 <= k <= 60 for k in kc) and all(16 <= r <= 64 for r in rc)
    for s, r, k in zip(sc, rc, kc):
        assert s['shape_params'].shape == (k, 14) and s['tdist'].shape == (r, 9)
        assert s['stroke_ids'].shape == (k,)
    again, rc2, kc2 = W.make_multiscene(n_scenes=6, seed=3, stroke_range=(20, 60), ray_range=(16, 64), n_samples=8)
    assert rc == rc2 and kc == kc2 and torch.equal(sc[2]['shape_params'], again[2]['shape_params'])


def test_batch_plan_rejects_unpacked_tables():
    """The plan validates the scene records before it touches a device."""
    import ctypes
    from neural3dstrokes_amd import _lib
    lib = _lib.lib()
    t = np.zeros((2, 10), dtype=np.int32)
    t[0, 1], t[0, 3] = 10, 5
    t[1, 0], t[1, 1], t[1, 2], t[1, 3] = 11, 10, 5, 5  # second scene does not start after the first
    st = lib.n3ds_render_batch_plan(t.ctypes.data_as(ctypes.c_void_p), 2, 64, 14, 3, 0, None, 0)
    assert st < 0 and b'packed' in lib.n3ds_last_error()
    t[1, 0] = 10
    t[1, 3] = -1
    assert lib.n3ds_render_batch_plan(t.ctypes.data_as(ctypes.c_void_p), 2, 64, 14, 3, 0, None, 0) < 0
    assert lib.n3ds_render_batch_plan(None, 0, 64, 14, 3, 0, None, 0) < 0

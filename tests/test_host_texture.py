"""Host side of the image texture (no device): image set-up as in textures.py:36-49,
the device texel layout, the config factory, and argument validation of the C-ABI."""
import ctypes
import types

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from neural3dstrokes_amd import _lib
from neural3dstrokes_amd.textures import ImageTexture, get_stroke_texture


def _png(tmp_path, arr):
    from PIL import Image
    p = tmp_path / 'tex.png'
    Image.fromarray(arr).save(p)
    return str(p)


def test_image_setup_matches_reference_steps(tmp_path):
    g = np.random.default_rng(0)
    arr = g.integers(0, 256, (40, 30, 3)).astype(np.uint8)
    tex = ImageTexture(_png(tmp_path, arr), (20, 15), tint_ratio=0.5)
    want = F.interpolate((torch.from_numpy(arr.astype(np.float32)) / 255.).permute(2, 0, 1)[None], (20, 15),
                         mode='area')
    assert torch.equal(tex.image_texture, want)
    assert tex.size == (20, 15) and tex.channels == 3 and tex.tint_ratio == 0.5
    lib = _lib.lib()
    assert lib.n3ds_texture_pack_floats(20, 15, 3) == 21 * 16 * 16
    assert lib.n3ds_texture_pack_floats(8, 8, 1) == 9 * 9 * 4
    assert lib.n3ds_texture_pack_floats(8, 8, 4) == -1 and b'RGB or grayscale' in lib.n3ds_last_error()
    assert lib.n3ds_texture_pack_floats(1 << 14, 1 << 13, 3) == -1
    # a grey image keeps one channel; an array gives the same buffers as the file
    grey = g.integers(0, 256, (8, 8)).astype(np.uint8)
    t1, t2 = ImageTexture(_png(tmp_path, grey), (8, 8)), ImageTexture(grey.astype(np.float32), (8, 8))
    assert t1.channels == 1 and torch.equal(t1.image_texture, t2.image_texture)
    t3 = ImageTexture(torch.from_numpy(grey.astype(np.float32)), (8, 8))
    assert torch.equal(t3.image_texture, t2.image_texture)
    assert tex.modulate_alpha is False and ImageTexture(grey, (4, 4), modulate_alpha=True).modulate_alpha


def test_stroke_texture_factory(tmp_path):
    path = _png(tmp_path, np.zeros((4, 4, 3), dtype=np.uint8))
    assert get_stroke_texture(types.SimpleNamespace(stroke_texture='none')) is None
    t = get_stroke_texture(types.SimpleNamespace(stroke_texture='image', texture_image_path=path,
                                                 texture_image_size=(2, 2)))
    assert isinstance(t, ImageTexture) and t.image_texture.shape == (1, 3, 2, 2)
    with pytest.raises(AssertionError):
        get_stroke_texture(types.SimpleNamespace(stroke_texture='video'))


def test_texture_rejects_cpu_tensors_and_bad_arguments():
    tex = ImageTexture(np.zeros((4, 4, 3), dtype=np.float32), (4, 4))
    with pytest.raises(RuntimeError, match='CUDA'):
        tex(torch.rand(5, 2), torch.rand(5, 3), torch.rand(5))
    lib = _lib.lib()
    # validation fails before any device work
    assert lib.n3ds_texture_forward(None, None, None, None, None, None, 4, 4, 2, 0, 1.0, 0, None) == 1
    assert b'RGB or grayscale' in lib.n3ds_last_error()
    assert lib.n3ds_texture_forward(None, None, None, None, None, None, 0, 4, 3, 0, 1.0, 0, None) == 1
    assert lib.n3ds_texture_backward(None, None, None, None, None, None, None, None, None, 4, 4, 3, -1, 1.0, 0,
                                     None) == 1
    assert lib.n3ds_texture_forward(None, None, None, None, None, ctypes.c_void_p(8), 4, 4, 3, 5, 1.0, 0,
                                    None) == 1


def test_resampling_entry_points_validate_arguments():
    """n3ds_sample_intervals / n3ds_max_dilate_weights reject bad sizes before any device work."""
    lib = _lib.lib()
    assert lib.n3ds_sample_intervals(None, None, None, None, None, 0, 0.0, 4, 8, 1, 0.0, 1.0, None) == 1
    assert b'num_samples must be > 1' in lib.n3ds_last_error()
    assert lib.n3ds_sample_intervals(None, None, None, None, None, 0, 0.0, 4, 8192, 4, 0.0, 1.0, None) == 1
    assert lib.n3ds_sample_intervals(None, None, None, None, ctypes.c_void_p(16), 3, 0.0, 4, 8, 4, 0.0, 1.0,
                                     None) == 1
    assert b'jitter width' in lib.n3ds_last_error()
    assert lib.n3ds_sample_intervals(None, None, None, None, None, 0, 0.0, 0, 8, 4, 0.0, 1.0, None) == 0
    assert lib.n3ds_max_dilate_weights(None, None, None, None, 4, 0, 0.1, 0.0, 1.0, 1, None) == 1
    assert lib.n3ds_max_dilate_weights(None, None, None, None, 4, 4000, 0.1, 0.0, 1.0, 1, None) == 1
    assert lib.n3ds_max_dilate_weights(None, None, None, None, 0, 8, 0.1, 0.0, 1.0, 1, None) == 0
    from neural3dstrokes_amd import stepfun
    with pytest.raises(ValueError):
        stepfun.sample_intervals(False, torch.zeros(2, 5), torch.zeros(2, 4), 1)
    with pytest.raises(RuntimeError):
        stepfun.max_dilate_weights(torch.zeros(2, 5), torch.zeros(2, 5), 0.1)

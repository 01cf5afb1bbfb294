"""ctypes binding of the gfx950 C-ABI library ``libn3ds.so``.

The library is built in-tree by ``neural3dstrokes_amd/build.py``.  There is no
fallback: if the library is missing or cannot be loaded every operator raises.
Entry points return 0 on success; on failure the message from
``n3ds_last_error`` is raised as ``RuntimeError`` (the reference's TORCH_CHECK
failures surface as RuntimeError too).
"""
import ctypes
import os

import torch

LIB_PATH = os.environ.get('N3DS_LIBRARY') or os.path.join(os.path.dirname(os.path.abspath(__file__)), 'libn3ds.so')

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int
_U32 = ctypes.c_uint32
_F = ctypes.c_float
_D = ctypes.c_double

SIGNATURES = {
    'n3ds_version': (_I32, []),
    'n3ds_last_error': (ctypes.c_char_p, []),
    'n3ds_stroke_forward': (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _U32, _U32, _F,
                                   _I32, _I32, _P]),
    'n3ds_stroke_backward': (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _U32,
                                    _U32, _F, _I32, _I32, _I32, _P]),
    'n3ds_compose_forward': (_I32, [_P, _P, _P, _P, _P, _I64, _I64, _I64, _P]),
    'n3ds_compose_backward': (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I32, _P]),
    'n3ds_grid_encode_forward': (_I32, [_P, _P, _P, _P, _I64, _I32, _I32, _I32, _F, _I32, _P, _I32, _I32, _I32, _I32,
                                        _P]),
    'n3ds_grid_encode_backward': (_I32, [_P, _P, _P, _P, _P, _I64, _I64, _I32, _I32, _I32, _F, _I32, _P, _P, _I32, _I32,
                                         _I32, _I32, _P, _I64, _P]),
    'n3ds_grid_encode_backward_workspace_bytes': (_I64, [_I64, _I32, _I32, _I32, _I64]),
    'n3ds_grid_grad_tv': (_I32, [_P, _P, _P, _P, _F, _I64, _I32, _I32, _I32, _F, _I32, _I32, _I32, _I32, _P]),
    'n3ds_prop_field_forward': (_I32, [_P, _I64, _I32, _F, _P, _P, _I32, _I32, _F, _I32, _I32, _I32, _I32, _I32, _P,
                                       _P, _P, _P, _F, _P, _P, _P, _P, _P]),
    'n3ds_prop_field_backward': (_I32, [_P, _I64, _I32, _F, _P, _P, _I32, _I32, _F, _I32, _I32, _I32, _I32, _I32, _P,
                                        _P, _P, _P, _F, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    'n3ds_grad_sqnorm_accumulate': (_I32, [_P, _I64, _P, _P]),
    'n3ds_grad_clip_coef': (_I32, [_P, _F, _P, _P]),
    'n3ds_adam_step': (_I32, [_P, _P, _P, _P, _I64, _I32, _P, _P, _D, _P, _D, _D, _D, _I64, _P, _F, _P, _F, _F, _P]),
    'n3ds_render_table_floats': (_I64, [_I32]),
    'n3ds_render_prepare': (_I32, [_P, _P, _P, _I32, _I32, _P, _P]),
    'n3ds_render_forward': (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _I64, _I32, _I32, _I32, _I32, _I32, _F, _I32, _I32,
                                   _I32, _F, _P, _P, _I32, _F, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, _P, _I64, _P]),
    'n3ds_render_forward_workspace_bytes': (_I64, [_I64, _I32]),
    'n3ds_render_backward_split_floats': (_I64, [_I64, _I32, _I32, _I32]),
    'n3ds_render_extras': (_I32, [_P, _P, _P, _P, _I64, _I32, _I32, _P, _P]),
    'n3ds_render_tuning': (_I32, [_I32, _I32]),
    'n3ds_render_backward_workspace_floats': (_I64, [_I64, _I32, _I32, _I32]),
    'n3ds_render_backward': (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _I64, _I32, _I32, _I32, _I32, _I32, _F, _I32,
                                    _I32, _I32, _F, _P, _P, _I32, _F, _I32, _P, _P, _P, _P, _P, _P, _P, _I32, _I32,
                                    _P, _I64, _P, _P, _P, _P]),
    'n3ds_render_backward_noalloc_scratch_floats': (_I64, []),
    'n3ds_render_backward_noalloc': (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _I64, _I32, _I32, _I32, _I32, _I32, _F,
                                            _I32, _I32, _I32, _F, _P, _P, _I32, _F, _I32, _P, _P, _P, _P, _P, _P, _P,
                                            _I32, _I32, _P, _P, _P, _P]),
    'n3ds_render_batch_plan': (_I32, [_P, _I32, _I32, _I32, _I32, _I32, _P, _I32]),
    'n3ds_render_batch_forward': (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I64, _I32, _I32, _I32, _I32,
                                         _I32, _F, _I32, _I32, _I32, _F, _P, _P, _I32, _F, _P, _P, _P, _P, _P, _P, _P,
                                         _P, _I32, _I32, _P, _I64, _P]),
    'n3ds_render_batch_backward_workspace_floats': (_I64, [_P, _I32, _I32, _I32, _I32, _I32]),
    'n3ds_render_batch_backward': (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I64, _I32, _I32, _I32, _I32,
                                          _I32, _F, _I32, _I32, _I32, _F, _P, _P, _I32, _F, _I32, _P, _P, _P, _P, _P,
                                          _P, _P, _I32, _I32, _P, _I64, _P, _P, _P, _P]),
    'n3ds_sample_intervals': (_I32, [_P, _P, _P, _P, _P, _I32, _F, _I64, _I32, _I32, _F, _F, _P]),
    'n3ds_max_dilate_weights': (_I32, [_P, _P, _P, _P, _I64, _I32, _F, _F, _F, _I32, _P]),
    'n3ds_texture_pack_floats': (_I64, [_I64, _I64, _I64]),
    'n3ds_texture_pack': (_I32, [_P, _P, _I64, _I64, _I64, _P]),
    'n3ds_texture_forward': (_I32, [_P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _D, _I32, _P]),
    'n3ds_texture_backward': (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _D, _I32, _P]),
}

_lib = None


def lib():
    """Load (once) and return the ctypes library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f'{LIB_PATH} is not built; run `python -m neural3dstrokes_amd.build`')
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status):
    if status != 0:
        msg = lib().n3ds_last_error().decode(errors='replace')
        raise RuntimeError(f'n3ds error {status}: {msg}')


def ptr(t):
    """Device pointer of a tensor, or NULL for None / empty tensors (absent optional)."""
    if t is None or t.numel() == 0:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def require_cuda(name, t, dtype=torch.float32, contiguous=True):
    """TORCH_CHECK-style argument validation (common.h:10-23 of the reference)."""
    if t is None:
        return
    if not t.is_cuda:
        raise RuntimeError(f'{name} must be a CUDA tensor')
    if contiguous and not t.is_contiguous():
        raise RuntimeError(f'{name} must be a contiguous tensor')
    if dtype is not None and t.dtype != dtype:
        raise RuntimeError(f'{name} must be a {dtype} tensor')

"""Reads the stroke field of a reference checkpoint for the fused renderer.

A reference run saves its model with ``accelerator.save_state``
(``train.py``); the stroke field's tensors are ``<prefix>shape_params``
``[max_num_strokes, P]``, ``<prefix>color_params``, ``<prefix>density_params``
and the ``<prefix>stroke_step`` buffer (``source/models.py:558-571``).  The
strokes in use are ``[0, stroke_step)`` and the composited density parameter is
``density_params * density_scale`` (``models.py:666-675``); the stroke type is
the run's ``shape_type`` (one per field, ``models.py:529``).

Files are read only with loaders that execute nothing from them:
safetensors, or ``torch.load(weights_only=True)``.
"""
import torch

from .strokelib.strokes import _color_dict, _make_sdf_id, _sdf_dict, get_stroke


def stroke_sdf_id(shape_type):
    """The reference's ``sdf_id`` for a shape type (``strokes.py:95-106``)."""
    base_name, _, _, use_t, use_r, use_ss, use_ms = _sdf_dict[shape_type]
    return _make_sdf_id(base_name, use_t, use_r, use_ss, use_ms)


def load_state(path):
    """Tensor dict of a checkpoint file (``.safetensors`` or a torch pickle of tensors)."""
    if str(path).endswith('.safetensors'):
        from safetensors.torch import load_file
        return load_file(str(path))
    state = torch.load(str(path), map_location='cpu', weights_only=True)
    if isinstance(state, dict) and 'model' in state and isinstance(state['model'], dict):
        state = state['model']
    return state


def _prefix(state):
    hits = [k[:-len('shape_params')] for k in state if k.endswith('shape_params')
            and k[:-len('shape_params')] + 'stroke_step' in state]
    if len(hits) != 1:
        raise RuntimeError(f'expected one stroke field in the checkpoint, found prefixes {hits}')
    return hits[0]


def stroke_field_from_state(state, shape_type, color_type, density_scale=20.0, prefix=None):
    """Arguments of ``render_strokes`` for a reference stroke field's state dict.

    Returns dict(stroke_ids int32 [n], shape_params [n, P], color_params [n, Q],
    density_params [n] (already scaled), color_id, n_strokes) for the n =
    stroke_step strokes in use.
    """
    prefix = _prefix(state) if prefix is None else prefix
    sp = state[prefix + 'shape_params'].float()
    cp = state[prefix + 'color_params'].float()
    dp = state[prefix + 'density_params'].float()
    n = int(state[prefix + 'stroke_step'])
    _, dim_s, dim_c, _, _, _, _ = get_stroke(shape_type, color_type, 'recon')
    if sp.shape[1] != dim_s or cp.shape[1] != dim_c:
        raise RuntimeError(f'checkpoint widths ({sp.shape[1]}, {cp.shape[1]}) do not match {shape_type} / '
                           f'{color_type} ({dim_s}, {dim_c})')
    if not 0 <= n <= sp.shape[0]:
        raise RuntimeError(f'stroke_step {n} outside [0, {sp.shape[0]}]')
    return dict(stroke_ids=torch.full((n,), stroke_sdf_id(shape_type), dtype=torch.int32),
                shape_params=sp[:n].contiguous(), color_params=cp[:n].contiguous(),
                density_params=(dp[:n] * density_scale).contiguous(), color_id=_color_dict[color_type][0],
                n_strokes=n)


def load_stroke_field(path, shape_type, color_type, density_scale=20.0, prefix=None):
    return stroke_field_from_state(load_state(path), shape_type, color_type, density_scale, prefix)

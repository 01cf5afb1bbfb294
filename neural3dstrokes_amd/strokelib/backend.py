"""``_backend`` shim: the reference's pybind functions, re-implemented over the C-ABI.

Signatures follow ``source/strokelib/src/strokes.h:14-54`` and
``compositing.h:11-32`` (bound in ``source/strokelib/src/bindings.cpp:6-12``):
outputs are allocated by the caller, an empty tensor means "absent", all
tensors are contiguous float32 CUDA tensors, launches are asynchronous on the
current stream.  One extra trailing argument selects the adjoint
(``exact``: true derivative of the forward; ``reference``: the reference's
hand-written adjoint term for term).
"""
from .. import _lib


def _mode_flag(adjoint):
    if adjoint not in ('exact', 'reference'):
        raise ValueError(f"adjoint must be 'exact' or 'reference', got {adjoint!r}")
    return 1 if adjoint == 'exact' else 0


def stroke_forward(alpha_output, color_output, sdf_output, texcoord_output, x, radius, viewdir, shape_params,
                   color_params, sdf_id, color_id, sdf_delta, use_laplace_transform, inv_scale_radius):
    """strokes.h:14-27 / strokes_forward.cu:170-225."""
    for name, t in (('alpha_output', alpha_output), ('color_output', color_output), ('sdf_output', sdf_output),
                    ('texcoord_output', texcoord_output), ('x', x), ('radius', radius), ('viewdir', viewdir),
                    ('shape_params', shape_params), ('color_params', color_params)):
        _lib.require_cuda(name, t)
    _lib.check(_lib.lib().n3ds_stroke_forward(
        _lib.ptr(alpha_output), _lib.ptr(color_output), _lib.ptr(sdf_output), _lib.ptr(texcoord_output),
        _lib.ptr(x), _lib.ptr(radius), _lib.ptr(viewdir), _lib.ptr(shape_params), _lib.ptr(color_params),
        x.shape[0], viewdir.shape[0], shape_params.shape[0], shape_params.shape[1], color_params.shape[1],
        int(sdf_id), int(color_id), float(sdf_delta), int(bool(use_laplace_transform)), int(bool(inv_scale_radius)),
        _lib.stream()))


def stroke_backward(grad_shape_params, grad_color_params, grad_x, grad_alpha, grad_color, grad_sdf, x, radius, viewdir,
                    alpha, shape_params, color_params, sdf_id, color_id, sdf_delta, use_laplace_transform,
                    inv_scale_radius, adjoint='exact'):
    """strokes.h:38-54 / strokes_backward.cu:224-287 (accumulates into the zero-initialised grads)."""
    for name, t in (('grad_shape_params', grad_shape_params), ('grad_color_params', grad_color_params),
                    ('grad_x', grad_x), ('grad_alpha', grad_alpha), ('grad_color', grad_color),
                    ('grad_sdf', grad_sdf), ('x', x), ('radius', radius), ('viewdir', viewdir), ('alpha', alpha),
                    ('shape_params', shape_params), ('color_params', color_params)):
        _lib.require_cuda(name, t)
    _lib.check(_lib.lib().n3ds_stroke_backward(
        _lib.ptr(grad_shape_params), _lib.ptr(grad_color_params), _lib.ptr(grad_x), _lib.ptr(grad_alpha),
        _lib.ptr(grad_color), _lib.ptr(grad_sdf), _lib.ptr(x), _lib.ptr(radius), _lib.ptr(viewdir), _lib.ptr(alpha),
        _lib.ptr(shape_params), _lib.ptr(color_params), x.shape[0], viewdir.shape[0], shape_params.shape[0],
        shape_params.shape[1], color_params.shape[1], int(sdf_id), int(color_id), float(sdf_delta),
        int(bool(use_laplace_transform)), int(bool(inv_scale_radius)), _mode_flag(adjoint), _lib.stream()))


def compose_forward(density_output, color_output, alphas, colors, density_params):
    """compositing.h:11-18 / compositing.cu:159-204."""
    for name, t in (('density_output', density_output), ('color_output', color_output), ('alphas', alphas),
                    ('colors', colors), ('density_params', density_params)):
        _lib.require_cuda(name, t)
    _lib.check(_lib.lib().n3ds_compose_forward(
        _lib.ptr(density_output), _lib.ptr(color_output), _lib.ptr(alphas), _lib.ptr(colors),
        _lib.ptr(density_params), alphas.shape[0], alphas.shape[1], colors.shape[-1], _lib.stream()))


def compose_backward(grad_alphas, grad_colors, grad_density_params, grad_density_output, grad_color_output, alphas,
                     colors, density_params, adjoint='exact'):
    """compositing.h:24-32 / compositing.cu:206-263.  Empty upstream grads are treated as zeros."""
    for name, t in (('grad_alphas', grad_alphas), ('grad_colors', grad_colors),
                    ('grad_density_params', grad_density_params), ('grad_density_output', grad_density_output),
                    ('grad_color_output', grad_color_output), ('alphas', alphas), ('colors', colors),
                    ('density_params', density_params)):
        _lib.require_cuda(name, t)
    _lib.check(_lib.lib().n3ds_compose_backward(
        _lib.ptr(grad_alphas), _lib.ptr(grad_colors), _lib.ptr(grad_density_params), _lib.ptr(grad_density_output),
        _lib.ptr(grad_color_output), _lib.ptr(alphas), _lib.ptr(colors), _lib.ptr(density_params),
        alphas.shape[0], alphas.shape[1], colors.shape[-1], _mode_flag(adjoint), _lib.stream()))

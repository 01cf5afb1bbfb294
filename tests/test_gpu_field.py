"""A reference-layout checkpoint rendered on the GPU, and step_update with the error field.

The checkpoint holds a model's ``nerf.*`` (StrokeField), ``prop_mlp_0.*``
(PropMLP) and ``error_field.*`` (ErrorMLP) tensors under the reference's names
(``models.py:77-96``).  After loading: the stroke field renders through the
fused operator and is checked against the float64 oracle of the chain
predict_density -> compute_alpha_weights -> volumetric_rendering on the same
arguments StrokeField derives (density scale, the ``_warp_coords`` divisor,
cone radii, the training-time sdf_delta, strokes behind ``max_opt_strokes``
fixed); the proposal and error fields against their plain torch formulation.
"""
import pytest
import torch

from conftest import rel_err
from neural3dstrokes_amd import scenes
from neural3dstrokes_amd.field import ErrorField, StrokeField, load_model_state, unwarp_coords
from neural3dstrokes_amd.propfield import PropDensityField, reference_forward
from oracle import field as ofield

D64 = torch.float64


def _checkpoint(tmp_path):
    from safetensors.torch import save_file
    torch.manual_seed(0)
    f = StrokeField(shape_type='ellipsoid', max_num_strokes=64, max_opt_strokes=30, init_num_strokes=50,
                    use_laplace_transform=True, bbox_size=3.0, sdf_delta=12.0, sdf_delta_eval=4.0)
    f.step_update(0, 64 * 100)  # the initial 50 strokes from the reference's samplers
    f.density_params.data[:50] = torch.rand(50) * 0.5 + 0.02
    prop = PropDensityField(grid_disired_resolution=512, grid_log2_hashmap_size=16)
    err = ErrorField(grid_log2_hashmap_size=14)
    for m in (prop, err):
        torch.nn.init.uniform_(m.encoder.embeddings, -1e-1, 1e-1)
    state = {}
    for name, m in (('nerf', f), ('prop_mlp_0', prop), ('error_field', err)):
        state.update({f'{name}.{k}': v.detach().contiguous() for k, v in m.state_dict().items()})
    path = tmp_path / 'model.safetensors'
    save_file(state, str(path))
    return path


@pytest.mark.gpu
def test_loaded_stroke_field_renders_as_the_reference_chain(tmp_path, dev):
    parts = load_model_state(_checkpoint(tmp_path))
    f = StrokeField(shape_type='ellipsoid', max_num_strokes=64, max_opt_strokes=30, use_laplace_transform=True,
                    bbox_size=3.0, sdf_delta=12.0, sdf_delta_eval=4.0).to(dev)
    f.load_state_dict(parts['nerf'])
    assert int(f.stroke_step) == 50 and f.stroke_range() == (20, 50)
    gen = torch.Generator().manual_seed(5)
    rays = scenes.make_rays(96, 32, gen, cam_radius=3.0, radius=0.002)
    g = torch.Generator().manual_seed(6)
    g_rgb, g_acc = torch.randn(96, 3, generator=g), torch.randn(96, generator=g)
    f.train()
    d = lambda k: rays[k].to(dev)  # noqa: E731
    rgb, acc, depth = f.render(d('origins'), d('directions'), d('viewdirs'), d('radii'), d('tdist'))
    ((rgb * g_rgb.to(dev)).sum() + (acc * g_acc.to(dev)).sum()).backward()
    delta = 12.0 * (1 - 50 / 64) + 4.0 * 50 / 64
    ids = torch.full((50,), f.sdf_id, dtype=torch.int32)
    args = tuple(rays[k].to(dev, D64) for k in ('origins', 'directions', 'viewdirs', 'radii', 'tdist'))
    out, gs, gc, gd = ofield.render_field_grads_chunked(
        g_rgb.to(dev, D64), g_acc.to(dev, D64), *args, ids, f.shape_params.detach()[:50].to(D64),
        f.color_params.detach()[:50].to(D64), (f.density_params.detach()[:50] * f.density_scale).to(D64), chunk=32,
        sdf_delta=delta, laplace=True, inv_scale_radius=True, cone=True, coord_div=1.5)
    assert float(out['acc'].max()) > 0.3
    assert rel_err(rgb, out['rgb']) < 1e-5 and rel_err(acc, out['acc']) < 1e-5
    # strokes 0..19 are behind the optimised window: no gradient; 20..49 get the chain's
    assert float(f.shape_params.grad[:20].abs().max()) == 0.0 and float(f.density_params.grad[:20].abs().max()) == 0.0
    assert float(f.shape_params.grad[50:].abs().max()) == 0.0
    assert rel_err(f.shape_params.grad[20:50], gs[20:]) < 1e-4
    assert rel_err(f.color_params.grad[20:50], gc[20:]) < 1e-4
    assert rel_err(f.density_params.grad[20:50], gd[20:] * f.density_scale) < 1e-4


@pytest.mark.gpu
def test_loaded_proposal_and_error_fields(tmp_path, dev):
    parts = load_model_state(_checkpoint(tmp_path))
    prop = PropDensityField(grid_disired_resolution=512, grid_log2_hashmap_size=16).to(dev)
    prop.load_state_dict(parts['prop_mlp_0'])
    err = ErrorField(grid_log2_hashmap_size=14).to(dev)
    err.load_state_dict(parts['error_field'])
    x = (torch.rand(4096, 3, generator=torch.Generator().manual_seed(7)) * 6 - 3).to(dev)
    prop.eval()
    with torch.no_grad():
        assert rel_err(prop(x)['density'], reference_forward(prop, x)) < 1e-5
        assert rel_err(err.sample_error(x), reference_forward(err, x)) < 1e-5


@pytest.mark.gpu
def test_step_update_places_new_strokes_at_the_largest_errors(dev):
    torch.manual_seed(3)
    f = StrokeField(shape_type='ellipsoid', max_num_strokes=40, init_num_strokes=10, error_point_samples=5000,
                    warp_fn=None, bbox_size=3.0).to(dev)
    err = ErrorField(grid_log2_hashmap_size=14, warp_fn=None, bbox_size=3.0).to(dev)
    torch.nn.init.uniform_(err.encoder.embeddings, -1.0, 1.0)
    f.step_update(0, 4000)
    f.step_update(400, 4000)  # past the initial strokes
    prev = int(f.stroke_step)
    assert prev > 10
    torch.cuda.manual_seed(11)
    upd = f.step_update(2000, 4000, error_field=err)
    nxt = upd[1]
    torch.cuda.manual_seed(11)  # replay the sample points
    pts = torch.rand((5000, 3), device=dev) * 2 - 1
    with torch.no_grad():
        e = reference_forward(err, unwarp_coords(None, pts, 3.0))
    top = torch.topk(e, k=nxt - prev).indices
    # recon samplers centre a new stroke on its error point: translation = the point (clipped to [-1, 1])
    assert torch.allclose(f.shape_params.data[prev:nxt, 6:9], pts[top].clamp(-1, 1), atol=1e-6)

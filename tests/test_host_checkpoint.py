"""Reading a reference-layout stroke-field checkpoint (safe loaders only)."""
import pytest
import torch

from neural3dstrokes_amd import checkpoint
from neural3dstrokes_amd.strokelib.strokes import get_stroke


class _Field(torch.nn.Module):  # the reference StrokeField's parameters and buffer names
    def __init__(self, shape_type, color_type, n=12):
        super().__init__()
        _, ds, dc, _, _, _, _ = get_stroke(shape_type, color_type, 'recon')
        g = torch.Generator().manual_seed(0)
        self.shape_params = torch.nn.Parameter(torch.rand(n, ds, generator=g))
        self.color_params = torch.nn.Parameter(torch.rand(n, dc, generator=g))
        self.density_params = torch.nn.Parameter(torch.rand(n, generator=g))
        self.register_buffer('stroke_step', torch.tensor(7, dtype=torch.int32))
        self.register_buffer('shape_params_grad', torch.zeros(n, ds))


class _Model(torch.nn.Module):
    def __init__(self, *a):
        super().__init__()
        self.nerf = _Field(*a)
        self.prop = torch.nn.Linear(3, 4)


@pytest.mark.parametrize('fmt', ['safetensors', 'torch'])
@pytest.mark.parametrize('shape_type,color_type', [('ellipsoid', 'constant_rgb'), ('cubic_bezier', 'constant_sh2')])
def test_load_reference_layout(tmp_path, fmt, shape_type, color_type):
    m = _Model(shape_type, color_type)
    state = {k: v.detach().clone() for k, v in m.state_dict().items()}
    if fmt == 'safetensors':
        from safetensors.torch import save_file
        path = tmp_path / 'model.safetensors'
        save_file(state, str(path))
    else:
        path = tmp_path / 'pytorch_model.bin'
        torch.save(state, str(path))
    f = checkpoint.load_stroke_field(path, shape_type, color_type, density_scale=20.0)
    assert f['n_strokes'] == 7
    assert torch.equal(f['shape_params'], m.nerf.shape_params.detach()[:7])
    assert torch.equal(f['color_params'], m.nerf.color_params.detach()[:7])
    assert torch.allclose(f['density_params'], m.nerf.density_params.detach()[:7] * 20.0)
    assert f['stroke_ids'].dtype == torch.int32 and f['stroke_ids'].shape == (7,)
    base = f['stroke_ids'][0].item() >> 4
    assert base == (0 if shape_type == 'ellipsoid' else 10)
    with pytest.raises(RuntimeError):
        checkpoint.stroke_field_from_state(state, 'sphere' if shape_type != 'sphere' else 'cube', color_type)

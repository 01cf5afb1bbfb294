"""CPU pins of the optimizer oracle (oracle/optim.py) against torch's own Adam and clipping."""
import pytest
import torch

from oracle import optim as O


def _grads(seed, shapes, bad=True):
    g = torch.Generator().manual_seed(seed)
    gs = [torch.randn(*s, generator=g, dtype=torch.float64) for s in shapes]
    if bad:
        gs[0].view(-1)[3] = float('nan')
        gs[0].view(-1)[5] = float('inf')
        gs[-1].view(-1)[0] = -float('inf')
    return gs


@pytest.mark.parametrize('max_norm,max_val', [(0.0, 0.0), (0.5, 0.0), (0.0, 0.3), (2.0, 0.4)])
def test_clip_gradients_matches_torch(max_norm, max_val):
    shapes = [(7, 9), (7, 3), (7,)]
    gs = _grads(0, shapes, bad=False)
    ps = [torch.zeros(s, dtype=torch.float64, requires_grad=True) for s in shapes]
    for p, g in zip(ps, gs):
        p.grad = g.clone()
    if max_norm > 0:
        torch.nn.utils.clip_grad_norm_(ps, max_norm)
    if max_val > 0:
        torch.nn.utils.clip_grad_value_(ps, max_val)
    for p in ps:
        p.grad.nan_to_num_()
    ours = O.clip_gradients(gs, max_norm, max_val)
    for p, o in zip(ps, ours):
        torch.testing.assert_close(o, p.grad, rtol=1e-12, atol=1e-15)


def test_nan_to_num_semantics():
    g = O.clip_gradients([torch.tensor([float('nan'), float('inf'), -float('inf'), 1.5], dtype=torch.float32)])[0]
    assert g.tolist() == [0.0, torch.finfo(torch.float32).max, -torch.finfo(torch.float32).max, 1.5]


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_adam_matches_torch_optim(dtype):
    g = torch.Generator().manual_seed(1)
    p0 = torch.randn(11, 9, generator=g, dtype=dtype)
    ref = p0.clone().requires_grad_(True)
    opt = torch.optim.Adam([ref], lr=1e-2, betas=(0.9, 0.99), eps=1e-15)
    p, m, v = p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0)
    for t in range(1, 8):
        grad = torch.randn(11, 9, generator=g, dtype=dtype) * (10.0 ** (t % 3 - 1))
        ref.grad = grad.clone()
        opt.step()
        p, m, v = O.adam_step(p, grad, m, v, t, 1e-2)
        torch.testing.assert_close(p, ref.detach(), rtol=0, atol=0)  # same formulas, same order
    torch.testing.assert_close(m, opt.state[ref]['exp_avg'], rtol=0, atol=0)


def test_clip_params_columns_and_bounds():
    p = torch.linspace(-2, 2, 24, dtype=torch.float64).reshape(6, 4)
    out = O.clip_params(p, ranges=[(0.01, 0.5), (None, 0.0), (-1.0, None), (None, None)])
    assert out[:, 0].min() >= 0.01 and out[:, 0].max() <= 0.5
    assert out[:, 1].max() <= 0.0 and torch.equal(out[:, 1][p[:, 1] <= 0], p[:, 1][p[:, 1] <= 0])
    assert out[:, 2].min() >= -1.0
    assert torch.equal(out[:, 3], p[:, 3])
    d = O.clip_params(torch.tensor([-1.0, 0.5]), min_value=0.0)
    assert d.tolist() == [0.0, 0.5]

"""HIP compose_forward / compose_backward (drop-in ``compose_strokes``) against the CPU oracle."""
import pytest
import torch

from conftest import rel_err
from helpers import rand_like_seeded
from neural3dstrokes_amd.strokelib import strokes as dstrokes
from oracle import compose as ocompose


def _inputs(n, S, D, seed, zero_frac=0.5, sat=False):
    g = torch.Generator().manual_seed(seed)
    a = torch.rand(n, S, generator=g) * 0.9999
    a = torch.where(torch.rand(n, S, generator=g) < zero_frac, torch.zeros_like(a), a)
    if sat:
        a = torch.where(torch.rand(n, S, generator=g) < 0.05, torch.full_like(a, 0.9999), a)
    c = torch.rand(n, S, D, generator=g) * 1.2 - 0.1
    d = torch.exp(torch.rand(S, generator=g) * 4.6)
    return a, c, d


@pytest.mark.gpu
@pytest.mark.parametrize('S', [1, 5, 50, 64, 65, 200])
@pytest.mark.parametrize('D', [1, 3])
def test_compose_forward(S, D, dev):
    a, c, d = _inputs(257, S, D, S + D)
    dens, col = dstrokes.compose_strokes(a.to(dev), c.to(dev), d.to(dev), 'over')
    rd, rc, _, _ = ocompose.compose_forward(a.double(), c.double(), d.double())
    assert rel_err(dens, rd) < 1e-5
    assert rel_err(col, rc) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize('S', [1, 5, 64, 130])
@pytest.mark.parametrize('adjoint', ['exact', 'reference'])
@pytest.mark.parametrize('D', [1, 3])
def test_compose_backward(S, adjoint, D, dev):
    a, c, d = _inputs(300, S, D, 100 + S, sat=True)
    gd = rand_like_seeded((300,), S)
    gc = rand_like_seeded((300, D), S + 1)
    ad, cd, dd = (t.to(dev).requires_grad_(True) for t in (a, c, d))
    dstrokes.set_adjoint(adjoint)
    try:
        dens, col = dstrokes.compose_strokes(ad, cd, dd, 'over')
        ((dens * gd.to(dev)).sum() + (col * gc.to(dev)).sum()).backward()
    finally:
        dstrokes.set_adjoint('exact')
    if adjoint == 'exact':
        ra, rc, rdp = ocompose.compose_backward_exact(gd.double(), gc.double(), a.double(), c.double(), d.double())
    else:
        ra, rc, rdp = ocompose.compose_backward_reference(gd.double(), gc.double(), a.double(), c.double(),
                                                          d.double())
    assert rel_err(ad.grad, ra) < 1e-4
    assert rel_err(cd.grad, rc) < 1e-4
    assert rel_err(dd.grad, rdp) < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize('S', [8200, 20000])
@pytest.mark.parametrize('adjoint', ['exact', 'reference'])
def test_compose_backward_many_strokes(S, adjoint, dev):
    """Beyond 8192 strokes per point (no limit in the reference): checkpointed transmittance and
    density-parameter gradients accumulated in global memory.  Small alphas keep the strokes at the
    back visible (total optical depth about 1)."""
    n = 24
    g = torch.Generator().manual_seed(S)
    a = torch.rand(n, S, generator=g) * (4.0 / S)
    a = torch.where(torch.rand(n, S, generator=g) < 0.5, torch.zeros_like(a), a)
    c = torch.rand(n, S, 3, generator=g)
    d = torch.exp(torch.rand(S, generator=g) * 4.6)
    gd = rand_like_seeded((n,), S)
    gc = rand_like_seeded((n, 3), S + 1)
    ad, cd, dd = (t.to(dev).requires_grad_(True) for t in (a, c, d))
    dstrokes.set_adjoint(adjoint)
    try:
        dens, col = dstrokes.compose_strokes(ad, cd, dd, 'over')
        ((dens * gd.to(dev)).sum() + (col * gc.to(dev)).sum()).backward()
    finally:
        dstrokes.set_adjoint('exact')
    rd, rcol, _, _ = ocompose.compose_forward(a.double(), c.double(), d.double())
    # float32 running sums over S / 64 chunks of about 32 nonzero terms: relative error ~ S / 64 ulp
    assert rel_err(dens, rd) < 5e-5 and rel_err(col, rcol) < 5e-5
    if adjoint == 'exact':  # autograd of the closed form (cumulative product), the same derivative
        leaves = [t.double().requires_grad_(True) for t in (a, c, d)]
        cd_, cc_ = ocompose.compose_closed_form(*leaves)
        torch.autograd.backward([cd_, cc_], [gd.double(), gc.double()])
        ra, rc, rdp = (t.grad for t in leaves)
    else:
        ra, rc, rdp = ocompose.compose_backward_reference(gd.double(), gc.double(), a.double(), c.double(),
                                                          d.double())
    assert rel_err(ad.grad, ra) < 1e-4
    assert rel_err(cd.grad, rc) < 1e-4
    assert rel_err(dd.grad, rdp) < 1e-4


@pytest.mark.gpu
def test_compose_missing_grad_is_zero(dev):
    """An absent upstream grad (empty tensor in the reference) is treated as zeros."""
    a, c, d = _inputs(64, 7, 3, 5)
    ad = a.to(dev).requires_grad_(True)
    dens, col = dstrokes.compose_strokes(ad, c.to(dev), d.to(dev), 'over')
    dens.sum().backward()
    ra, _, _ = ocompose.compose_backward_exact(torch.ones(64, dtype=torch.float64), None, a.double(), c.double(),
                                               d.double())
    assert rel_err(ad.grad, ra) < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize('mode', ['max', 'max_density_weighted', 'softmax', 'softmax_density_weighted'])
def test_compose_other_modes(mode, dev):
    a, c, d = _inputs(33, 9, 3, 11, zero_frac=0.0)
    dens, col = dstrokes.compose_strokes(a.to(dev), c.to(dev), d.to(dev), mode)
    rd, rc = ocompose.compose_other(a.double(), c.double(), d.double(), mode)
    assert rel_err(dens, rd) < 1e-5
    assert rel_err(col, rc) < 1e-5

"""HIP interval resampling (n3ds_sample_intervals through ``stepfun.sample_intervals``)
against the float64 oracle restatement of stepfun.py:5-317, on the same jitter draws
(the wrapper draws them with the reference's torch calls).  Outputs 1e-5."""
import numpy as np
import pytest
import torch

from neural3dstrokes_amd import stepfun
from oracle import stepfun as O

pytestmark = pytest.mark.gpu


def _case(seed, r, m):
    g = np.random.default_rng(seed)
    t = np.sort(g.uniform(0, 1, (r, m + 1)), -1)
    t[:, 0], t[:, -1] = 0.0, 1.0
    if m > 5:
        t[:, 5] = t[:, 4]  # a zero-width bin
    w = g.uniform(0, 1, (r, m))
    w[g.uniform(size=(r, m)) < 0.2] = 0.0  # zero-weight bins
    return t.astype(np.float32), w.astype(np.float32)


@pytest.mark.parametrize('m,n,rand,single', [(64, 64, False, True), (64, 64, True, True), (128, 193, True, False),
                                             (3000, 48, True, True), (1, 2, True, True), (31, 7, False, True)])
def test_sample_intervals_matches_oracle(m, n, rand, single, dev):
    r = 300
    t, w = _case(m * 7 + n, r, m)
    sd, wd = torch.from_numpy(t).to(dev), torch.from_numpy(w).to(dev)
    logits = stepfun.resample_logits(sd, wd, 0.0)
    torch.manual_seed(5)
    out = stepfun.sample_intervals(rand, sd.reshape(3, 100, m + 1), logits.reshape(3, 100, m), n,
                                   single_jitter=single, domain=(0.0, 1.0))
    assert out.shape == (3, 100, n + 1) and out.dtype == torch.float32
    torch.manual_seed(5)
    draws = torch.rand((3, 100, 1 if single else n), device=dev).reshape(r, -1).double().cpu().numpy() \
        if rand else None
    u = O.sample_u(n, draws)
    want = O.sample_intervals(u, t.astype(np.float64), logits.double().cpu().numpy(), domain=(0.0, 1.0))
    got = out.reshape(r, n + 1).double().cpu().numpy()
    assert np.abs(got - want).max() < 1e-5
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5
    assert np.all(np.diff(got, axis=-1) >= 0)


def test_two_proposal_levels(dev):
    """models.py:125-181 without dilation: [0, 1] with weight 1, then the rendered
    weights of the first level drive the second."""
    r = 512
    sd = torch.tensor([0.0, 1.0], device=dev).repeat(r, 1)
    w = torch.ones(r, 1, device=dev)
    torch.manual_seed(1)
    s1 = stepfun.sample_intervals(True, sd, stepfun.resample_logits(sd, w), 64, single_jitter=True, domain=(0., 1.))
    g = torch.Generator().manual_seed(2)
    w1 = torch.rand(r, 64, generator=g).to(dev)
    s2 = stepfun.sample_intervals(True, s1, stepfun.resample_logits(s1, w1), 32, single_jitter=True, domain=(0., 1.))
    torch.manual_seed(1)
    d1 = torch.rand(r, 1, device=dev).double().cpu().numpy()
    d2 = torch.rand(r, 1, device=dev).double().cpu().numpy()
    o1 = O.sample_intervals(O.sample_u(64, d1), sd.double().cpu().numpy(),
                            stepfun.resample_logits(sd, w).double().cpu().numpy(), (0.0, 1.0))
    assert np.abs(s1.double().cpu().numpy() - o1).max() < 1e-5
    o2 = O.sample_intervals(O.sample_u(32, d2), s1.double().cpu().numpy(),
                            stepfun.resample_logits(s1, w1).double().cpu().numpy(), (0.0, 1.0))
    assert np.abs(s2.double().cpu().numpy() - o2).max() < 1e-5


def test_sample_intervals_errors(dev):
    t = torch.linspace(0, 1, 9, device=dev).repeat(4, 1)
    lg = torch.zeros(4, 8, device=dev)
    with pytest.raises(ValueError):
        stepfun.sample_intervals(False, t, lg, 1)
    with pytest.raises(RuntimeError):
        stepfun.sample_intervals(False, t, lg[:, :7], 4)
    big = torch.linspace(0, 1, 8193, device=dev)[None]
    with pytest.raises(RuntimeError):
        stepfun.sample_intervals(False, big, torch.zeros(1, 8192, device=dev), 4)
    with pytest.raises(RuntimeError, match='CUDA'):
        stepfun.sample_intervals(False, t.cpu(), lg.cpu(), 4)
    empty = stepfun.sample_intervals(True, t[:0], lg[:0], 4)
    assert empty.shape == (0, 5)

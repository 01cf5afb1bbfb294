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


@pytest.mark.parametrize('m,dilation,renorm,domain', [(64, 0.01, True, (0.0, 1.0)), (1, 0.25, True, (0.0, 1.0)),
                                                      (200, 0.002, False, (-np.inf, np.inf)),
                                                      (1500, 0.0005, True, (0.0, 1.0)), (33, 0.0, True, (0.0, 1.0))])
def test_max_dilate_weights_matches_oracle(m, dilation, renorm, domain, dev):
    r = 200
    t, w = _case(m + 3, r, m)
    w = w / np.maximum(w.sum(-1, keepdims=True), 1e-12)
    td, wd = stepfun.max_dilate_weights(torch.from_numpy(t).to(dev).reshape(2, 100, m + 1),
                                        torch.from_numpy(w.astype(np.float32)).to(dev).reshape(2, 100, m), dilation,
                                        domain=domain, renormalize=renorm)
    assert td.shape == (2, 100, 3 * m + 1) and wd.shape == (2, 100, 3 * m)
    # the dilated edges t -+ d are float32 in the reference too: with 1500 bins a piece
    # between neighbouring dilated edges can be ~1e-5 wide, so their rounding is kept
    ot, ow = O.max_dilate_weights(t.astype(np.float64), w.astype(np.float32).astype(np.float64),
                                  np.float32(dilation).item(), domain, renorm, edge_dtype=np.float32)
    gt, gw = td.reshape(r, -1).double().cpu().numpy(), wd.reshape(r, -1).double().cpu().numpy()
    assert np.abs(gt - ot).max() < 1e-6
    assert np.linalg.norm(gw - ow) / np.linalg.norm(ow) < 1e-5
    assert np.all(np.diff(gt, axis=-1) >= 0)


def test_dilate_then_resample_like_models(dev):
    """models.py:159-181: dilate the first level's weights, drop the end fenceposts, resample."""
    r, m = 256, 64
    t, w = _case(77, r, m)
    sd, wd = torch.from_numpy(t).to(dev), torch.from_numpy(w).to(dev)
    td, wdl = stepfun.max_dilate_weights(sd, wd, 0.005, domain=(0.0, 1.0), renormalize=True)
    td, wdl = td[..., 1:-1], wdl[..., 1:-1]
    torch.manual_seed(3)
    out = stepfun.sample_intervals(True, td, stepfun.resample_logits(td, wdl), 128, single_jitter=True,
                                   domain=(0.0, 1.0))
    ot, ow = O.max_dilate_weights(t.astype(np.float64), w.astype(np.float64), np.float32(0.005).item(), (0.0, 1.0),
                                  True, edge_dtype=np.float32)
    torch.manual_seed(3)
    draws = torch.rand(r, 1, device=dev).double().cpu().numpy()
    lg = stepfun.resample_logits(td, wdl).double().cpu().numpy()
    want = O.sample_intervals(O.sample_u(128, draws), td.double().cpu().numpy(), lg, (0.0, 1.0))
    assert np.abs(td.double().cpu().numpy() - ot[..., 1:-1]).max() < 1e-6
    # the inverse CDF jumps across a bin of zero weight and positive width (a flat CDF
    # level); the float32 CDF (torch's cumsum in the reference, the wave scan here) can put
    # a u within ~1e-6 of such a level on the other side: compare fenceposts away from them
    tdn = td.double().cpu().numpy()
    cw = O.integrate_weights(O.softmax(lg))
    flat = (np.diff(cw, axis=-1) == 0) & (np.diff(tdn, axis=-1) > 0)  # [r, bins]
    u = np.broadcast_to(O.sample_u(128, draws), (r, 128))
    near = np.zeros((r, 128), dtype=bool)
    for i in range(r):
        levels = cw[i, :-1][flat[i]]
        if levels.size:
            near[i] = (np.abs(u[i][:, None] - levels[None]) < 1e-5).any(-1)
    bad = np.zeros((r, 129), dtype=bool)
    bad[:, :-1] |= near
    bad[:, 1:] |= near
    bad[:, 0] |= near[:, 1]
    bad[:, -1] |= near[:, -2]
    assert bad.mean() < 0.01
    # elsewhere a centre moves by the CDF error times the inverse CDF's slope (bin width /
    # bin weight), which the dilated low-density pieces make large: scale the tolerance by it
    m2 = tdn.shape[-1] - 1
    slope = np.zeros((r, 128))
    for i in range(r):
        k = np.clip(np.searchsorted(cw[i], u[i], side='right'), 1, m2) - 1
        slope[i] = (tdn[i, k + 1] - tdn[i, k]) / np.maximum(cw[i, k + 1] - cw[i, k], 1e-30)
    sf = np.zeros((r, 129))
    sf[:, 1:-1] = np.maximum(slope[:, 1:], slope[:, :-1])
    sf[:, 0] = np.maximum(slope[:, 0], slope[:, 1])
    sf[:, -1] = np.maximum(slope[:, -1], slope[:, -2])
    err = np.abs(out.double().cpu().numpy() - want)
    assert np.all((err <= 1e-5 * (1 + sf)) | bad)
    assert (err < 1e-5).mean() > 0.99

"""Pins of the oracle's eval-time ray-march extras (volumetric_rendering with compute_extras)."""
import numpy as np
import torch

from oracle import compose as ocompose

D64 = torch.float64


def test_sorted_interp_is_numpy_interp_on_strictly_increasing_points():
    g = torch.Generator().manual_seed(0)
    for _ in range(20):
        xp = torch.cumsum(torch.rand(9, generator=g, dtype=D64) + 0.01, 0)
        fp = torch.cumsum(torch.rand(9, generator=g, dtype=D64), 0)
        x = torch.sort(torch.rand(6, generator=g, dtype=D64) * (xp[-1] - xp[0] + 2) + xp[0] - 1).values
        got = ocompose.sorted_interp(x, xp, fp)
        exp = np.interp(x.numpy(), xp.numpy(), fp.numpy())  # clamps outside [xp0, xp_last] like the mask rule
        assert np.allclose(got.numpy(), exp, atol=1e-12)


def _rays(n, R, seed):
    g = torch.Generator().manual_seed(seed)
    td = torch.sort(torch.rand(R, n + 1, generator=g, dtype=D64) * 4 + 2, -1).values
    far = td[:, -1] + 0.5
    return td, far


def test_one_opaque_sample_puts_every_percentile_inside_its_interval():
    n, R = 16, 5
    td, far = _rays(n, R, 1)
    w = torch.zeros(R, n, dtype=D64)
    ks = torch.tensor([0, 3, 7, 11, 15])
    w[torch.arange(R), ks] = 1.0
    ex = ocompose.volumetric_extras(w, td, far)
    a, b = td[torch.arange(R), ks], td[torch.arange(R), ks + 1]
    for p, key in ((0.05, 'distance_percentile_5'), (0.5, 'distance_median'), (0.95, 'distance_percentile_95')):
        assert torch.allclose(ex[key], a + p * (b - a), atol=1e-12)
    assert torch.allclose(ex['distance_mean'], (a + b) / 2, atol=1e-12)


def test_empty_ray_uses_the_far_distance():
    n, R = 12, 4
    td, far = _rays(n, R, 2)
    ex = ocompose.volumetric_extras(torch.zeros(R, n, dtype=D64), td, far)
    for p, key in ((0.05, 'distance_percentile_5'), (0.5, 'distance_median'), (0.95, 'distance_percentile_95')):
        assert torch.allclose(ex[key], td[:, -1] + p * (far - td[:, -1]), atol=1e-12)
    assert torch.allclose(ex['distance_mean'], td[:, 0])  # exp(0) = 1 clipped up to the first fencepost


def test_uniform_weights_give_the_geometric_mean_depth():
    n, R = 32, 3
    td, far = _rays(n, R, 3)
    w = torch.full((R, n), 0.5 / n, dtype=D64)
    mids = (td[:, 1:] + td[:, :-1]) / 2
    ex = ocompose.volumetric_extras(w, td, far)
    assert torch.allclose(ex['distance_mean'], torch.exp(torch.log(mids).mean(-1)), atol=1e-12)
    # half the weight in the samples, half on [td_n, far]: the median is td_n
    assert torch.allclose(ex['distance_median'], td[:, -1], atol=1e-12)
    # 5th percentile: 10% of the way through the samples' uniform cumulative weight
    exp5 = torch.tensor([np.interp(0.05, np.concatenate([[0], np.cumsum(w[r].numpy())]), td[r].numpy())
                         for r in range(R)], dtype=D64)
    assert torch.allclose(ex['distance_percentile_5'], exp5, atol=1e-12)

"""CPU pins of the interval-sampler oracle (oracle/stepfun.py): numpy.interp (the
reference's own invert_cdf_np, stepfun.py:205-213) and closed forms."""
import numpy as np
import pytest

from oracle import stepfun as O


def _rays(seed, r, m, zero_bins=True):
    g = np.random.default_rng(seed)
    t = np.sort(g.uniform(0, 1, (r, m + 1)), -1)
    t[:, 0], t[:, -1] = 0.0, 1.0
    logits = g.standard_normal((r, m))
    if zero_bins:
        logits[g.uniform(size=(r, m)) < 0.2] = -np.inf
    return t, logits


@pytest.mark.parametrize('m,n', [(1, 2), (7, 5), (64, 33), (200, 64)])
def test_invert_cdf_is_numpy_interp(m, n):
    t, logits = _rays(m, 6, m)
    u = np.sort(np.random.default_rng(1).uniform(0, 1, (6, n)), -1)
    got = O.invert_cdf(u, t, logits)
    cw = O.integrate_weights(O.softmax(logits))
    for r in range(6):
        # numpy.interp takes the last of repeated cw values (zero-weight bins), as the
        # masked max / min does for sorted t
        np.testing.assert_allclose(got[r], np.interp(u[r], cw[r], t[r]), rtol=0, atol=1e-12)


def test_closed_forms():
    m, n = 16, 9
    t = np.linspace(2.0, 6.0, m + 1)[None]
    # uniform weights on uniform bins: centres are t0 + u (t_m - t0)
    u = O.sample_u(n)
    c = O.invert_cdf(u[None], t, np.zeros((1, m)))
    np.testing.assert_allclose(c[0], 2.0 + 4.0 * u, atol=1e-12)
    s = O.sample_intervals(u[None], t, np.zeros((1, m)))
    np.testing.assert_allclose(s[0, 1:-1], (c[0, 1:] + c[0, :-1]) / 2)
    np.testing.assert_allclose(s[0, 0], c[0, 0] - (c[0, 1] - c[0, 0]) / 2)
    np.testing.assert_allclose(s[0, -1], c[0, -1] + (c[0, -1] - c[0, -2]) / 2)
    # all mass in one bin: every centre lies in it, spread by u
    lg = np.full((1, m), -np.inf)
    lg[0, 5] = 0.0
    c = O.invert_cdf(u[None], t, lg)
    np.testing.assert_allclose(c[0], t[0, 5] + u * (t[0, 6] - t[0, 5]), atol=1e-12)
    # domain clamp of the reflected ends
    s = O.sample_intervals(u[None], t, lg, domain=(t[0, 5] + 1e-3, t[0, 6] - 1e-3))
    assert s[0, 0] == t[0, 5] + 1e-3 and s[0, -1] == t[0, 6] - 1e-3
    # sample positions of the reference: linspace (no jitter) and jittered within [0, 1)
    u4 = O.sample_u(4)
    eps = float(np.finfo(np.float32).eps)
    assert u4[0] == 1 / 8 and abs(u4[-1] - (7 / 8 - eps)) < 1e-15 and np.allclose(np.diff(u4, 2), 0, atol=1e-15)
    uj = O.sample_u(4, np.array([[0.0], [1.0 - 1e-9]]))
    assert uj.min() >= 0 and uj.max() < 1 and np.all(np.diff(uj, axis=-1) > 0)


def test_outputs_sorted_and_zero_width_bins_get_no_centres():
    t, logits = _rays(3, 20, 50)
    zero = np.zeros_like(t[:, :-1], dtype=bool)
    t[:, 10] = t[:, 9]  # a zero-width bin
    zero[:, 9] = True
    logits[zero] = -np.inf
    u = O.sample_u(40, np.random.default_rng(2).uniform(0, 1, (20, 1)))
    s = O.sample_intervals(u, t, logits, domain=(0.0, 1.0))
    assert np.all(np.diff(s, axis=-1) >= 0) and s.min() >= 0 and s.max() <= 1


def test_max_dilate_closed_forms():
    # no dilation keeps the step function: every weight survives on its own bin
    t, _ = _rays(4, 5, 12, zero_bins=False)
    w = np.random.default_rng(5).uniform(0.1, 1, (5, 12))
    w /= w.sum(-1, keepdims=True)
    td, wd = O.max_dilate_weights(t, w, 0.0, renormalize=True)
    assert td.shape == (5, 37) and wd.shape == (5, 36)
    np.testing.assert_allclose(wd.sum(-1), 1.0)
    for r in range(5):
        keep = np.diff(td[r]) > 0
        np.testing.assert_allclose(wd[r][keep], w[r], rtol=1e-12)
    # one bin [0, 1] dilated by 0.25 inside the domain [0, 1]: all mass stays on [0, 1]
    td, wd = O.max_dilate_weights(np.array([[0.0, 1.0]]), np.array([[1.0]]), 0.25, domain=(0.0, 1.0),
                                  renormalize=True)
    np.testing.assert_array_equal(td, [[0.0, 0.0, 1.0, 1.0]])
    np.testing.assert_array_equal(wd, [[0.0, 1.0, 0.0]])
    # max dilation: the dilated pdf dominates the original pdf at every original edge
    t, _ = _rays(6, 3, 20, zero_bins=False)
    w = np.random.default_rng(7).uniform(0, 1, (3, 20))
    p = w / np.diff(t, axis=-1)
    td, wd = O.max_dilate_weights(t, w, 0.03)
    pd = wd / np.maximum(np.diff(td, axis=-1), 1e-30)
    for r in range(3):
        for j in range(20):
            mid = 0.5 * (t[r, j] + t[r, j + 1])
            i = np.searchsorted(td[r], mid, side='right') - 1
            assert pd[r, i] >= p[r, j] * (1 - 1e-12)
    # total mass never shrinks without renormalisation
    assert np.all(wd.sum(-1) >= w.sum(-1) * (1 - 1e-12))

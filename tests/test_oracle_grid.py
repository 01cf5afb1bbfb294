"""CPU pins of the hash-grid oracle (oracle/grid.py).

Checks: the level table of the 16-level 2^19 configuration, the prime-XOR hash
against Python integers, the dense index x + y R + z R^2 (and its uint32
stride wrap), interpolation identities (a lattice vertex returns its row,
constant tables give constants, trilinear interpolation reproduces linear
functions), the backward as the exact transpose of the forward, input
gradients against finite differences, out-of-range inputs, and the TV
gradient of a constant table.
"""
import math

import numpy as np
import pytest

from oracle import grid as G


def test_layout_16_levels_2e19():
    offsets, res, pls = G.grid_layout(3, 16, 2, base_resolution=16, log2_hashmap_size=19, desired_resolution=2048)
    assert abs(pls - 2 ** (math.log2(2048 / 16) / 15)) < 1e-15
    assert res[:5] == [17, 24, 32, 44, 60]
    assert res[-1] == 2049
    assert int(offsets[-1]) == 6_119_864
    sizes = np.diff(offsets)
    for r, n in zip(res, sizes):
        assert n == int(math.ceil(min(2 ** 19, r ** 3) / 8) * 8)
    assert offsets.dtype == np.int32


def test_hash_matches_python_integers():
    rng = np.random.default_rng(0)
    pos = rng.integers(0, 2 ** 31, size=(500, 3)).astype(np.uint32)
    h = G.fast_hash(pos)
    for p, v in zip(pos.tolist(), h.tolist()):
        assert v == ((p[0] * 1) ^ (p[1] * 2654435761) ^ (p[2] * 805459861)) & 0xFFFFFFFF


def test_dense_and_hashed_index():
    rng = np.random.default_rng(1)
    R = 17  # resolution + 1 lattice points per axis, R^3 <= T: dense
    pos = rng.integers(0, R, size=(300, 3)).astype(np.uint32)
    idx = G.grid_index(0, False, 5000, R - 1, pos)
    np.testing.assert_array_equal(idx, (pos[:, 0] + pos[:, 1] * R + pos[:, 2] * R * R) % 5000)
    # R^3 > T: hashed level
    pos = rng.integers(0, 100, size=(300, 3)).astype(np.uint32)
    idx = G.grid_index(0, False, 4096, 99, pos)
    np.testing.assert_array_equal(idx, G.fast_hash(pos).astype(np.int64) % 4096)
    # tiled grid type never hashes: dimensions stop contributing once the stride exceeds T
    idx = G.grid_index(1, False, 4096, 99, pos)
    np.testing.assert_array_equal(idx, (pos[:, 0] + pos[:, 1] * 100) % 4096)


def test_stride_wraps_in_uint32():
    """Resolution + 1 = 65537: stride 65537^2 wraps to 131073 in uint32, which is <= T = 2^19,
    so the level keeps a (wrapped) dense index instead of hashing."""
    R = 65537
    assert (R * R) % 2 ** 32 == 131073
    pos = np.array([[5, 7, 9]], dtype=np.uint32)
    idx = G.grid_index(0, False, 2 ** 19, R - 1, pos)
    expect = ((5 + 7 * R + 9 * ((R * R) % 2 ** 32)) % 2 ** 32) % 2 ** 19
    assert int(idx[0]) == expect


def _setup(levels=4, C=2, base=4, pls=1.5, log2T=10, seed=0):
    offsets, res, _ = G.grid_layout(3, levels, C, per_level_scale=pls, base_resolution=base, log2_hashmap_size=log2T)
    rng = np.random.default_rng(seed)
    emb = rng.uniform(-1, 1, size=(int(offsets[-1]), C))
    return offsets, res, emb, rng


def test_vertex_query_returns_table_row():
    offsets, res, emb, rng = _setup(levels=1, base=9)
    scale = float(G.level_scale_f32(0, np.float32(np.log2(1.5)), 9))
    R = int(math.ceil(scale)) + 1
    k = rng.integers(0, R - 1, size=(50, 3))
    x = k / scale  # align_corners: pos = x * scale lands on the vertex
    out, _ = G.grid_forward(x, emb, offsets, 1.5, 9, align_corners=True)
    assert R ** 3 <= int(offsets[1])
    np.testing.assert_allclose(out, emb[k[:, 0] + k[:, 1] * R + k[:, 2] * R * R], atol=1e-12)


def test_partition_of_unity_and_linear_reproduction():
    offsets, res, emb, rng = _setup(levels=2, base=4, log2T=14)
    x = rng.uniform(0, 1, size=(200, 3))
    emb_c = np.full_like(emb, 0.75)
    out, _ = G.grid_forward(x, emb_c, offsets, 1.5, 4)
    np.testing.assert_allclose(out, 0.75, atol=1e-12)
    # dense level 0: table(v) = a . v + b  ->  output = a . pos + b
    S = np.float32(np.log2(1.5))
    scale = float(G.level_scale_f32(0, S, 4))
    R = int(math.ceil(scale)) + 2  # lattice points per axis without align_corners: resolution + 1
    assert R ** 3 <= offsets[1]
    a, b = np.array([0.3, -0.2, 0.7]), 0.1
    lin = np.zeros_like(emb)
    v = np.stack(np.meshgrid(np.arange(R), np.arange(R), np.arange(R), indexing='ij'), -1).reshape(-1, 3)
    lin[v[:, 0] + v[:, 1] * R + v[:, 2] * R * R, 0] = v @ a + b
    out, _ = G.grid_forward(x, lin, offsets, 1.5, 4)
    np.testing.assert_allclose(out[:, 0], (x * scale + 0.5) @ a + b, atol=1e-12)


def test_backward_is_transpose_of_forward():
    offsets, res, emb, rng = _setup(levels=3, base=4, log2T=8)
    x = rng.uniform(-0.05, 1.05, size=(300, 3))  # includes out-of-range rows
    g = rng.normal(size=(300, 3 * 2))
    out, _ = G.grid_forward(x, emb, offsets, 1.5, 4)
    gemb, _ = G.grid_backward(g, x, emb, offsets, 1.5, 4)
    # the forward is linear in the table: <g, F(e)> = <F^T g, e>
    assert abs((g * out).sum() - (gemb * emb).sum()) < 1e-10 * (1 + abs((g * out).sum()))
    oob = np.any((x < 0) | (x > 1), axis=1)
    assert oob.any() and np.all(out[oob] == 0)


@pytest.mark.parametrize('interp', [0, 1])
def test_input_gradient_matches_finite_differences(interp):
    offsets, res, emb, rng = _setup(levels=3, base=4, log2T=8, seed=2)
    x = rng.uniform(0.05, 0.95, size=(40, 3))
    g = rng.normal(size=(40, 3 * 2))
    out, dydx = G.grid_forward(x, emb, offsets, 1.5, 4, interp=interp, calc_dy_dx=True)
    _, gin = G.grid_backward(g, x, emb, offsets, 1.5, 4, interp=interp, dy_dx=dydx)
    h = 1e-7
    fd = np.zeros_like(x)
    for d in range(3):
        e = np.zeros_like(x)
        e[:, d] = h
        fp, _ = G.grid_forward(x + e, emb, offsets, 1.5, 4, interp=interp)
        fm, _ = G.grid_forward(x - e, emb, offsets, 1.5, 4, interp=interp)
        fd[:, d] = ((fp - fm) * g).sum(1) / (2 * h)
    # rows whose FD stencil crosses a cell boundary are excluded (a kink of trilinear interpolation)
    S = np.float32(np.log2(1.5))
    safe = np.ones(40, bool)
    for l in range(3):
        sc = float(G.level_scale_f32(l, S, 4))
        for s in (-h, h):
            safe &= np.all(np.floor((x + s) * sc + 0.5) == np.floor(x * sc + 0.5), axis=1)
    assert safe.sum() > 30
    np.testing.assert_allclose(gin[safe], fd[safe], rtol=1e-5, atol=1e-6)


def test_total_variation_of_constant_table_is_zero():
    offsets, res, emb, rng = _setup(levels=2, base=4, log2T=8)
    x = rng.uniform(0, 1, size=(50, 3))
    grad = rng.normal(size=emb.shape)
    out = G.grad_total_variation(x, np.full_like(emb, 0.3), grad, offsets, 0.5, 1.5, 4)
    np.testing.assert_allclose(out, grad, atol=1e-12)
    # a nonconstant table changes the gradient only at rows the queries touch
    out2 = G.grad_total_variation(x, emb, grad, offsets, 0.5, 1.5, 4)
    assert np.abs(out2 - grad).max() > 0

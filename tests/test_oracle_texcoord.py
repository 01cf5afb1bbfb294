"""Pins of the oracle's texture coordinates (``BaseSDF<...>::texcoord``, strokes_sdf.h:258-526).

Checked against what fixes them independently of the restated formulas:
closed-form points (poles and equator of the unit sphere, box corners,
capsule / line caps, quarter turns of the azimuth), invariance of the sphere
coordinates to the point's distance, the multiscale axis permutation as a
sort of the axes by decreasing scale (numpy's argsort), and the clamps.
"""
import math

import numpy as np
import torch

from oracle import strokes as ostrokes

D = torch.float64


def _p(*xyz):
    return [torch.tensor([float(c)], dtype=D) for c in xyz]


def _row(*vals):
    return torch.tensor([[float(v) for v in vals]], dtype=D)


def test_sphere_closed_forms():
    row = _row(0)
    cases = {(1, 0, 0): (1.0, 0.5), (-1, 0, 0): (0.0, 0.5), (0, 1, 0): (0.5, 0.5), (0, 0, 1): (0.5, 0.75),
             (0, 0, -1): (0.5, 0.25), (0, -1, 1e-300): (0.5, 1.0)}
    for xyz, (u, v) in cases.items():
        uu, vv = ostrokes.texcoord(0, _p(*xyz), row, False)
        assert abs(float(uu) - u) < 1e-12 and abs(float(vv) - v) < 1e-12, (xyz, float(uu), float(vv))
    # latitude measured from +x: u = 1 - angle(p, +x) / pi; azimuth of (y, z) around x
    g = torch.Generator().manual_seed(0)
    P = torch.randn(200, 3, generator=g, dtype=D)
    for scale in (0.3, 1.0, 7.5):
        u, v = ostrokes.texcoord(0, [P[:, i] * scale for i in range(3)], _row(0).expand(200, -1), False)
        ang = np.arccos(np.clip(P[:, 0].numpy() / np.linalg.norm(P.numpy(), axis=1), -1, 1))
        assert np.allclose(u.numpy(), 1.0 - ang / np.pi, atol=1e-12)
        az = np.angle(P[:, 1].numpy() + 1j * P[:, 2].numpy())
        assert np.allclose(v.numpy(), az / (2 * np.pi) + 0.5, atol=1e-12)


def test_box_and_round_box_closed_forms():
    for base in (1, 2):
        row = _row(0.2)
        assert [float(c) for c in ostrokes.texcoord(base, _p(-1, -1, 0), row, False)] == [0.0, 0.0]
        assert [float(c) for c in ostrokes.texcoord(base, _p(1, 1, 0.3), row, False)] == [1.0, 1.0]
        assert [float(c) for c in ostrokes.texcoord(base, _p(0, 0.5, 9), row, False)] == [0.5, 0.75]
        assert [float(c) for c in ostrokes.texcoord(base, _p(3, -4, 0), row, False)] == [1.0, 0.0]  # clamped


def test_capsule_and_line_closed_forms():
    h = 0.6
    # capsule: u runs over the whole length including the caps, y in [-(h+1), h+1]
    assert abs(float(ostrokes.texcoord(4, _p(0.3, h + 1, 0), _row(h), False)[0]) - 1.0) < 1e-12
    assert abs(float(ostrokes.texcoord(4, _p(0.3, -(h + 1), 0), _row(h), False)[0])) < 1e-12
    assert abs(float(ostrokes.texcoord(4, _p(0.3, 0, 0), _row(h), False)[0]) - 0.5) < 1e-12
    # azimuth of (x, z): +x -> 0.5, +z -> 0.75, -z -> 0.25
    for xz, v in (((1, 0), 0.5), ((0, 1), 0.75), ((0, -1), 0.25)):
        for base, row in ((4, _row(h)), (5, _row(h, 0.1))):
            vv = ostrokes.texcoord(base, _p(xz[0], 0.1, xz[1]), row, False)[1]
            assert abs(float(vv) - v) < 1e-12
    # line: u = (y + h + 1 - r_diff) / (2h + 2): 0 at y = -(h + 1) + r_diff
    rd = 0.2
    assert abs(float(ostrokes.texcoord(5, _p(0, -(h + 1) + rd, 0), _row(h, rd), False)[0])) < 1e-12
    assert abs(float(ostrokes.texcoord(5, _p(0, h + 1 + rd, 0), _row(h, rd), False)[0]) - 1.0) < 1e-12


def test_multiscale_permutation_sorts_axes_by_decreasing_scale():
    g = torch.Generator().manual_seed(1)
    S = torch.rand(500, 3, generator=g, dtype=D) + 0.05
    d = torch.randn(500, 3, generator=g, dtype=D)
    out = torch.stack(ostrokes.permute_multiscale_axis([d[:, i] for i in range(3)], [S[:, i] for i in range(3)]), -1)
    order = np.argsort(-S.numpy(), axis=1, kind='stable')
    want = np.take_along_axis(d.numpy(), order, axis=1)
    assert np.array_equal(out.numpy(), want)


def test_multiscale_reads_the_scale_where_the_reference_does():
    """Sphere / box read the scale at params + 0 (their base has no parameters, so it is the
    scale); the round box reads params + 0 too, i.e. (radius, s_x, s_y) (H14).  Capsule and
    line compare params[3] with params[1]: the scale's z and x for the capsule, the scale's y
    and r_diff for the line."""
    p = _p(0.3, 0.2, 0.1)
    # box with scale (0.1, 0.5, 0.2): the y axis first
    u, v = ostrokes.texcoord(1, p, _row(0.1, 0.5, 0.2), True)
    assert abs(float(u) - (0.2 * 0.5 + 0.5)) < 1e-12 and abs(float(v) - (0.1 * 0.5 + 0.5)) < 1e-12
    # round box row (r=0.9, s=(0.1, 0.5, 0.2)): the permutation sees (0.9, 0.1, 0.5)
    u, v = ostrokes.texcoord(2, p, _row(0.9, 0.1, 0.5, 0.2), True)
    assert abs(float(u) - (0.3 * 0.5 + 0.5)) < 1e-12 and abs(float(v) - (0.1 * 0.5 + 0.5)) < 1e-12
    # capsule row (h, sx, sy, sz): sz > sx swaps x and z in the azimuth
    _, v_sw = ostrokes.texcoord(4, _p(1, 0, 0), _row(0.5, 0.1, 0.2, 0.3), True)
    _, v_no = ostrokes.texcoord(4, _p(1, 0, 0), _row(0.5, 0.3, 0.2, 0.1), True)
    assert abs(float(v_sw) - 0.75) < 1e-12 and abs(float(v_no) - 0.5) < 1e-12


def test_bases_without_texture_coordinates_return_zero():
    for base in (3, 6, 7, 8, 9, 10, 11):
        u, v = ostrokes.texcoord(base, _p(0.3, -0.2, 0.5), _row(*range(14)), False)
        assert float(u) == 0.0 and float(v) == 0.0

"""CPU pins of the stroke oracle (oracle/strokes.py).

Each oracle function is checked against something other than its own
formula: closed-form distances at chosen points, brute-force Bezier
distances by dense parameter sampling, scipy's rotation and spherical-harmonic
routines, Gauss-Legendre quadrature (SH orthonormality), the Barry-Goldman
construction of centripetal Catmull-Rom curves, and central finite
differences in float64 for every exact adjoint.  The reference-adjoint
restatement is checked to equal the true derivative wherever the reference is
exact, and to deviate by the documented factor where it is not.
"""
import math

import numpy as np
import pytest
import scipy.special
import scipy.stats
import torch
from scipy.spatial.transform import Rotation

from helpers import points_near, stroke_params
from oracle import strokes as S

D64 = torch.float64


def _pts(*rows):
    t = torch.tensor(rows, dtype=D64)
    return [t[:, 0], t[:, 1], t[:, 2]]


def _no_prm(n):
    return torch.zeros(n, 0, dtype=D64)


# ----------------------------------------------------------------------------
# registry
# ----------------------------------------------------------------------------

# (sdf id, shape-param count) of the reference registry for the shapes the
# workloads use: id = base<<4 | T | R<<1 | SS<<2 | MS<<3
REGISTRY = {
    'sphere': ((0 << 4) | 1 | 4, 4),
    'ellipsoid': ((0 << 4) | 1 | 2 | 8, 9),
    'cube': ((1 << 4) | 1 | 2 | 4, 7),
    'obb': ((1 << 4) | 1 | 2 | 8, 9),
    'roundcube': ((2 << 4) | 1 | 2 | 4, 8),
    'roundbox': ((2 << 4) | 1 | 2 | 8, 10),
    'capsule': ((4 << 4) | 1 | 2 | 4, 8),
    'scapsule': ((4 << 4) | 1 | 2 | 8, 10),
    'octahedron': ((7 << 4) | 1 | 2 | 4, 7),
    'tetrahedron': ((8 << 4) | 1 | 2 | 4, 7),
    'quadratic_bezier': (9 << 4, 11),
    'cubic_bezier': (10 << 4, 14),
    'catmull_rom': (11 << 4, 14),
}


@pytest.mark.parametrize('shape', sorted(REGISTRY))
def test_registry(shape):
    sid, width = REGISTRY[shape]
    assert S.shape_sdf_id(shape) == sid
    assert S.num_shape_params(sid) == width
    assert S.decode_sdf_id(sid)[0] == sid >> 4


def test_color_widths():
    assert S.NUM_COLOR_PARAMS == [3, 12, 12, 27]
    assert [len(S.sh_basis(2, _pts((0, 0, 1)))), len(S.sh_basis(3, _pts((0, 0, 1))))] == [4, 9]


# ----------------------------------------------------------------------------
# rotation and transforms
# ----------------------------------------------------------------------------

def test_rotation_matches_scipy_extrinsic_xyz():
    g = torch.Generator().manual_seed(0)
    ang = (torch.rand(64, 3, generator=g, dtype=D64) * 2 - 1) * math.pi
    R = S.rotation_matrix(ang).numpy()
    ref = Rotation.from_euler('xyz', ang.numpy()).as_matrix()  # extrinsic x, then y, then z: Rz Ry Rx
    np.testing.assert_allclose(R, ref, atol=1e-14)
    np.testing.assert_allclose(np.einsum('nij,nkj->nik', R, R), np.broadcast_to(np.eye(3), R.shape), atol=1e-14)
    np.testing.assert_allclose(np.linalg.det(R), 1.0, atol=1e-13)
    np.testing.assert_allclose(S.rotation_matrix(torch.zeros(1, 3, dtype=D64))[0].numpy(), np.eye(3))


@pytest.mark.parametrize('shape', ['sphere', 'ellipsoid', 'cube', 'obb', 'capsule'])
def test_inverse_transform_round_trip(shape):
    sid = S.shape_sdf_id(shape)
    g = torch.Generator().manual_seed(1)
    n = 32
    P = S.num_shape_params(sid)
    prm = torch.rand(n, P, generator=g, dtype=D64) + 0.2
    p = torch.randn(n, 3, generator=g, dtype=D64)
    _, scale, angle, trans = S.split_shape_params(prm, sid)
    x = p * scale if scale is not None else p.clone()
    if angle is not None:
        x = torch.einsum('nij,nj->ni', S.rotation_matrix(angle), x)
    if trans is not None:
        x = x + trans
    back = torch.stack(S.inverse_transform(x, prm, sid), -1)
    torch.testing.assert_close(back, p, atol=1e-12, rtol=0)


def test_split_reads_transform_from_the_end():
    sid = S.shape_sdf_id('roundbox')  # base(1) + scale(3) + euler(3) + translation(3)
    prm = torch.arange(10, dtype=D64)[None]
    base, scale, angle, trans = S.split_shape_params(prm, sid)
    assert base.tolist() == [[0.0]]
    assert scale.tolist() == [[1.0, 2.0, 3.0]]
    assert angle.tolist() == [[4.0, 5.0, 6.0]]
    assert trans.tolist() == [[7.0, 8.0, 9.0]]


# ----------------------------------------------------------------------------
# base SDF closed forms
# ----------------------------------------------------------------------------

def test_sphere_is_distance_minus_one():
    g = torch.Generator().manual_seed(2)
    p = torch.randn(200, 3, generator=g, dtype=D64)
    p = p / p.norm(dim=-1, keepdim=True) * (0.3 + 3 * torch.rand(200, 1, generator=g, dtype=D64))
    s = S.sdf_sphere([p[:, 0], p[:, 1], p[:, 2]], _no_prm(200))
    torch.testing.assert_close(s, p.norm(dim=-1) - 1, atol=2e-8, rtol=0)


def test_cube_distance_brute_force():
    g = torch.Generator().manual_seed(3)
    p = (torch.rand(400, 3, generator=g, dtype=D64) * 2 - 1) * 2.5
    s = S.sdf_cube([p[:, 0], p[:, 1], p[:, 2]], _no_prm(400))
    inside = (p.abs() <= 1).all(-1)
    outside_d = (p - p.clamp(-1, 1)).norm(dim=-1)           # closest point of a box is the clamp
    inside_d = -(1 - p.abs()).min(-1).values                 # nearest face
    torch.testing.assert_close(s, torch.where(inside, inside_d, outside_d), atol=1e-12, rtol=0)


def test_solid_values_at_axis_points():
    c = 1 / math.sqrt(3)
    # octahedron |x|+|y|+|z| = 1: distance along a face normal
    s = S.sdf_octahedron(_pts((2, 0, 0), (1, 1, 1), (0.1, 0.1, 0.1)), _no_prm(3))
    torch.testing.assert_close(s, torch.tensor([c, 2 / math.sqrt(3), -0.7 * c], dtype=D64), atol=1e-8, rtol=0)
    s = S.sdf_tetrahedron(_pts((0, 0, 2), (0, 0, -2), (0, 0, 0)), _no_prm(3))
    torch.testing.assert_close(s, torch.tensor([c, c, -c], dtype=D64), atol=1e-8, rtol=0)
    # capsule of half-height h and unit radius: distance to the segment minus 1
    h = torch.full((4, 1), 0.7, dtype=D64)
    s = S.sdf_capsule(_pts((0, 2.7, 0), (3, 0.35, 0), (0, -1.7, 0), (0.5, 0, 0)), h)
    torch.testing.assert_close(s, torch.tensor([1.0, 2.0, 0.0, -0.5], dtype=D64), atol=1e-7, rtol=0)
    # round cube = box distance minus the rounding radius
    r = torch.full((2, 1), 0.2, dtype=D64)
    s = S.sdf_round_cube(_pts((3, 0, 0), (2, 2, 1)), r)
    torch.testing.assert_close(s, torch.tensor([1.8, math.sqrt(2) - 0.2], dtype=D64), atol=1e-7, rtol=0)
    # line: capsule whose radius goes 1-rd .. 1+rd along y in [-h, h]
    prm = torch.tensor([[1.0, 0.25]] * 3, dtype=D64)
    s = S.sdf_line(_pts((2, -1, 0), (2, 1, 0), (0, 3, 0)), prm)
    torch.testing.assert_close(s, torch.tensor([2 - 0.75, 2 - 1.25, 2 - 1.25], dtype=D64), atol=1e-7, rtol=0)
    # triangular prism: below the base plane the z term dominates
    s = S.sdf_triprism(_pts((0, 0, -2)), torch.tensor([[0.5]], dtype=D64))
    torch.testing.assert_close(s, torch.tensor([1.5], dtype=D64), atol=1e-12, rtol=0)


def _bezier_quad(t, A, B, C):
    u = 1 - t
    return u[:, None] ** 2 * A + 2 * (u * t)[:, None] * B + t[:, None] ** 2 * C


def test_quadratic_bezier_distance_brute_force():
    """Closed form vs dense sampling (20001 parameters) and the exact stationary points.

    One-root branch: the closed form is the true distance.  Three-root branch:
    the reference evaluates the smallest root of the stationarity cubic and the
    candidate (m + n) z - kx, which is not a root (Inigo Quilez's sdBezier uses
    2 m z - kx); the result is never below the true distance and equals it
    whenever the true closest point is the smallest root.
    """
    g = torch.Generator().manual_seed(4)
    n = 256
    ctrl = torch.rand(n, 9, generator=g, dtype=D64) * 2 - 1
    p = torch.rand(n, 3, generator=g, dtype=D64) * 2 - 1
    d2, t = S.quad_bezier_dist2_t([p[:, 0], p[:, 1], p[:, 2]], ctrl)
    ts = torch.linspace(0, 1, 20001, dtype=D64)
    A, B, C = ctrl[:, 0:3], ctrl[:, 3:6], ctrl[:, 6:9]
    best = torch.full((n,), float('inf'), dtype=D64)
    tbest = torch.zeros(n, dtype=D64)
    for chunk in ts.split(2001):
        pts = (1 - chunk)[None, :, None] ** 2 * A[:, None] + 2 * ((1 - chunk) * chunk)[None, :, None] * B[:, None] \
            + chunk[None, :, None] ** 2 * C[:, None]
        dd = ((pts - p[:, None]) ** 2).sum(-1)
        v, i = dd.min(-1)
        better = v < best
        best = torch.where(better, v, best)
        tbest = torch.where(better, chunk[i], tbest)
    # t is the parameter of the point whose distance is reported
    torch.testing.assert_close(((_bezier_quad(t, A, B, C) - p) ** 2).sum(-1), d2, atol=1e-12, rtol=0)
    assert (d2.sqrt() >= best.sqrt() - 1e-7).all()  # never closer than the curve
    # exact stationary points: d/dt |d + c t + b t^2|^2 / 2 is a cubic in t
    a, b = B - A, A - 2 * B + C
    c, dv = 2 * a, A - p
    dot = lambda u, w: (u * w).sum(-1)
    coef = torch.stack([2 * dot(b, b), 3 * dot(b, c), dot(c, c) + 2 * dot(b, dv), dot(c, dv)], -1).numpy()
    n_real = np.array([(np.abs(np.roots(cf).imag) < 1e-9).sum() for cf in coef])
    smallest = torch.tensor([np.roots(cf).real.min() if k == 3 else np.nan for cf, k in zip(coef, n_real)])
    one = torch.tensor(n_real == 1)
    assert one.sum() > 20 and (~one).sum() > 20
    err = (d2.sqrt() - best.sqrt()).abs()
    # near p = 0 the cube roots amplify the 1e-8 regulariser inside sqrt(h + 1e-8)
    pcoef = S._quad_setup([p[:, 0], p[:, 1], p[:, 2]], ctrl)[7]
    assert err[one & (pcoef.abs() > 0.01)].max() < 2e-5
    assert err[one].max() < 1e-3
    near_smallest = (~one) & ((tbest - smallest.clamp(0, 1)).abs() < 1e-3)
    assert near_smallest.sum() > 10
    assert err[near_smallest].max() < 2e-5


def test_cubic_bezier_straight_line_is_exact():
    """Evenly spaced collinear control points: the curve is its own polyline."""
    g = torch.Generator().manual_seed(5)
    n = 64
    a = torch.rand(n, 3, generator=g, dtype=D64) * 2 - 1
    b = torch.rand(n, 3, generator=g, dtype=D64) * 2 - 1
    prm = torch.cat([a, a + (b - a) / 3, a + 2 * (b - a) / 3, b, torch.full((n, 2), 0.05, dtype=D64)], -1)
    p = torch.rand(n, 3, generator=g, dtype=D64) * 2 - 1
    s = S.sdf_cubic_bezier([p[:, 0], p[:, 1], p[:, 2]], prm)
    ab = b - a
    tt = ((p - a) * ab).sum(-1) / (ab * ab).sum(-1)
    seg = ((a + tt.clamp(0, 1)[:, None] * ab) - p).norm(dim=-1)
    torch.testing.assert_close(s, seg - 0.05, atol=1e-12, rtol=0)
    t = S.cubic_bezier_t([p[:, 0], p[:, 1], p[:, 2]], prm)
    torch.testing.assert_close(t, tt.clamp(0, 1), atol=1e-12, rtol=0)


def test_cubic_bezier_picks_closest_polyline_segment():
    g = torch.Generator().manual_seed(6)
    n = 48
    prm = torch.cat([torch.rand(n, 12, generator=g, dtype=D64) * 2 - 1, torch.zeros(n, 2, dtype=D64)], -1)
    p = torch.rand(n, 3, generator=g, dtype=D64) * 2 - 1
    pl = [p[:, 0], p[:, 1], p[:, 2]]
    t = S.cubic_bezier_t(pl, prm)
    # brute force over the 10-segment polyline (each segment densely sampled)
    knots = torch.linspace(0, 1, 11, dtype=D64)
    A, B, C, Dp = prm[:, 0:3], prm[:, 3:6], prm[:, 6:9], prm[:, 9:12]

    def bez(tv):
        u = 1 - tv
        return (u ** 3)[..., None] * A[:, None] + (3 * u * u * tv)[..., None] * B[:, None] + \
            (3 * u * tv * tv)[..., None] * C[:, None] + (tv ** 3)[..., None] * Dp[:, None]

    kp = bez(knots[None].expand(n, 11))                     # [n, 11, 3]
    s = torch.linspace(0, 1, 4001, dtype=D64)
    seg_pts = kp[:, :-1, None] + s[None, None, :, None] * (kp[:, 1:] - kp[:, :-1])[:, :, None]
    dist = (seg_pts - p[:, None, None]).norm(dim=-1)        # [n, 10, 4001]
    flat = dist.reshape(n, -1)
    idx = flat.argmin(-1)
    seg, si = idx // 4001, idx % 4001
    t_brute = (seg + s[si]) / 10
    sel = S.curve_selection_margin(10, pl, prm) > 1e-3      # skip near-ties between segments
    assert sel.sum() > n // 2
    assert (t[sel] - t_brute[sel]).abs().max() < 1e-3 / 10 + 1e-9


def test_catmull_rom_is_barry_goldman_centripetal():
    g = torch.Generator().manual_seed(7)
    n = 16
    pts = torch.rand(n, 4, 3, generator=g, dtype=D64) * 2 - 1
    prm = torch.cat([pts.reshape(n, 12), torch.zeros(n, 2, dtype=D64)], -1)
    A, B, C, Dc = S.catmull_coeffs(prm)
    P = [pts[:, i] for i in range(4)]
    knots = [torch.zeros(n, dtype=D64)]
    for i in range(3):
        knots.append(knots[-1] + (P[i + 1] - P[i]).norm(dim=-1).sqrt())
    for u in [0.0, 0.13, 0.5, 0.77, 1.0]:
        tt = knots[1] + u * (knots[2] - knots[1])

        def lerp(pa, pb, ta, tb):
            return ((tb - tt) / (tb - ta))[:, None] * pa + ((tt - ta) / (tb - ta))[:, None] * pb

        a1 = lerp(P[0], P[1], knots[0], knots[1])
        a2 = lerp(P[1], P[2], knots[1], knots[2])
        a3 = lerp(P[2], P[3], knots[2], knots[3])
        b1 = lerp(a1, a2, knots[0], knots[2])
        b2 = lerp(a2, a3, knots[1], knots[3])
        c = lerp(b1, b2, knots[1], knots[2])
        uu = torch.full((n,), u, dtype=D64)
        ours = torch.stack(S._catmull_point(uu, A, B, C, Dc), -1)
        torch.testing.assert_close(ours, c, atol=1e-6, rtol=0)  # 1e-8 regulariser inside the knot sqrt


# ----------------------------------------------------------------------------
# colour fields, SH and alpha transfer
# ----------------------------------------------------------------------------

def _real_sh(l, m, x, y, z):
    """Real SH that keep the Condon-Shortley phase of scipy's complex SH (the
    graphics convention: Y_1,-1 = -c y, Y_1,1 = -c x)."""
    polar = np.arccos(np.clip(z, -1, 1))
    azim = np.arctan2(y, x)
    Y = scipy.special.sph_harm_y(l, abs(m), polar, azim)
    if m == 0:
        return Y.real
    return math.sqrt(2) * (Y.imag if m < 0 else Y.real)


def test_sh_basis_matches_real_spherical_harmonics():
    g = np.random.default_rng(8)
    d = g.normal(size=(100, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ours = torch.stack(S.sh_basis(3, [torch.tensor(d[:, i]) for i in range(3)]), -1).numpy()
    order = [(0, 0), (1, -1), (1, 0), (1, 1), (2, -2), (2, -1), (2, 0), (2, 1), (2, 2)]
    for j, (l, m) in enumerate(order):
        np.testing.assert_allclose(ours[:, j], _real_sh(l, m, d[:, 0], d[:, 1], d[:, 2]), atol=1e-12)


def test_sh_basis_is_orthonormal():
    zs, wz = np.polynomial.legendre.leggauss(12)
    phi = np.arange(24) * 2 * math.pi / 24
    Z, PH = np.meshgrid(zs, phi, indexing='ij')
    W = np.meshgrid(wz, np.full(24, 2 * math.pi / 24), indexing='ij')
    W = W[0] * W[1]
    r = np.sqrt(1 - Z ** 2)
    d = [torch.tensor((r * np.cos(PH)).ravel()), torch.tensor((r * np.sin(PH)).ravel()), torch.tensor(Z.ravel())]
    Y = torch.stack(S.sh_basis(3, d), -1).numpy()
    gram = (Y * W.ravel()[:, None]).T @ Y
    np.testing.assert_allclose(gram, np.eye(9), atol=1e-12)


def test_laplace_cdf_matches_scipy():
    s = torch.linspace(-30, 30, 1001, dtype=D64)
    np.testing.assert_allclose(S.laplace_cdf(s).numpy(), scipy.stats.laplace.cdf(s.numpy()), rtol=1e-13, atol=1e-300)


def test_gradient_rgb_endpoints():
    """The gradient colour is sigmoid(5 t' - 2.5) between c0 and c1 with t' the projection on p0->p1."""
    cp = torch.tensor([[0, 0, 0, 1, 0, 0, 0.1, 0.2, 0.3, 0.9, 0.8, 0.7]], dtype=D64)
    pos = _pts((0, 0, 0), (1, 0, 0), (0.5, 3, 0))
    col = S.color_forward(1, None, pos, None, None, cp)
    sg = lambda v: 1 / (1 + math.exp(-v))
    for row, tp in zip(col, [0.0, 1.0, 0.5]):
        w = sg(5 * tp - 2.5)
        torch.testing.assert_close(row, torch.tensor([0.1 + 0.8 * w, 0.2 + 0.6 * w, 0.3 + 0.4 * w], dtype=D64))


def test_alpha_transfer_closed_form():
    """Sphere stroke (centre c, scale s) at world distance rho: alpha from the sdf rho/s - 1."""
    sid = S.shape_sdf_id('sphere')
    c = torch.tensor([0.3, -0.2, 0.1], dtype=D64)
    s = 0.5
    prm = torch.tensor([[s, *c.tolist()]], dtype=D64)
    rho = torch.linspace(0.3, 0.7, 41, dtype=D64)
    x = c + rho[:, None] * torch.tensor([0.0, 0.6, 0.8], dtype=D64)
    r = torch.full((41,), 0.01, dtype=D64)
    vd = torch.tensor([[0.0, 0.0, 1.0]], dtype=D64).expand(41, 3)
    cp = torch.tensor([[0.2, 0.4, 0.6]], dtype=D64)
    sdf = rho / s - 1
    for laplace in (False, True):
        out = S.stroke_forward(x, r, vd, prm, cp, sid, 0, 1.0, laplace)
        if laplace:
            expect = torch.tensor(scipy.stats.laplace.cdf((-sdf * 2 / 0.01).numpy()), dtype=D64)
        else:
            expect = (-sdf * 0.5 / 0.01 + 0.5).clamp(0, 0.9999)
        torch.testing.assert_close(out['alpha'][:, 0], expect, atol=1e-6, rtol=0)
        torch.testing.assert_close(out['color'][:, 0], cp.expand(41, 3))


# ----------------------------------------------------------------------------
# adjoints
# ----------------------------------------------------------------------------

SHAPES = ['sphere', 'ellipsoid', 'aacube', 'cube', 'aabb', 'obb', 'roundcube', 'roundbox', 'cappedtorus', 'capsule',
          'scapsule', 'line', 'triprism', 'octahedron', 'tetrahedron', 'quadratic_bezier', 'cubic_bezier',
          'catmull_rom']


def _setup(shape, color, seed, n_strokes=3, n_rays=4, n_samples=10):
    sp, cp = stroke_params(shape, color, n_strokes, seed)
    x, vd = points_near(shape, sp, n_rays, n_samples, seed)
    return sp.double(), cp.double(), x.reshape(-1, 3).double(), torch.full((n_rays * n_samples,), 0.05, dtype=D64), \
        vd.double()


def _fd_grad(fn, params, h=1e-6):
    out = torch.zeros_like(params)
    flat = params.reshape(-1)
    for i in range(flat.numel()):
        e = torch.zeros_like(flat)
        e[i] = h
        out.view(-1)[i] = (fn((flat + e).view_as(params)) - fn((flat - e).view_as(params))) / (2 * h)
    return out


@pytest.mark.parametrize('shape', SHAPES)
@pytest.mark.parametrize('laplace', [True, False])
def test_exact_adjoint_matches_finite_differences(shape, laplace):
    color = 'gradient_rgb' if shape == 'ellipsoid' else 'constant_rgb'
    cid = S.COLOR_TYPES[color]
    sid = S.shape_sdf_id(shape)
    sp, cp, x, r, vd = _setup(shape, color, 11)
    g = torch.Generator().manual_seed(12)
    ga = torch.randn(x.shape[0], sp.shape[0], generator=g, dtype=D64)
    gc = torch.randn(x.shape[0], sp.shape[0], 3, generator=g, dtype=D64)

    def loss(spv, cpv=cp, xv=x):
        fw = S.stroke_forward(xv, r, vd, spv, cpv, sid, cid, 1.0, laplace)
        return (fw['alpha'] * ga).sum() + (fw['color'] * gc).sum()

    gs, gcp, gx = S.stroke_backward_exact(ga, gc, None, x, r, vd, sp, cp, sid, cid, 1.0, laplace)
    fd = _fd_grad(loss, sp)
    assert (gs - fd).norm() <= 1e-6 * fd.norm() + 1e-9, ((gs - fd).norm(), fd.norm())
    fdx = _fd_grad(lambda xv: loss(sp, cp, xv), x)
    assert (gx - fdx).norm() <= 1e-6 * fdx.norm() + 1e-9
    fdc = _fd_grad(lambda cv: loss(sp, cv), cp)
    assert (gcp - fdc).norm() <= 1e-6 * fdc.norm() + 1e-9


@pytest.mark.parametrize('shape', ['sphere', 'ellipsoid', 'capsule', 'scapsule', 'roundcube', 'octahedron',
                                   'tetrahedron', 'triprism'])
def test_reference_adjoint_exact_where_the_reference_is(shape):
    """Solids whose hand-written adjoint is the true derivative (constant colour, no radius rescale)."""
    sid = S.shape_sdf_id(shape)
    sp, cp, x, r, vd = _setup(shape, 'constant_rgb', 13, n_rays=8)
    g = torch.Generator().manual_seed(14)
    ga = torch.randn(x.shape[0], sp.shape[0], generator=g, dtype=D64)
    gc = torch.randn(x.shape[0], sp.shape[0], 3, generator=g, dtype=D64)
    alpha = S.stroke_forward(x, r, vd, sp, cp, sid, 0, 1.0, True)['alpha']
    rs, rc, rx = S.stroke_backward_reference(ga, gc, None, x, r, vd, alpha, sp, cp, sid, 0, 1.0, True)
    es, ec, ex = S.stroke_backward_exact(ga, gc, None, x, r, vd, sp, cp, sid, 0, 1.0, True)
    torch.testing.assert_close(rs, es, rtol=1e-9, atol=1e-9)
    torch.testing.assert_close(rc, ec, rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(rx, ex, rtol=1e-9, atol=1e-9)


def test_cube_reference_adjoint_outside_is_squared_length():
    g = torch.Generator().manual_seed(15)
    p = (torch.rand(300, 3, generator=g, dtype=D64) * 2 - 1) * 2.0
    pl = [p[:, i].clone().requires_grad_(True) for i in range(3)]
    s = S.sdf_cube(pl, _no_prm(300))
    ex = torch.stack(torch.autograd.grad(s.sum(), pl), -1)
    ref = torch.stack(S.grad_sdf_reference(1, [p[:, i] for i in range(3)], None)[0], -1)
    inside = (p.abs() < 1).all(-1)
    outside_len = (p.abs() - 1).clamp_min(0).norm(dim=-1)
    torch.testing.assert_close(ref[inside], ex[inside], atol=1e-12, rtol=0)
    torch.testing.assert_close(ref[~inside], ex[~inside] * 2 * outside_len[~inside, None], atol=1e-12, rtol=1e-12)


def test_cubic_reference_grad_x_is_squared_distance_gradient():
    """On a straight evenly parametrised curve (exact polyline, equal radii) the
    reference's dL/dx is g^2 * 2 (x - P) for upstream dL/dsdf = g (the gradient
    of dist^2, with the upstream factor applied twice), the true one g (x - P) / dist."""
    g = torch.Generator().manual_seed(16)
    S_ = 4
    a = torch.rand(S_, 3, generator=g, dtype=D64) * 2 - 1
    b = torch.rand(S_, 3, generator=g, dtype=D64) * 2 - 1
    sp = torch.cat([a, a + (b - a) / 3, a + 2 * (b - a) / 3, b, torch.full((S_, 2), 0.04, dtype=D64)], -1)
    x = (a[:, None] + torch.rand(S_, 12, 1, generator=g, dtype=D64) * (b - a)[:, None] +
         0.05 * torch.randn(S_, 12, 3, generator=g, dtype=D64)).reshape(-1, 3)
    n = x.shape[0]
    r = torch.full((n,), 0.05, dtype=D64)
    vd = torch.tensor([[0.0, 0.0, 1.0]], dtype=D64).expand(n, 3).contiguous()
    cp = torch.rand(S_, 3, generator=g, dtype=D64)
    sid = S.shape_sdf_id('cubic_bezier')
    gs_up = torch.randn(n, S_, generator=g, dtype=D64)
    zero_a = torch.zeros(n, S_, dtype=D64)
    zero_c = torch.zeros(n, S_, 3, dtype=D64)
    alpha = S.stroke_forward(x, r, vd, sp, cp, sid, 0, 1.0)['alpha']
    _, _, rx = S.stroke_backward_reference(zero_a, zero_c, gs_up, x, r, vd, alpha, sp, cp, sid, 0, 1.0)
    t = S.cubic_bezier_t([x[:, None, 0], x[:, None, 1], x[:, None, 2]], sp[None])
    Pt = a[None] + t[..., None] * (b - a)[None]
    v = x[:, None] - Pt
    expect = (gs_up[..., None] ** 2 * 2 * v).sum(1)
    torch.testing.assert_close(rx, expect, atol=1e-10, rtol=1e-10)
    _, _, ex = S.stroke_backward_exact(zero_a, zero_c, gs_up, x, r, vd, sp, cp, sid, 0, 1.0)
    true = (gs_up[..., None] * v / v.norm(dim=-1, keepdim=True)).sum(1)
    torch.testing.assert_close(ex, true, atol=1e-9, rtol=1e-9)

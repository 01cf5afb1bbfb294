"""Oracle restatement of the reference stroke kernels (test infrastructure only).

Follows, in order of the reference's own files:

* type registry ``source/strokelib/strokes.py:12-106`` (``_base_sdf_id``,
  ``_sdf_dict``, ``_color_dict``, ``_make_sdf_id``);
* rotation / inverse transform ``source/strokelib/src/strokes_sdf.h:25-222``;
* the twelve base SDFs ``strokes_sdf.h:242-1456`` (forward values);
* the hand-written base-SDF adjoints ``strokes_sdf.h`` ``grad_sdf`` members
  (restated in :func:`grad_sdf_reference`, including their known inexact terms);
* colour fields ``strokes_color.h:34-171`` and SH basis ``sh.h:8-30``;
* per-(point, stroke) forward ``strokes_forward.cu:35-109``;
* per-(point, stroke) adjoint ``strokes_backward.cu:38-157``;
* texture coordinates ``BaseSDF<...>::texcoord`` (``strokes_sdf.h:258-526``).

Two adjoints are provided:

``exact``
    torch.autograd of the forward restatement (run it in float64 for the
    oracle).  Non-differentiable points follow the conventions of the forward
    code (``sign(0) = +1``; ``sqrt`` of an exact zero has zero gradient).
``reference``
    a restatement of the reference's hand-written CUDA adjoint, including the
    terms in which it differs from the true derivative (cube outside,
    capped torus, cubic Bezier / Catmull-Rom squared-distance gradients,
    gradient_rgb position dependence, inv_scale_radius radius dependence).

Pinned by ``tests/test_oracle_strokes.py``: closed forms (sphere/cube/octahedron
values at axis points, capsule on its axis), brute-force Bezier distances by
dense parameter sampling, rotation orthogonality and angle conventions,
finite differences of the forward in float64 for every exact adjoint, and
hand-derived derivatives for every reference-adjoint term that is exact.
``tests/test_oracle_reference_surrogates.py`` rebuilds every reference-mode
adjoint, the inexact terms included, as torch.autograd of a surrogate forward
(the real values, the reference's local derivatives), so the hand chaining
here is checked by an independent route; ``tests/test_oracle_texcoord.py``
pins the texture coordinates by closed-form points, invariances and a sort of
the scale axes.
"""
import math

import torch

BASE_SDF_ID = {
    'unit_sphere': 0,
    'unit_cube': 1,
    'unit_round_cube': 2,
    'unit_capped_torus': 3,
    'unit_capsule': 4,
    'unit_line': 5,
    'unit_triprism': 6,
    'unit_octahedron': 7,
    'unit_tetrahedron': 8,
    'quadratic_bezier': 9,
    'cubic_bezier': 10,
    'catmull_rom': 11,
}
BASE_NAMES = {v: k for k, v in BASE_SDF_ID.items()}

# number of base (pre-transform) parameters per base SDF (strokes_sdf.h param reads)
NUM_BASE_PARAMS = [0, 0, 1, 2, 1, 2, 1, 0, 0, 11, 14, 14]

# shape name -> (base, T, R, SS, MS), strokes.py:40-82
SHAPE_TYPES = {
    'sphere': ('unit_sphere', True, False, True, False),
    'ellipsoid': ('unit_sphere', True, True, False, True),
    'aacube': ('unit_cube', True, False, True, False),
    'cube': ('unit_cube', True, True, True, False),
    'aabb': ('unit_cube', True, False, False, True),
    'obb': ('unit_cube', True, True, False, True),
    'roundcube': ('unit_round_cube', True, True, True, False),
    'roundbox': ('unit_round_cube', True, True, False, True),
    'cappedtorus': ('unit_capped_torus', True, True, True, False),
    'capsule': ('unit_capsule', True, True, True, False),
    'scapsule': ('unit_capsule', True, True, False, True),
    'line': ('unit_line', True, True, True, False),
    'triprism': ('unit_triprism', True, True, True, False),
    'octahedron': ('unit_octahedron', True, True, True, False),
    'tetrahedron': ('unit_tetrahedron', True, True, True, False),
    'quadratic_bezier': ('quadratic_bezier', False, False, False, False),
    'cubic_bezier': ('cubic_bezier', False, False, False, False),
    'catmull_rom': ('catmull_rom', False, False, False, False),
}

COLOR_TYPES = {'constant_rgb': 0, 'gradient_rgb': 1, 'constant_sh2': 2, 'constant_sh3': 3}
NUM_COLOR_PARAMS = [3, 12, 12, 27]


def make_sdf_id(base_name, enable_translation, enable_rotation, enable_singlescale, enable_multiscale):
    """strokes.py:95-106: ``base << 4 | T | R<<1 | SS<<2 | MS<<3``."""
    sdf_id = BASE_SDF_ID[base_name] << 4
    if enable_translation:
        sdf_id |= 1
    if enable_rotation:
        sdf_id |= 2
    if enable_singlescale:
        sdf_id |= 4
    if enable_multiscale:
        sdf_id |= 8
    return sdf_id


def decode_sdf_id(sdf_id):
    return (sdf_id >> 4, bool(sdf_id & 1), bool(sdf_id & 2), bool(sdf_id & 4), bool(sdf_id & 8))


def shape_sdf_id(shape_type):
    base, t, r, ss, ms = SHAPE_TYPES[shape_type]
    return make_sdf_id(base, t, r, ss, ms)


def num_shape_params(sdf_id):
    base, t, r, ss, ms = decode_sdf_id(sdf_id)
    return NUM_BASE_PARAMS[base] + 3 * t + 3 * r + (1 if ss else (3 if ms else 0))


# ----------------------------------------------------------------------------
# small vector helpers on [..., 3] tensors (component lists keep the reference's
# left-to-right operation order, which matters for the float32 emulation)
# ----------------------------------------------------------------------------

def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _v(t):
    return [t[..., 0], t[..., 1], t[..., 2]]


def _sign(v):
    """Reference convention ``v < 0 ? -1 : 1`` (sign of zero is +1)."""
    return torch.where(v < 0, -torch.ones_like(v), torch.ones_like(v))


def _safe_sqrt(v):
    """sqrt with zero gradient at exactly zero (reference ``length`` has no epsilon)."""
    pos = v > 0
    return torch.where(pos, torch.sqrt(torch.where(pos, v, torch.ones_like(v))), torch.zeros_like(v))


def _clamp(v, lo, hi):
    """helper_math.h:1166-1169 ``fmaxf(a, fminf(f, b))``."""
    return torch.clamp(v, lo, hi)


def _lerp(a, b, t):
    """helper_math.h:1144-1147 ``a + t * (b - a)``."""
    return a + t * (b - a)


def _sigmoid(v):
    """common.h:41-45 ``1 / (1 + exp(-v))``.

    The value is that formula; its autograd derivative is sigma (1 - sigma)
    (torch.sigmoid's), which stays finite where exp(-v) overflows and the
    formula's own derivative would be inf / inf.
    """
    s = torch.sigmoid(v)
    return s + (1.0 / (1.0 + torch.exp(-v)) - s).detach()


# ----------------------------------------------------------------------------
# rotation (strokes_sdf.h:25-56)
# ----------------------------------------------------------------------------

def rotation_entries(angle):
    """Entries r00..r22 of R = Rz(z) Ry(y) Rx(x), strokes_sdf.h:28-39."""
    sx, sy, sz = torch.sin(angle[..., 0]), torch.sin(angle[..., 1]), torch.sin(angle[..., 2])
    cx, cy, cz = torch.cos(angle[..., 0]), torch.cos(angle[..., 1]), torch.cos(angle[..., 2])
    r00 = cy * cz
    r01 = sx * sy * cz - cx * sz
    r02 = sx * sz + cx * sy * cz
    r10 = cy * sz
    r11 = cx * cz + sx * sy * sz
    r12 = cx * sy * sz - sx * cz
    r20 = -sy
    r21 = sx * cy
    r22 = cx * cy
    return [[r00, r01, r02], [r10, r11, r12], [r20, r21, r22]]


def rotation_matrix(angle):
    r = rotation_entries(angle)
    return torch.stack([torch.stack(row, -1) for row in r], -2)


def rotate_inverse(p, r):
    """rotate_point<true>: Rᵀ p (strokes_sdf.h:43-46)."""
    return [r[0][0] * p[0] + r[1][0] * p[1] + r[2][0] * p[2],
            r[0][1] * p[0] + r[1][1] * p[1] + r[2][1] * p[2],
            r[0][2] * p[0] + r[1][2] * p[1] + r[2][2] * p[2]]


def split_shape_params(params, sdf_id):
    """Split [..., P] shape params into (base, scale, angle, translation).

    The transform is read from the end backwards (strokes_sdf.h:164-191):
    translation = last 3, euler angles = previous 3, scale = previous 1 or 3.
    """
    base, t, r, ss, ms = decode_sdf_id(sdf_id)
    end = params.shape[-1]
    trans = angle = scale = None
    if t:
        trans = params[..., end - 3:end]
        end -= 3
    if r:
        angle = params[..., end - 3:end]
        end -= 3
    if ss:
        scale = params[..., end - 1:end]
        end -= 1
    elif ms:
        scale = params[..., end - 3:end]
        end -= 3
    return params[..., :end], scale, angle, trans


def inverse_transform(x, params, sdf_id):
    """strokes_sdf.h:164-191: p = Rᵀ (x - t) / s, returned as a component list."""
    _, scale, angle, trans = split_shape_params(params, sdf_id)
    p = _v(x)
    if trans is not None:
        p = [p[i] - trans[..., i] for i in range(3)]
    if angle is not None:
        p = rotate_inverse(p, rotation_entries(angle))
    if scale is not None:
        if scale.shape[-1] == 1:
            p = [p[i] / scale[..., 0] for i in range(3)]
        else:
            p = [p[i] / scale[..., i] for i in range(3)]
    return p


def inverse_transform_direction(d, params, sdf_id):
    """strokes_sdf.h:193-210 (rotation, then multiscale divide + normalize)."""
    _, scale, angle, _ = split_shape_params(params, sdf_id)
    _, _, _, _, ms = decode_sdf_id(sdf_id)
    d = _v(d) if torch.is_tensor(d) else d
    if angle is not None:
        d = rotate_inverse(d, rotation_entries(angle))
    if ms:
        d = [d[i] / scale[..., i] for i in range(3)]
        inv = torch.rsqrt(_dot(d, d))
        d = [d[i] * inv for i in range(3)]
    return d


# ----------------------------------------------------------------------------
# base SDF values (strokes_sdf.h)
# ----------------------------------------------------------------------------

def sdf_sphere(p, prm):
    """strokes_sdf.h:245-249."""
    return torch.sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + 1e-8) - 1.0


def sdf_cube(p, prm):
    """strokes_sdf.h:273-277."""
    q = [torch.abs(p[i]) - 1.0 for i in range(3)]
    qp = [torch.clamp_min(q[i], 0.0) for i in range(3)]
    return _safe_sqrt(_dot(qp, qp)) + torch.clamp_max(torch.maximum(q[0], torch.maximum(q[1], q[2])), 0.0)


def sdf_round_cube(p, prm):
    """strokes_sdf.h:316-325."""
    d = [torch.abs(p[i]) - 1.0 for i in range(3)]
    dp = [torch.clamp_min(d[i], 0.0) for i in range(3)]
    norm = torch.sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2] + 1e-8)
    dmin = torch.clamp_max(torch.maximum(d[0], torch.maximum(d[1], d[2])), 0.0)
    return norm + dmin - prm[..., 0]


def _torus_k(p, prm, eps_in_length=True):
    ax = torch.abs(p[0])
    s, c = torch.sin(prm[..., 0]), torch.cos(prm[..., 0])
    cross = ax * c - p[1] * s
    if eps_in_length:
        ln = torch.sqrt(ax * ax + p[1] * p[1] + 1e-8)
    else:
        ln = _safe_sqrt(ax * ax + p[1] * p[1])
    return torch.where(cross > 0, ax * s + p[1] * c, ln), cross, ax, s, c


def sdf_capped_torus(p, prm):
    """strokes_sdf.h:373-380."""
    k, _, ax, _, _ = _torus_k(p, prm)
    return torch.sqrt(ax * ax + p[1] * p[1] + p[2] * p[2] + 1.0 - 2.0 * k + 1e-8) - prm[..., 1]


def sdf_capsule(p, prm):
    """strokes_sdf.h:434-440 (the reference evaluates the sqrt in double)."""
    h = prm[..., 0]
    y = p[1] - torch.maximum(-h, torch.minimum(p[1], h))
    return torch.sqrt(p[0] * p[0] + y * y + p[2] * p[2] + 1e-8) - 1.0


def sdf_line(p, prm):
    """strokes_sdf.h:474-483."""
    h, rd = prm[..., 0], prm[..., 1]
    t = _clamp((p[1] + h) / (2.0 * h), 0.0, 1.0)
    r = _lerp(1.0 - rd, 1.0 + rd, t)
    y = p[1] - torch.maximum(-h, torch.minimum(p[1], h))
    return torch.sqrt(p[0] * p[0] + y * y + p[2] * p[2] + 1e-8) - r


TRI_C = 0.866025


def sdf_triprism(p, prm):
    """strokes_sdf.h:531-535."""
    qx, qy = torch.abs(p[0]), torch.abs(p[1])
    return torch.maximum(qy - prm[..., 0], torch.maximum(qx * TRI_C + p[2] * 0.5, -p[2]) - 0.5)


OCTA_C = 0.57735027


def sdf_octahedron(p, prm):
    """strokes_sdf.h:584-588."""
    return (torch.abs(p[0]) + torch.abs(p[1]) + torch.abs(p[2]) - 1.0) * OCTA_C


def sdf_tetrahedron(p, prm):
    """strokes_sdf.h:610-613."""
    return (torch.maximum(torch.abs(p[0] + p[1]) - p[2], torch.abs(p[0] - p[1]) + p[2]) - 1.0) * OCTA_C


def _quad_setup(p, prm):
    A = [prm[..., 0], prm[..., 1], prm[..., 2]]
    B = [prm[..., 3], prm[..., 4], prm[..., 5]]
    C = [prm[..., 6], prm[..., 7], prm[..., 8]]
    a = [B[i] - A[i] for i in range(3)]
    b = [A[i] - 2.0 * B[i] + C[i] for i in range(3)]
    c = [a[i] * 2.0 for i in range(3)]
    d = [A[i] - p[i] for i in range(3)]
    kk = 1.0 / _dot(b, b)
    kx = kk * _dot(a, b)
    ky = kk * (2.0 * _dot(a, a) + _dot(d, b)) / 3.0
    kz = kk * _dot(d, a)
    pp = ky - kx * kx
    p3 = pp * pp * pp
    q = kx * (2.0 * kx * kx - 3.0 * ky) + kz
    h = q * q + 4.0 * p3
    return a, b, c, d, kx, ky, kz, pp, q, h


def _quad_dist2(d, c, b, t):
    m = [d[i] + (c[i] + b[i] * t) * t for i in range(3)]
    return _dot(m, m)


def _cbrt_signed(v):
    """copysignf(powf(fabs(x), 1/3), x), strokes_sdf.h:671-673."""
    av = torch.abs(v)
    pos = av > 0
    r = torch.where(pos, torch.pow(torch.where(pos, av, torch.ones_like(av)), 1.0 / 3.0), torch.zeros_like(av))
    return torch.where(v < 0, -r, r)


def quad_bezier_dist2_t(p, prm):
    """Closed-form closest point, strokes_sdf.h:649-695.  Returns (dist², t)."""
    a, b, c, d, kx, ky, kz, pp, q, h = _quad_setup(p, prm)
    one = h >= 0.0
    # branch 1 (one real root); inputs sanitised so the unused branch stays finite
    h1 = torch.sqrt(torch.where(one, h, torch.zeros_like(h)) + 1e-8)
    x0 = (h1 - q) * 0.5
    x1 = -(h1 + q) * 0.5
    t1 = _clamp((_cbrt_signed(x0) + _cbrt_signed(x1)) - kx, 0.0, 1.0)
    d1 = _quad_dist2(d, c, b, t1)
    # branch 2 (three real roots, two candidates evaluated)
    pn = torch.where(one, -torch.ones_like(pp), pp)
    qn = torch.where(one, torch.zeros_like(q), q)
    z = torch.sqrt(-pn + 1e-8)
    arg = qn / (pn * z * 2.0 + 1e-8)
    # clamped: the regularisers can push |arg| past 1 when h ~ 0, where the
    # reference's acosf returns NaN (the only deviation from strokes_sdf.h:680)
    arg = torch.clamp(torch.where(one, torch.zeros_like(arg), arg), -1.0, 1.0)
    v = torch.acos(arg) / 3.0
    m = torch.cos(v)
    n = torch.sin(v) * 1.732050808
    ta = _clamp((m + n) * z - kx, 0.0, 1.0)
    tb = _clamp((-n - m) * z - kx, 0.0, 1.0)
    da = _quad_dist2(d, c, b, ta)
    db = _quad_dist2(d, c, b, tb)
    use_b = db < da
    d2 = torch.where(use_b, db, da)
    t2 = torch.where(use_b, tb, ta)
    return torch.where(one, d1, d2), torch.where(one, t1, t2)


def sdf_quadratic_bezier(p, prm):
    """strokes_sdf.h:697-705."""
    d2, t = quad_bezier_dist2_t(p, prm)
    dist = torch.sqrt(d2 + 1e-8)
    return dist - _lerp(prm[..., 9], prm[..., 10], t)


CURVE_SEGMENTS = 10


def _cubic_point(t, A, B, C, D):
    """strokes_sdf.h:1243-1251."""
    u = 1.0 - t
    w0 = u * u * u
    w1 = 3.0 * u * u * t
    w2 = 3.0 * u * t * t
    w3 = t * t * t
    return [w0 * A[i] + w1 * B[i] + w2 * C[i] + w3 * D[i] for i in range(3)]


def _catmull_point(t, A, B, C, D):
    """strokes_sdf.h:1326-1331."""
    t2 = t * t
    t3 = t2 * t
    return [t3 * A[i] + t2 * B[i] + t * C[i] + D[i] for i in range(3)]


def _polyline_t(p, point_fn, A, B, C, D, like):
    """Closest polyline parameter, strokes_sdf.h:1228-1269 / 1349-1364."""
    step = 1.0 / CURVE_SEGMENTS
    dist_min = torch.full_like(like, 1e9)
    t_min = torch.zeros_like(like)
    t_start = 0.0
    s0 = point_fn(torch.zeros_like(like) + t_start, A, B, C, D)
    for i in range(1, CURVE_SEGMENTS + 1):
        t_end = i * step
        s1 = point_fn(torch.zeros_like(like) + t_end, A, B, C, D)
        ab = [s1[k] - s0[k] for k in range(3)]
        pa = [p[k] - s0[k] for k in range(3)]
        ts = _clamp(_dot(pa, ab) / _dot(ab, ab), 0.0, 1.0)
        proj = [s0[k] + ts * ab[k] for k in range(3)]
        diff = [p[k] - proj[k] for k in range(3)]
        dist = _safe_sqrt(_dot(diff, diff))
        better = dist < dist_min
        dist_min = torch.where(better, dist, dist_min)
        t_min = torch.where(better, t_start * (1.0 - ts) + t_end * ts, t_min)
        t_start = t_end
        s0 = s1
    return t_min


def curve_selection_margin(base, p, prm):
    """Relative gap between the closest and the runner-up curve-parameter candidate.

    Cubic / Catmull-Rom: distances to the ten polyline segments (candidates with
    a different parameter t); quadratic: the two real roots of the three-root
    branch (inf in the one-root branch).  A gap below float32 resolution means
    any float32 evaluation may select either candidate.
    """
    if base == 9:
        a, b, c, d, kx, ky, kz, pp, q, h = _quad_setup(p, prm)
        one = h >= 0.0
        pn = torch.where(one, -torch.ones_like(pp), pp)
        z = torch.sqrt(-pn + 1e-8)
        arg = torch.clamp(torch.where(one, torch.zeros_like(q), q) / (pn * z * 2.0 + 1e-8), -1.0, 1.0)
        v = torch.acos(arg) / 3.0
        m, n = torch.cos(v), torch.sin(v) * 1.732050808
        ta = _clamp((m + n) * z - kx, 0.0, 1.0)
        tb = _clamp((-n - m) * z - kx, 0.0, 1.0)
        da, db = _quad_dist2(d, c, b, ta), _quad_dist2(d, c, b, tb)
        gap = (da - db).abs() / torch.maximum(da, db).clamp_min(1e-30)
        gap = torch.where((ta - tb).abs() > 1e-9, gap, torch.full_like(gap, float('inf')))
        return torch.where(one, torch.full_like(gap, float('inf')), gap)
    if base == 10:
        A, B, C, D = _cubic_ctrl(prm)
        fn = _cubic_point
    else:
        A, B, C, D = catmull_coeffs(prm)
        fn = _catmull_point
    like = p[0] + prm[..., 0] * 0
    step = 1.0 / CURVE_SEGMENTS
    dists, ts = [], []
    s0 = fn(torch.zeros_like(like), A, B, C, D)
    for i in range(1, CURVE_SEGMENTS + 1):
        s1 = fn(torch.zeros_like(like) + i * step, A, B, C, D)
        ab = [s1[k] - s0[k] for k in range(3)]
        pa = [p[k] - s0[k] for k in range(3)]
        ts_ = _clamp(_dot(pa, ab) / _dot(ab, ab), 0.0, 1.0)
        diff = [p[k] - (s0[k] + ts_ * ab[k]) for k in range(3)]
        dists.append(_safe_sqrt(_dot(diff, diff)))
        ts.append((i - 1) * step * (1.0 - ts_) + i * step * ts_)
        s0 = s1
    dd = torch.stack(dists, -1)
    tt = torch.stack(ts, -1)
    best = dd.argmin(-1, keepdim=True)
    d1 = torch.gather(dd, -1, best)
    t1 = torch.gather(tt, -1, best)
    other = torch.where((tt - t1).abs() > 1e-9, dd, torch.full_like(dd, float('inf')))
    d2 = other.min(-1, keepdim=True).values
    return ((d2 - d1) / d1.clamp_min(1e-30)).squeeze(-1)


def _cubic_ctrl(prm):
    return ([prm[..., 0], prm[..., 1], prm[..., 2]], [prm[..., 3], prm[..., 4], prm[..., 5]],
            [prm[..., 6], prm[..., 7], prm[..., 8]], [prm[..., 9], prm[..., 10], prm[..., 11]])


def cubic_bezier_t(p, prm):
    A, B, C, D = _cubic_ctrl(prm)
    return _polyline_t(p, _cubic_point, A, B, C, D, p[0] + prm[..., 0] * 0)


def sdf_cubic_bezier(p, prm):
    """strokes_sdf.h:1253-1283."""
    A, B, C, D = _cubic_ctrl(prm)
    t = cubic_bezier_t(p, prm)
    P = _cubic_point(t, A, B, C, D)
    diff = [P[k] - p[k] for k in range(3)]
    return _safe_sqrt(_dot(diff, diff)) - _lerp(prm[..., 12], prm[..., 13], t)


def catmull_coeffs(prm):
    """Centripetal Catmull-Rom segment coefficients, strokes_sdf.h:1335-1345."""
    p0, p1, p2, p3 = _cubic_ctrl(prm)

    def _len(u, v):
        w = [u[k] - v[k] for k in range(3)]
        return _safe_sqrt(_dot(w, w))

    t01 = torch.sqrt(_len(p0, p1) + 1e-8)
    t12 = torch.sqrt(_len(p1, p2) + 1e-8)
    t23 = torch.sqrt(_len(p2, p3) + 1e-8)
    m1 = [p2[k] - p1[k] + t12 * ((p1[k] - p0[k]) / t01 - (p2[k] - p0[k]) / (t01 + t12)) for k in range(3)]
    m2 = [p2[k] - p1[k] + t12 * ((p3[k] - p2[k]) / t23 - (p3[k] - p1[k]) / (t12 + t23)) for k in range(3)]
    A = [2.0 * (p1[k] - p2[k]) + m1[k] + m2[k] for k in range(3)]
    B = [3.0 * (p2[k] - p1[k]) - 2.0 * m1[k] - m2[k] for k in range(3)]
    return A, B, m1, p1


def catmull_rom_t(p, prm):
    A, B, C, D = catmull_coeffs(prm)
    return _polyline_t(p, _catmull_point, A, B, C, D, p[0] + prm[..., 0] * 0)


def sdf_catmull_rom(p, prm):
    """strokes_sdf.h:1333-1383."""
    A, B, C, D = catmull_coeffs(prm)
    t = _polyline_t(p, _catmull_point, A, B, C, D, p[0] + prm[..., 0] * 0)
    P = _catmull_point(t, A, B, C, D)
    diff = [P[k] - p[k] for k in range(3)]
    return _safe_sqrt(_dot(diff, diff)) - _lerp(prm[..., 12], prm[..., 13], t)


BASE_SDF_FNS = [sdf_sphere, sdf_cube, sdf_round_cube, sdf_capped_torus, sdf_capsule, sdf_line,
                sdf_triprism, sdf_octahedron, sdf_tetrahedron, sdf_quadratic_bezier,
                sdf_cubic_bezier, sdf_catmull_rom]


def base_sdf(base, p, prm):
    return BASE_SDF_FNS[base](p, prm)


# ----------------------------------------------------------------------------
# texture coordinates (BaseSDF<...>::texcoord, written by strokes_forward.cu:105-107)
# ----------------------------------------------------------------------------

def permute_multiscale_axis(d, scale):
    """strokes_sdf.h:211-222: the components reordered by decreasing scale (component lists)."""
    sx, sy, sz = scale
    xy, yz, xz = sx > sy, sy > sz, sx > sz
    zy = sz > sy
    first = [d[0], d[1], d[2]], [d[0], d[2], d[1]], [d[2], d[0], d[1]]
    second = [d[2], d[1], d[0]], [d[1], d[0], d[2]], [d[1], d[2], d[0]]
    out = []
    for i in range(3):
        when_xy = torch.where(yz, first[0][i], torch.where(xz, first[1][i], first[2][i]))
        when_not = torch.where(zy, second[0][i], torch.where(xz, second[1][i], second[2][i]))
        out.append(torch.where(xy, when_xy, when_not))
    return out


def texcoord(base, p, row, multiscale):
    """(u, v) of a unit-space point p (component list) for a stroke whose parameter row is ``row``
    ([..., P], the whole row as the reference's ``params`` pointer); (0, 0) for the bases without
    texture coordinates.  Sphere strokes_sdf.h:258-268, box / round box :301-311 / :357-367
    (multiscale permutation from params + 0), capsule :456-469 and line :511-526 (swap of x and
    z when scale_z > scale_x, scale read from params + 1)."""
    zero = torch.zeros_like(p[0])
    if base == 0:
        n = torch.rsqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2])
        d = [p[i] * n for i in range(3)]
        if multiscale:
            d = permute_multiscale_axis(d, [row[..., i] for i in range(3)])
        u = torch.asin(torch.clamp(d[0], -1.0, 1.0)) / math.pi + 0.5
        v = torch.atan2(d[2], d[1]) / (2.0 * math.pi) + 0.5
        return u, v
    if base in (1, 2):
        uvt = [p[i] * 0.5 + 0.5 for i in range(3)]
        if multiscale:
            uvt = permute_multiscale_axis(uvt, [row[..., i] for i in range(3)])
        return torch.clamp(uvt[0], 0.0, 1.0), torch.clamp(uvt[1], 0.0, 1.0)
    if base in (4, 5):
        h = row[..., 0]
        if base == 4:
            u = torch.clamp((p[1] / (h + 1.0)) * 0.5 + 0.5, 0.0, 1.0)
        else:
            u = torch.clamp((p[1] + h + 1.0 - row[..., 1]) / (2.0 * h + 2.0), 0.0, 1.0)
        px, pz = p[0], p[2]
        if multiscale:
            swap = row[..., 3] > row[..., 1]
            px, pz = torch.where(swap, p[2], p[0]), torch.where(swap, p[0], p[2])
        v = torch.atan2(pz, px) / (2.0 * math.pi) + 0.5
        return u, v
    return zero, zero


def stroke_texcoords(x, shape_params, sdf_id):
    """texcoord [N, S, 2] of points x [N, 3] for strokes [S, P] of one type."""
    base, _, _, _, ms = decode_sdf_id(sdf_id)
    p = inverse_transform(x[:, None, :], shape_params[None], sdf_id)
    p = [pi.expand(x.shape[0], shape_params.shape[0]) for pi in p]
    u, v = texcoord(base, p, shape_params[None].expand(x.shape[0], -1, -1), ms)
    return torch.stack([u, v], -1)


# ----------------------------------------------------------------------------
# reference hand-written base adjoints (strokes_sdf.h grad_sdf members)
# returns (dSDF/dpos as component list, dSDF/dbase_params [..., nb] or None)
# where the param part is what the reference multiplies by grad_SDF.
# ----------------------------------------------------------------------------

def grad_sdf_reference(base, p, prm):
    z = torch.zeros_like(p[0])
    one = torch.ones_like(p[0])
    if base == 0:  # sphere, strokes_sdf.h:251-256
        inv = torch.rsqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + 1e-8)
        return [p[i] * inv for i in range(3)], None
    if base == 1:  # cube, strokes_sdf.h:279-299 (outside: gradient of squared length)
        q = [torch.abs(p[i]) - 1.0 for i in range(3)]
        inside = torch.maximum(q[0], torch.maximum(q[1], q[2])) <= 0.0
        gx_in = torch.where((q[0] >= q[1]) & (q[0] >= q[2]), _sign(p[0]), z)
        gy_in = torch.where((q[0] < q[1]) & (q[1] >= q[2]), _sign(p[1]), z)
        gz_in = torch.where((q[0] < q[2]) & (q[1] < q[2]), _sign(p[2]), z)
        go = [torch.where(q[i] > 0, 2.0 * torch.clamp_min(q[i], 0.0), z) * _sign(p[i]) for i in range(3)]
        return [torch.where(inside, gx_in, go[0]), torch.where(inside, gy_in, go[1]),
                torch.where(inside, gz_in, go[2])], None
    if base == 2:  # round cube, strokes_sdf.h:327-355
        d = [torch.abs(p[i]) - 1.0 for i in range(3)]
        dp = [torch.clamp_min(d[i], 0.0) for i in range(3)]
        sq = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]
        inv = torch.rsqrt(sq + 1e-8)
        sg = [_sign(p[i]) for i in range(3)]
        gdp = [dp[i] * inv * sg[i] for i in range(3)]
        dmin = torch.clamp_max(torch.maximum(d[0], torch.maximum(d[1], d[2])), 0.0)
        gm = torch.where(dmin < 0.0, one, z)
        gmx = gm * torch.where((d[0] >= d[1]) & (d[0] >= d[2]), one, z) * sg[0]
        gmy = gm * torch.where((d[0] < d[1]) & (d[1] >= d[2]), one, z) * sg[1]
        gmz = gm * torch.where((d[0] < d[2]) & (d[1] < d[2]), one, z) * sg[2]
        return [gdp[0] + gmx, gdp[1] + gmy, gdp[2] + gmz], (-one)[..., None]
    if base == 3:  # capped torus, strokes_sdf.h:382-420 (inexact)
        ax = torch.abs(p[0])
        s, c = torch.sin(prm[..., 0]), torch.cos(prm[..., 0])
        cross = ax * c - p[1] * s
        k = torch.where(cross > 0, ax * s + p[1] * c, _safe_sqrt(ax * ax + p[1] * p[1]))
        inv = torch.rsqrt(ax * ax + p[1] * p[1] + p[2] * p[2] + 1.0 - 2.0 * k + 1e-8)
        kpos = k > 0.0
        gx = torch.where(kpos, p[0] + (-_sign(p[0]) * s), p[0] + p[0]) * inv
        gy = torch.where(kpos, p[1] + (-c), p[1] + p[1]) * inv
        gz = p[2] * inv
        g0 = torch.where(kpos, -(ax * c + p[1] * s), z)
        return [gx, gy, gz], torch.stack([g0, -one], -1)
    if base == 4:  # capsule, strokes_sdf.h:442-454
        h = prm[..., 0]
        y = p[1] - torch.maximum(-h, torch.minimum(p[1], h))
        inv = torch.rsqrt(p[0] * p[0] + y * y + p[2] * p[2] + 1e-8)
        g = [p[0] * inv, y * inv, p[2] * inv]
        inside = (-h <= p[1]) & (p[1] <= h)
        gy = torch.where(inside, z, g[1])
        g0 = torch.where(inside, z, g[1] * torch.where(p[1] > 0, -one, one))
        return [g[0], gy, g[2]], g0[..., None]
    if base == 5:  # line, strokes_sdf.h:486-509
        h, rd = prm[..., 0], prm[..., 1]
        t = _clamp((p[1] + h) / (2.0 * h), 0.0, 1.0)
        y = p[1] - torch.maximum(-h, torch.minimum(p[1], h))
        inv = torch.rsqrt(p[0] * p[0] + y * y + p[2] * p[2] + 1e-8)
        g = [p[0] * inv, y * inv, p[2] * inv]
        inside = (-h <= p[1]) & (p[1] <= h)
        gy = torch.where(inside, -rd / h, g[1])
        g0 = torch.where(inside, p[1] * rd / (h * h), g[1] * torch.where(p[1] > 0, -one, one))
        return [g[0], gy, g[2]], torch.stack([g0, 1.0 - 2.0 * t], -1)
    if base == 6:  # triprism, strokes_sdf.h:537-572
        qx, qy = torch.abs(p[0]), torch.abs(p[1])
        e1 = qx * TRI_C + p[2] * 0.5
        c1 = e1 > -p[2]
        top1 = qy - prm[..., 0] > e1 - 0.5
        top2 = qy - prm[..., 0] > -p[2] - 0.5
        top = torch.where(c1, top1, top2)
        gx = torch.where(c1 & ~top1, TRI_C * _sign(p[0]), z)
        gy = torch.where(top, _sign(p[1]), z)
        gz = torch.where(c1, torch.where(top1, z, 0.5 * one), torch.where(top2, z, -one))
        return [gx, gy, gz], torch.where(top, -one, z)[..., None]
    if base == 7:  # octahedron, strokes_sdf.h:590-597
        return [OCTA_C * _sign(p[i]) for i in range(3)], None
    if base == 8:  # tetrahedron, strokes_sdf.h:615-633
        a = p[0] + p[1]
        cc = p[0] - p[1]
        first = torch.abs(a) - p[2] > torch.abs(cc) + p[2]
        gx = torch.where(first, OCTA_C * _sign(a), OCTA_C * _sign(cc))
        gy = torch.where(first, OCTA_C * _sign(a), -OCTA_C * _sign(cc))
        gz = torch.where(first, -OCTA_C * one, OCTA_C * one)
        return [gx, gy, gz], None
    if base == 9:
        return grad_quad_bezier_reference(p, prm)
    if base in (10, 11):  # strokes_sdf.h:1285-1318 / 1385-1445 (squared-distance gradient)
        if base == 10:
            A, B, C, D = _cubic_ctrl(prm)
            t = _polyline_t(p, _cubic_point, A, B, C, D, p[0] + prm[..., 0] * 0)
            P = _cubic_point(t, A, B, C, D)
        else:
            A, B, C, D = catmull_coeffs(prm)
            t = _polyline_t(p, _catmull_point, A, B, C, D, p[0] + prm[..., 0] * 0)
            P = _catmull_point(t, A, B, C, D)
        v = [P[k] - p[k] for k in range(3)]
        u = 1.0 - t
        # the reference returns -grad_v with grad_v = grad_SDF * 2 v; the caller
        # multiplies by grad_SDF once more.  Return 2v so the caller can apply both.
        if base == 10:
            w = [u * u * u, 3.0 * u * u * t, 3.0 * u * t * t, t * t * t]
            gp = []
            for wi in w:
                gp += [2.0 * v[k] * wi for k in range(3)]
        else:
            gp = _catmull_param_grad_reference(prm, t, v)
        gp += [-u, -t]
        return ('squared', v), torch.stack(gp, -1)
    raise ValueError(base)


def _catmull_param_grad_reference(prm, t, v):
    """strokes_sdf.h:1408-1445 with grad_P = 2 v (the caller scales by grad_SDF)."""
    p0, p1, p2, p3 = _cubic_ctrl(prm)
    sub = lambda u, w: [u[k] - w[k] for k in range(3)]
    p0p1, p1p2, p2p3 = sub(p0, p1), sub(p1, p2), sub(p2, p3)
    nrm = lambda u: _safe_sqrt(_dot(u, u))
    t01 = torch.sqrt(nrm(p0p1) + 1e-8)
    t12 = torch.sqrt(nrm(p1p2) + 1e-8)
    t23 = torch.sqrt(nrm(p2p3) + 1e-8)
    gP = [2.0 * v[k] for k in range(3)]
    gA = [gP[k] * (t * t * t) for k in range(3)]
    gB = [gP[k] * (t * t) for k in range(3)]
    gC = [gP[k] * t for k in range(3)]
    gD = gP
    gm1 = [gA[k] - gB[k] * 2.0 + gC[k] for k in range(3)]
    gm2 = [gA[k] - gB[k] for k in range(3)]
    dm1_dp0 = t12 / (t01 + t12) - t12 / t01
    dm1_dp1 = -1.0 + t12 / t01
    dm1_dp2 = 1.0 - t12 / (t01 + t12)
    dm2_dp1 = -1.0 + t12 / (t12 + t23)
    dm2_dp2 = 1.0 - t12 / t23
    dm2_dp3 = t12 / t23 - t12 / (t12 + t23)
    s01 = (t01 + t12) * (t01 + t12)
    s23 = (t12 + t23) * (t12 + t23)
    dm1_dt01 = [t12 * ((p2[k] - p0[k]) / s01 - (p1[k] - p0[k]) / (t01 * t01)) for k in range(3)]
    dm1_dt12 = [(p1[k] - p0[k]) / t01 + t12 * (p2[k] - p0[k]) / s01 - (p2[k] - p0[k]) / (t01 + t12) for k in range(3)]
    dm2_dt12 = [(p2[k] - p1[k]) / t23 + t12 * (p3[k] - p1[k]) / s23 - (p3[k] - p1[k]) / (t12 + t23) for k in range(3)]
    dm2_dt23 = [t12 * ((p2[k] - p3[k]) / (t23 * t23) + (p3[k] - p1[k]) / s23) for k in range(3)]
    den = lambda u: 2.0 * torch.pow(_dot(u, u) + 1e-8, 0.75)
    d01, d12, d23 = den(p0p1), den(p1p2), den(p2p3)
    dt01 = [p0p1[k] / d01 for k in range(3)]
    dt12 = [p1p2[k] / d12 for k in range(3)]
    dt23 = [p2p3[k] / d23 for k in range(3)]
    g0 = [gm1[k] * (dm1_dp0 + dm1_dt01[k] * dt01[k]) for k in range(3)]
    g1 = [gA[k] * 2.0 + gB[k] * -3.0 + gD[k] + gm1[k] * (dm1_dp1 - dm1_dt01[k] * dt01[k] + dm1_dt12[k] * dt12[k])
          + gm2[k] * (dm2_dp1 + dm2_dt12[k] * dt12[k]) for k in range(3)]
    g2 = [gA[k] * -2.0 + gB[k] * 3.0 + gm1[k] * (dm1_dp2 - dm1_dt12[k] * dt12[k])
          + gm2[k] * (dm2_dp2 - dm2_dt12[k] * dt12[k] + dm2_dt23[k] * dt23[k]) for k in range(3)]
    g3 = [gm2[k] * (dm2_dp3 - dm2_dt23[k] * dt23[k]) for k in range(3)]
    return g0 + g1 + g2 + g3


def grad_quad_bezier_reference(p, prm):
    """Restatement of the hand-written quadratic Bezier adjoint, strokes_sdf.h:708-1172.

    Reproduces its code-evident deviations from the true derivative: the comma
    operator in ``grad_uv*_h`` (``(powf(..), x.x) / 6`` in the xyz path and
    ``/ 3`` in the params path, :1063-1064 / :1119-1120) and the
    ``(res.y + 1.0f)`` radius term of the three-root/t.x branch (:1046).
    Returns (dSDF/dpos, dSDF/dparams[..., 11]).
    """
    A = [prm[..., 0], prm[..., 1], prm[..., 2]]
    B = [prm[..., 3], prm[..., 4], prm[..., 5]]
    C = [prm[..., 6], prm[..., 7], prm[..., 8]]
    r1, r2 = prm[..., 9], prm[..., 10]
    a = [B[i] - A[i] for i in range(3)]
    b = [A[i] - 2.0 * B[i] + C[i] for i in range(3)]
    c = [a[i] * 2.0 for i in range(3)]
    d = [A[i] - p[i] for i in range(3)]
    kk = 1.0 / _dot(b, b)
    kx = kk * _dot(a, b)
    ky = kk * (2.0 * _dot(a, a) + _dot(d, b)) / 3.0
    kz = kk * _dot(d, a)
    pp = ky - kx * kx
    q = kx * (2.0 * kx * kx - 3.0 * ky) + kz
    s = q * q + 4.0 * pp * pp * pp
    b2 = _dot(b, b)
    g_kk_b2 = -1.0 / (b2 * b2)
    zero = torch.zeros_like(kx)

    # d(a,b,c,d)/d(param j) for j = 0..8: unit-vector Jacobians (strokes_sdf.h:732-777)
    def unit(j, scale):
        comp = j % 3
        return [scale if k == comp else 0.0 for k in range(3)]

    ga = [unit(j, -1.0) if j < 3 else (unit(j, 1.0) if j < 6 else [0.0] * 3) for j in range(9)]
    gb = [unit(j, 1.0) if j < 3 else (unit(j, -2.0) if j < 6 else unit(j, 1.0)) for j in range(9)]
    gc = [unit(j, -2.0) if j < 3 else (unit(j, 2.0) if j < 6 else [0.0] * 3) for j in range(9)]
    gd = [unit(j, 1.0) if j < 3 else [0.0] * 3 for j in range(9)]
    dotc = lambda u, w: u[0] * w[0] + u[1] * w[1] + u[2] * w[2]
    g_kx, g_ky, g_kz = [], [], []
    aa, db, da, ab = _dot(a, a), _dot(d, b), _dot(d, a), _dot(a, b)
    for j in range(9):
        g_b2 = dotc([2.0 * b[k] for k in range(3)], gb[j])
        g_kk = g_kk_b2 * g_b2
        g_ab = dotc(b, ga[j]) + dotc(a, gb[j])
        g_kx.append(kk * g_ab + ab * g_kk)
        g_aa = dotc([2.0 * a[k] for k in range(3)], ga[j])
        g_db = dotc(b, gd[j]) + dotc(d, gb[j])
        g_ky.append(2.0 / 3.0 * kk * g_aa + 1.0 / 3.0 * kk * g_db + 1.0 / 3.0 * (2 * aa + db) * g_kk)
        g_da = dotc(d, ga[j]) + dotc(a, gd[j])
        g_kz.append(kk * g_da + da * g_kk)
    g_p = [g_ky[j] - 2.0 * kx * g_kx[j] for j in range(9)]
    g_q = [6.0 * kx * kx * g_kx[j] - 3.0 * ky * g_kx[j] - 3.0 * kx * g_ky[j] + g_kz[j] for j in range(9)]
    # xyz derivatives: grad_d_xyz = -I; grad_db_xyz = b * (-I) (row-vector * mat3)
    g_db_xyz = [-b[k] for k in range(3)]
    g_da_xyz = [-a[k] for k in range(3)]
    g_ky_xyz = [1.0 / 3.0 * kk * g_db_xyz[k] for k in range(3)]
    g_kz_xyz = [kk * g_da_xyz[k] for k in range(3)]
    g_p_xyz = g_ky_xyz
    g_q_xyz = [-3.0 * kx * g_ky_xyz[k] + g_kz_xyz[k] for k in range(3)]

    one_root = s >= 0.0
    # ---- branch one root (category 1 / 3)
    s1 = torch.where(one_root, s, torch.zeros_like(s))
    h = torch.sqrt(s1 + 1e-8)
    xx = (h - q) * 0.5
    xy = -(h + q) * 0.5
    t1 = _clamp((_cbrt_signed(xx) + _cbrt_signed(xy)) - kx, 0.0, 1.0)
    m1v = [d[k] + (c[k] + b[k] * t1) * t1 for k in range(3)]
    res1 = _dot(m1v, m1v)
    # ---- branch three roots (category 2 / 3)
    pn = torch.where(one_root, -torch.ones_like(pp), pp)
    qn = torch.where(one_root, zero, q)
    zz = torch.sqrt(-pn + 1e-8)
    vv = torch.acos(torch.clamp(torch.where(one_root, zero, qn / (pn * zz * 2.0 + 1e-8)), -1.0, 1.0)) / 3.0
    mm = torch.cos(vv)
    nn = torch.sin(vv) * 1.732050808
    ta = _clamp((mm + nn) * zz - kx, 0.0, 1.0)
    tb = _clamp((-mm - nn) * zz - kx, 0.0, 1.0)
    ma = [d[k] + (c[k] + b[k] * ta) * ta for k in range(3)]
    mb = [d[k] + (c[k] + b[k] * tb) * tb for k in range(3)]
    disa, disb = _dot(ma, ma), _dot(mb, mb)
    use_b = disb < disa

    # category 1 (one root, interior t)
    rs1 = 0.5 * torch.rsqrt(res1 + 1e-8)
    gm_t1 = [c[k] + 2.0 * b[k] * t1 for k in range(3)]
    ax_ = torch.abs(xx)
    ay_ = torch.abs(xy)
    # comma-operator results: the xyz path divides x.x by 6, the params path by 3
    g_uvx_h_xyz = xx / 6.0
    g_uvy_h_xyz = -1.0 * xy / 6.0
    g_uvx_h_prm = xx / 3.0
    g_uvy_h_prm = -1.0 * xy / 3.0
    del ax_, ay_
    g_h_s = 0.5 * torch.rsqrt(s1 + 1e-8)
    g_s_q = 2.0 * q
    g_s_p = 12.0 * pp * pp

    def t_grad_cat1(gux, guy, gq, gp):
        gus = gux * g_h_s
        gvs = guy * g_h_s
        return (-gux * gq + gus * (g_s_q * gq + g_s_p * gp)) + (-guy * gq + gvs * (g_s_q * gq + g_s_p * gp))

    g_t1_xyz = [t_grad_cat1(g_uvx_h_xyz, g_uvy_h_xyz, g_q_xyz[k], g_p_xyz[k]) for k in range(3)]
    # grad_ansx_m * (grad_d_xyz + outer(grad_t_xyz, grad_m_t)), row vector times mat3:
    # result_j = sum_i 2 m_i (-delta_ij + g_t_i * gm_t_j)
    def row_times(mv, gt, gmt):
        two_m = [2.0 * mv[k] for k in range(3)]
        return [-two_m[j] + (two_m[0] * gt[0] + two_m[1] * gt[1] + two_m[2] * gt[2]) * gmt[j] for j in range(3)]

    cat1_xyz = [e * rs1 for e in row_times(m1v, g_t1_xyz, gm_t1)]
    cat1_xyz = [cat1_xyz[k] - (r2 - r1) * g_t1_xyz[k] for k in range(3)]
    cat1_prm = []
    for j in range(9):
        g_tp = t_grad_cat1(g_uvx_h_prm, g_uvy_h_prm, g_q[j], g_p[j]) - g_kx[j]
        gmp = [gd[j][k] + gc[j][k] * t1 + gm_t1[k] * g_tp + gb[j][k] * (t1 * t1) for k in range(3)]
        val = (2.0 * m1v[0] * gmp[0] + 2.0 * m1v[1] * gmp[1] + 2.0 * m1v[2] * gmp[2]) * rs1
        cat1_prm.append(val - (r2 - r1) * g_tp)

    # category 2 (three roots, interior t): shared partials, strokes_sdf.h:1073-1090
    def cat2(t, mv, res):
        rs = 0.5 * torch.rsqrt(res + 1e-8)
        gh_t = [c[k] + 2.0 * b[k] * t for k in range(3)]
        g_t_n = -1.0 * zz
        g_t_m = -1.0 * zz
        g_t_z = -mm - nn
        g_z_p = -0.5 * torch.rsqrt(-pn + 1e-8)
        g_n_v = torch.cos(vv) * 1.732050808
        g_m_v = -torch.sin(vv)
        den = pn * zz * 2.0 + 1e-8
        rq = torch.rsqrt(1.0 - qn * qn / (pn * pn * zz * zz * 4.0) + 1e-8)
        g_v_q = 1.0 / den * (-1.0) * rq / 3.0
        g_v_p = 1.0 / pn * qn / den * rq / 3.0
        g_v_z = 1.0 / zz * qn / den * rq / 3.0
        g_v_xyz = [g_q_xyz[k] * g_v_q + g_p_xyz[k] * g_v_p + g_p_xyz[k] * (g_v_z * g_z_p) for k in range(3)]
        g_t_xyz = [(g_v_xyz[k] * (g_t_n * g_n_v) + g_v_xyz[k] * (g_t_m * g_m_v)) + g_p_xyz[k] * (g_t_z * g_z_p)
                   for k in range(3)]
        xyz = [e * rs for e in row_times(mv, g_t_xyz, gh_t)]
        xyz = [xyz[k] - (r2 - r1) * g_t_xyz[k] for k in range(3)]
        prm_g = []
        for j in range(9):
            g_vpi = g_v_q * g_q[j] + g_v_p * g_p[j] + g_v_z * g_z_p * g_p[j]
            g_tp = (g_t_n * g_n_v * g_vpi + g_t_m * g_m_v * g_vpi) + g_t_z * g_z_p * g_p[j] - g_kx[j]
            ghp = [gd[j][k] + c[k] * g_tp + gc[j][k] * t + b[k] * 2.0 * t * g_tp + gb[j][k] * t * t for k in range(3)]
            val = (2.0 * mv[0] * ghp[0] + 2.0 * mv[1] * ghp[1] + 2.0 * mv[2] * ghp[2]) * rs
            prm_g.append(val - (r2 - r1) * g_tp)
        return xyz, prm_g

    cat2a_xyz, cat2a_prm = cat2(ta, ma, disa)
    cat2b_xyz, cat2b_prm = cat2(tb, mb, disb)

    # category 3 (clamped t), strokes_sdf.h:1110-1114, 1166-1170
    def cat3(t, res):
        rs = 0.5 * torch.rsqrt(res + 1e-8)
        xyz = [d[k] * (-1.0 * torch.rsqrt(res + 1e-8)) for k in range(3)]
        at0 = t == 0.0
        prm_g = []
        for j in range(9):
            v0 = dotc([2.0 * d[k] for k in range(3)], gd[j])
            dcb = [d[k] + c[k] + b[k] for k in range(3)]
            gsum = [gd[j][k] + gc[j][k] + gb[j][k] for k in range(3)]
            v1 = dotc([2.0 * dcb[k] for k in range(3)], gsum)
            prm_g.append(torch.where(at0, v0, v1) * rs)
        return xyz, prm_g

    c31_xyz, c31_prm = cat3(t1, res1)
    c3a_xyz, c3a_prm = cat3(ta, disa)
    c3b_xyz, c3b_prm = cat3(tb, disb)

    t1_clamped = (t1 == 0.0) | (t1 == 1.0)
    ta_clamped = (ta == 0.0) | (ta == 1.0)
    tb_clamped = (tb == 0.0) | (tb == 1.0)

    def pick(v1c, v1i, vac, vai, vbc, vbi):
        one_v = torch.where(t1_clamped, v1c, v1i)
        a_v = torch.where(ta_clamped, vac, vai)
        b_v = torch.where(tb_clamped, vbc, vbi)
        return torch.where(one_root, one_v, torch.where(use_b, b_v, a_v))

    gxyz = [pick(c31_xyz[k], cat1_xyz[k], c3a_xyz[k], cat2a_xyz[k], c3b_xyz[k], cat2b_xyz[k]) for k in range(3)]
    gprm = [pick(c31_prm[j], cat1_prm[j], c3a_prm[j], cat2a_prm[j], c3b_prm[j], cat2b_prm[j]) for j in range(9)]
    # radius params: -(1 - t), -t, except the three-root / t.x interior branch
    # which uses -(t + 1) for r1 (strokes_sdf.h:1046)
    tsel = torch.where(one_root, t1, torch.where(use_b, tb, ta))
    bug = (~one_root) & (~use_b) & (~ta_clamped)
    g_r1 = torch.where(bug, -(tsel + 1.0), -(-tsel + 1.0))
    g_r2 = -tsel
    return gxyz, torch.stack(gprm + [g_r1, g_r2], -1)


# ----------------------------------------------------------------------------
# colour fields (strokes_color.h) and SH basis (sh.h)
# ----------------------------------------------------------------------------

_INV_SQRT_PI = 1.0 / math.sqrt(math.pi)
SH_C0 = 0.5 * _INV_SQRT_PI                      # 1/(2 sqrt(pi))
SH_C1 = 0.5 * math.sqrt(3.0) * _INV_SQRT_PI     # sqrt(3)/(2 sqrt(pi))
SH_C2 = 0.5 * math.sqrt(15.0) * _INV_SQRT_PI    # sqrt(15)/(2 sqrt(pi))
SH_C3 = 0.75 * math.sqrt(5.0) * _INV_SQRT_PI    # 3 sqrt(5)/(4 sqrt(pi))
SH_C4 = 0.25 * math.sqrt(5.0) * _INV_SQRT_PI    # sqrt(5)/(4 sqrt(pi))
SH_C5 = 0.25 * math.sqrt(15.0) * _INV_SQRT_PI   # sqrt(15)/(4 sqrt(pi))


def sh_basis(degree, d):
    """Real SH basis up to ``degree`` (4 or 9 terms), sh.h:8-30 layout and signs."""
    x, y, z = d
    out = [torch.full_like(x, SH_C0)]
    if degree <= 1:
        return out
    out += [-SH_C1 * y, SH_C1 * z, -SH_C1 * x]
    if degree <= 2:
        return out
    out += [SH_C2 * (x * y), -SH_C2 * (y * z), SH_C3 * (z * z) - SH_C4, -SH_C2 * (x * z),
            SH_C5 * (x * x) - SH_C5 * (y * y)]
    return out


def color_forward(color_id, pos_world, pos_unit, dir_raw, dir_unit, cp):
    """ColorField<...>::get_color, strokes_color.h:40-154.  Returns [..., 3]."""
    if color_id == 0:
        return torch.stack([cp[..., 0], cp[..., 1], cp[..., 2]], -1).expand(*pos_unit[0].shape, 3)
    if color_id == 1:
        p0 = [cp[..., 0], cp[..., 1], cp[..., 2]]
        p1 = [cp[..., 3], cp[..., 4], cp[..., 5]]
        v0 = [pos_unit[k] - p0[k] for k in range(3)]
        v1 = [p1[k] - p0[k] for k in range(3)]
        t = _sigmoid((_dot(v0, v1) / _dot(v1, v1)) * 5.0 - 2.5)
        return torch.stack([cp[..., 6 + k] * (1.0 - t) + cp[..., 9 + k] * t for k in range(3)], -1)
    deg = 2 if color_id == 2 else 3
    w = sh_basis(deg, dir_raw)
    rgb = []
    for k in range(3):
        acc = None
        for i, wi in enumerate(w):
            term = cp[..., 3 * i + k] * wi
            acc = term if acc is None else acc + term
        rgb.append(_sigmoid(acc))
    return torch.stack(rgb, -1)


def color_grad_reference(color_id, pos_unit, dir_raw, cp, gc):
    """ColorField<...>::grad_color, strokes_color.h:44-170 (params only, no pos term)."""
    if color_id == 0:
        return gc
    if color_id == 1:
        p0 = [cp[..., 0], cp[..., 1], cp[..., 2]]
        p1 = [cp[..., 3], cp[..., 4], cp[..., 5]]
        c0 = [cp[..., 6], cp[..., 7], cp[..., 8]]
        c1 = [cp[..., 9], cp[..., 10], cp[..., 11]]
        v0 = [pos_unit[k] - p0[k] for k in range(3)]
        v1 = [p1[k] - p0[k] for k in range(3)]
        d0, d1 = _dot(v0, v1), _dot(v1, v1)
        t = _sigmoid((d0 / d1) * 5.0 - 2.5)
        g = [gc[..., k] for k in range(3)]
        dt = _dot(g, [c1[k] - c0[k] for k in range(3)])
        dtr = 5.0 * t * (1.0 - t) * dt
        dd0 = dtr / d1
        dd1 = -dtr * d0 / (d1 * d1)
        dv1 = [dd0 * v0[k] + dd1 * 2.0 * v1[k] for k in range(3)]
        dv0 = [dd0 * v1[k] for k in range(3)]
        out = [-dv0[k] - dv1[k] for k in range(3)] + dv1 + [g[k] * (1.0 - t) for k in range(3)] + \
              [g[k] * t for k in range(3)]
        return torch.stack(out, -1)
    deg = 2 if color_id == 2 else 3
    w = sh_basis(deg, dir_raw)
    rgb = color_forward(color_id, None, [dir_raw[0]], dir_raw, None, cp)
    grgb = gc * rgb * (1.0 - rgb)
    out = []
    for wi in w:
        out += [grgb[..., k] * wi for k in range(3)]
    return torch.stack(out, -1)


# ----------------------------------------------------------------------------
# stroke forward (strokes_forward.cu:35-109)
# ----------------------------------------------------------------------------

def laplace_cdf(s):
    """common.h:47-51."""
    return torch.where(s <= 0, 0.5 * torch.exp(torch.clamp_max(s, 0.0)), 1.0 - 0.5 * torch.exp(-torch.clamp_min(s, 0.0)))


def _radius_scale_factor(base, p, prm_base, scale):
    """|∇sdf ⊙ s| / |∇sdf|, strokes_forward.cu:73-80 (direction of the reference grad)."""
    g, _ = grad_sdf_reference(base, p, prm_base)
    gs = [g[k] * scale[..., k] for k in range(3)]
    return _safe_sqrt(_dot(gs, gs)) / _safe_sqrt(_dot(g, g))


def stroke_forward(x, radius, viewdir, shape_params, color_params, sdf_id, color_id, sdf_delta,
                   use_laplace_transform=False, inv_scale_radius=False, exact_radius_grad=False):
    """Per-(point, stroke) alpha / colour / sdf, strokes_forward.cu:35-109.

    x [N,3], radius [N], viewdir [R,3] (N = R * samples), shape_params [S,P],
    color_params [S,Q].  Returns dict(alpha [N,S], color [N,S,3], sdf [N,S],
    unit_pos list, ...).
    """
    base, t, r, ss, ms = decode_sdf_id(sdf_id)
    n_points = x.shape[0]
    n_samples = n_points // viewdir.shape[0]
    xs = x[:, None, :]
    prm = shape_params[None, :, :]
    cp = color_params[None, :, :]
    dir_raw = _v(viewdir.repeat_interleave(n_samples, 0)[:, None, :])
    p = inverse_transform(xs, prm, sdf_id)
    p = [pi.expand(n_points, shape_params.shape[0]) for pi in p]
    nb = NUM_BASE_PARAMS[base]
    prm_base = prm[..., :nb]
    sdf = base_sdf(base, p, prm_base)
    radii = radius[:, None].expand_as(sdf)
    if inv_scale_radius:
        _, scale, _, _ = split_shape_params(prm, sdf_id)
        if ss:
            radii = radii / scale[..., 0]
        elif ms:
            if exact_radius_grad:
                fac = _radius_factor_exact(base, p, prm_base, scale)
            else:
                fac = _radius_scale_factor(base, [pi.detach() for pi in p], prm_base.detach(), scale)
            radii = radii / fac
    sdf_scale = (2.0 if use_laplace_transform else 0.5) / (sdf_delta * radii)
    if sdf_delta > 0:
        if use_laplace_transform:
            alpha = laplace_cdf(-sdf * sdf_scale)
        else:
            alpha = _clamp(-sdf * sdf_scale + 0.5, 0.0, 0.9999)
    else:
        alpha = (sdf <= 0).to(sdf.dtype)
    if color_id == 1:
        color = color_forward(color_id, None, p, dir_raw, None, cp)
    else:
        color = color_forward(color_id, None, p, dir_raw, None, cp)
    return dict(alpha=alpha, color=color, sdf=sdf, pos=p, radii=radii, sdf_scale=sdf_scale)


def _radius_factor_exact(base, p, prm_base, scale):
    """Differentiable |∇sdf ⊙ s| / |∇sdf| for the exact adjoint.

    For the sphere base ∇sdf ∝ p, so the factor is |p ⊙ s| / |p| and depends on
    the stroke parameters through p.  For the other bases the reference gradient
    direction is piecewise constant or is differentiated through torch here.
    """
    if base == 0:
        ps = [p[k] * scale[..., k] for k in range(3)]
        return torch.sqrt(_dot(ps, ps)) / torch.sqrt(_dot(p, p))
    g, _ = grad_sdf_reference(base, p, prm_base)
    gs = [g[k] * scale[..., k] for k in range(3)]
    return _safe_sqrt(_dot(gs, gs)) / _safe_sqrt(_dot(g, g))


def stroke_backward_reference(grad_alpha, grad_color, grad_sdf, x, radius, viewdir, alpha, shape_params,
                              color_params, sdf_id, color_id, sdf_delta, use_laplace_transform=False,
                              inv_scale_radius=False, need_grad_x=True):
    """Restatement of stroke_backward_kernel, strokes_backward.cu:38-157."""
    base, t, r, ss, ms = decode_sdf_id(sdf_id)
    fw = stroke_forward(x, radius, viewdir, shape_params, color_params, sdf_id, color_id, sdf_delta,
                        use_laplace_transform, inv_scale_radius)
    p = fw['pos']
    sdf_scale = fw['sdf_scale']
    nb = NUM_BASE_PARAMS[base]
    prm = shape_params[None, :, :]
    cp = color_params[None, :, :]
    n_points = x.shape[0]
    n_samples = n_points // viewdir.shape[0]
    dir_raw = _v(viewdir.repeat_interleave(n_samples, 0)[:, None, :])
    zero = torch.zeros_like(alpha)
    if sdf_delta > 0:
        if use_laplace_transform:
            da = torch.where(alpha < 0.5, alpha, 1.0 - alpha) * -sdf_scale
        else:
            da = torch.where((0.0 < alpha) & (alpha < 0.9999), -sdf_scale, zero)
    else:
        da = zero
    dl_dsdf = grad_alpha * da
    if grad_sdf is not None and grad_sdf.numel() > 0:
        dl_dsdf = dl_dsdf + grad_sdf
    # colour params
    gcp = color_grad_reference(color_id, p, dir_raw, cp, grad_color)
    grad_color_params = gcp.sum(0)
    if grad_color_params.dim() == 1:
        grad_color_params = grad_color_params[None].expand(shape_params.shape[0], -1)
    # base sdf adjoint
    gpos, gbase = grad_sdf_reference(base, p, prm[..., :nb])
    if isinstance(gpos, tuple):  # cubic / catmull: dSDF/dpos = -grad_SDF * 2 v
        v = gpos[1]
        dsdf_dpos = [-dl_dsdf * 2.0 * v[k] for k in range(3)]
        gbase = torch.cat([gbase[..., :-2] * dl_dsdf[..., None], gbase[..., -2:] * dl_dsdf[..., None]], -1)
    else:
        dsdf_dpos = gpos
        if gbase is not None:
            gbase = gbase * dl_dsdf[..., None]
    dl_dpos = [dl_dsdf * dsdf_dpos[k] for k in range(3)]
    S = shape_params.shape[0]
    grads = []
    if gbase is not None:
        grads.append(gbase.sum(0))
    _, scale, angle, trans = split_shape_params(prm, sdf_id)
    pos = list(p)
    if ss:
        s = scale[..., 0]
        dl_dpos = [g / s for g in dl_dpos]
        grads.append((-(dl_dpos[0] * pos[0]) + -(dl_dpos[1] * pos[1]) + -(dl_dpos[2] * pos[2])).sum(0)[:, None])
        pos = [pi * s for pi in pos]
    elif ms:
        dl_dpos = [dl_dpos[k] / scale[..., k] for k in range(3)]
        grads.append(torch.stack([(-dl_dpos[k] * pos[k]).sum(0) for k in range(3)], -1))
        pos = [pos[k] * scale[..., k] for k in range(3)]
    if r:
        re = rotation_entries(angle)
        q = [re[k][0] * pos[0] + re[k][1] * pos[1] + re[k][2] * pos[2] for k in range(3)]
        grads.append(torch.stack(_grad_angle(dl_dpos, q, angle), -1).sum(0))
        dl_dpos = [re[k][0] * dl_dpos[0] + re[k][1] * dl_dpos[1] + re[k][2] * dl_dpos[2] for k in range(3)]
        pos = q
    if t:
        grads.append(torch.stack([(-dl_dpos[k]).sum(0) for k in range(3)], -1))
    grad_shape = torch.cat(grads, -1) if grads else torch.zeros_like(shape_params)
    grad_x = torch.stack([dl_dpos[k].sum(1) for k in range(3)], -1) if need_grad_x else None
    return grad_shape.reshape(S, -1), grad_color_params, grad_x


def _grad_angle(g, q, angle):
    """grad_angle_rotate_point<true>, strokes_sdf.h:92-162: dL/dangle for p = Rᵀ q."""
    sx, sy, sz = torch.sin(angle[..., 0]), torch.sin(angle[..., 1]), torch.sin(angle[..., 2])
    cx, cy, cz = torch.cos(angle[..., 0]), torch.cos(angle[..., 1]), torch.cos(angle[..., 2])
    zero = torch.zeros_like(sx)
    dR = [[(zero, -sy * cz, -cy * sz),
           (cx * sy * cz + sx * sz, sx * cy * cz, -sx * sy * sz - cx * cz),
           (cx * sz - sx * sy * cz, cx * cy * cz, sx * cz - cx * sy * sz)],
          [(zero, -sy * sz, cy * cz),
           (-sx * cz + cx * sy * sz, sx * cy * sz, -cx * sz + sx * sy * cz),
           (-sx * sy * sz - cx * cz, cx * cy * sz, cx * sy * cz + sx * sz)],
          [(zero, -cy, zero),
           (cx * cy, -sx * sy, zero),
           (-sx * cy, -cx * sy, zero)]]
    out = [0.0, 0.0, 0.0]
    # p_j = sum_i R_ij q_i  ->  dL/da = sum_{i,j} g_j q_i dR_ij/da
    for j in range(3):
        for i in range(3):
            for a in range(3):
                out[a] = out[a] + g[j] * q[i] * dR[i][j][a]
    return out


def stroke_backward_exact(grad_alpha, grad_color, grad_sdf, x, radius, viewdir, shape_params, color_params,
                          sdf_id, color_id, sdf_delta, use_laplace_transform=False, inv_scale_radius=False,
                          need_grad_x=True):
    """True derivative of :func:`stroke_forward` by torch.autograd (use float64)."""
    with torch.enable_grad():
        xs = x.detach().clone().requires_grad_(need_grad_x)
        sp = shape_params.detach().clone().requires_grad_(True)
        cpp = color_params.detach().clone().requires_grad_(True)
        fw = stroke_forward(xs, radius, viewdir, sp, cpp, sdf_id, color_id, sdf_delta, use_laplace_transform,
                            inv_scale_radius, exact_radius_grad=True)
        loss = (fw['alpha'] * grad_alpha).sum() + (fw['color'] * grad_color).sum()
        if grad_sdf is not None and grad_sdf.numel() > 0:
            loss = loss + (fw['sdf'] * grad_sdf).sum()
        inputs = [sp, cpp] + ([xs] if need_grad_x else [])
        gr = torch.autograd.grad(loss, inputs, allow_unused=True)
    out = [g if g is not None else torch.zeros_like(i) for g, i in zip(gr, inputs)]
    return out[0], out[1], (out[2] if need_grad_x else None)

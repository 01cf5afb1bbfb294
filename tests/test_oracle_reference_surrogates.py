"""The oracle's reference-mode adjoints against autograd of surrogate forwards.

The reference's hand-written adjoints (``compositing.cu:61-157``,
``strokes_backward.cu:38-157`` and the ``grad_sdf`` members of
``strokes_sdf.h``) are not derivatives of the forward everywhere.  The oracle
restates them term by term (``oracle.compose.compose_backward_reference``,
``oracle.strokes.stroke_backward_reference``).  Here each one is rebuilt by a
different route: a surrogate forward that has the real forward's values but
whose *local* derivatives are the reference's deviating ones (straight-through
terms ``v.detach() + f - f.detach()``), differentiated by torch.autograd in
float64.  Autograd does all the chaining (transform, transfer, compositing,
cubic set-up), so a slip in the oracle's hand chaining, a wrong sign or a
dropped term fails here; only the local deviations are written by hand, each
next to the reference line it encodes:

* H1   compositing colour -> alpha term (``compositing.cu:143-155``);
* H3   inverse-scaled radius: no derivative through the radius
       (``strokes_forward.cu:69-81`` vs ``strokes_backward.cu``);
* H10  box outside: gradient of the squared length (``strokes_sdf.h:279-299``);
* H11  gradient_rgb: no position dependence of the colour
       (``strokes_color.h:66-92``);
* H12  capped torus (``strokes_sdf.h:382-420``);
* H2   cubic Bezier / Catmull-Rom parameters: squared-distance gradient at
       fixed t, Catmull-Rom knot lengths differentiated per component
       (``strokes_sdf.h:1287-1318, 1408-1445``);
* H2   quadratic Bezier parameters: the comma-operator cube-root slopes, the
       opposite q-sign of the second root's term, the three-root branch
       differentiated as its second candidate with its own regularisers, the
       (t + 1) radius factor of the first candidate
       (``strokes_sdf.h:1040-1172``).
"""
import math

import pytest
import torch

from oracle import compose as ocompose
from oracle import strokes as ostrokes

D64 = torch.float64
SQ3 = 1.732050808


def ste(value, surrogate):
    """value's value, surrogate's derivative."""
    return value.detach() + surrogate - surrogate.detach()


def slope(x, k):
    """x's value; derivative k (detached) with respect to x."""
    return x.detach() + (x - x.detach()) * k.detach()


def rel(a, b):
    return float((a - b).norm() / b.norm().clamp_min(1e-300))


# ----------------------------------------------------------------------------
# H1 compositing
# ----------------------------------------------------------------------------

@pytest.mark.parametrize('seed', [0, 1, 2])
def test_compose_reference_adjoint_is_the_autograd_of_its_surrogate(seed):
    g = torch.Generator().manual_seed(seed)
    N, S = 64, 9
    a = torch.rand(N, S, generator=g, dtype=D64) * 0.95
    a[torch.rand(N, S, generator=g) < 0.3] = 0.0
    c = torch.rand(N, S, 3, generator=g, dtype=D64) * 1.2 - 0.1  # some normalised colours leave [0, 1]
    dp = torch.exp(torch.rand(S, generator=g, dtype=D64) * 4.6)
    gD = torch.randn(N, generator=g, dtype=D64)
    gC = torch.randn(N, 3, generator=g, dtype=D64)
    ra, rc, rd = ocompose.compose_backward_reference(gD, gC, a, c, dp)

    av, cv, dv = (t.clone().requires_grad_(True) for t in (a, c, dp))
    # strokes are composited from index S-1 down to 0: transmittance before stroke k
    one_m = (1.0 - av).flip(-1)
    Tb = torch.cat([torch.ones(N, 1, dtype=D64), torch.cumprod(one_m, -1)[:, :-1]], -1).flip(-1)
    w = av * Tb
    D = (w * dv).sum(-1)
    Craw = (w[..., None] * cv).sum(-2)
    Tn = torch.prod(1.0 - av, -1)
    one_eps = float(torch.tensor(1.0 + 1e-6, dtype=torch.float32))  # 1.0f + 1e-6f in single precision
    Od = (one_eps - Tn).detach()[:, None]
    in_range = ((Craw / Od >= 0) & (Craw / Od <= 1)).detach()
    gCm = gC * in_range
    # compositing.cu:148-154: (T_after c_k (O - a_k) - suffix) / ((1 - a_k) O^2); with
    # d C_raw / d a_k = (T_after c_k - suffix) / (1 - a_k) this is
    # d/da_k [C_raw / O^2 + c_k T_before (a_k (O - 1) - a_k^2 / 2) / O^2] at frozen O, T_before
    Sa = Craw / Od ** 2 + ((cv.detach() * Tb.detach()[..., None]) *
                           (av * (Od - 1.0) - av * av / 2.0)[..., None]).sum(-2) / Od ** 2
    ga, = torch.autograd.grad((gD * D).sum() + (gCm * Sa).sum(), av, retain_graph=True)
    # colours: dL_dcolor * final_color_scale * weight; density params: dL_ddensity * weight
    gc, gd = torch.autograd.grad((gD * D).sum() + (gCm * Craw / Od).sum(), [cv, dv])
    # the reference forms the suffix sums by subtraction from the totals (compositing.cu:138,145),
    # which loses ~1e-9 relative in float64 where 1 / O^2 is large
    assert rel(ra, ga) < 1e-8
    assert rel(rc, gc) < 1e-12
    assert rel(rd, gd) < 1e-12
    # the colour term is not the derivative of the normalised colour (it is for one stroke)
    ex, _, _ = ocompose.compose_backward_exact(gD, gC, a, c, dp)
    assert rel(ra, ex) > 1e-3


# ----------------------------------------------------------------------------
# stroke adjoints
# ----------------------------------------------------------------------------

def _g_cube(p, prm):
    q = [torch.abs(p[i]) - 1.0 for i in range(3)]
    inside = torch.maximum(q[0], torch.maximum(q[1], q[2])) <= 0.0
    inner = torch.maximum(q[0], torch.maximum(q[1], q[2]))
    outer = sum(torch.clamp_min(q[i], 0.0) ** 2 for i in range(3))  # strokes_sdf.h:288-296: 2 max(q,0) sign
    return torch.where(inside, inner, outer)


def _g_capped_torus(p, prm):
    th, rb = prm[..., 0], prm[..., 1]
    s, c = torch.sin(th), torch.cos(th)
    ax = torch.abs(p[0])
    cross = ax.detach() * c.detach() - p[1].detach() * s.detach()
    k = torch.where(cross > 0, ax * s + p[1] * c, torch.sqrt(ax * ax + p[1] * p[1])).detach()
    inv = torch.rsqrt(p[0] ** 2 + p[1] ** 2 + p[2] ** 2 + 1.0 - 2.0 * k + 1e-8).detach()
    kpos = k > 0
    sg = torch.where(p[0] < 0, -1.0, 1.0).detach()
    # :394-404 (k > 0): grad = (p - (sign(x) sin, cos, 0)) inv, whichever form k took;
    # :406-414 (k <= 0): grad = (p + (x, y, 0)) inv
    ks = torch.where(kpos, sg * p[0] * s.detach() + p[1] * c.detach(), -(p[0] ** 2 + p[1] ** 2) / 2.0)
    e = 0.5 * (p[0] ** 2 + p[1] ** 2 + p[2] ** 2 - 2.0 * ks) * inv
    # :398 parameter 0: -(x cos + y sin) without the 1/length factor
    f = torch.where(kpos, -(ax.detach() * torch.sin(th) - p[1].detach() * torch.cos(th)), torch.zeros_like(th))
    return e + f - rb


def _curve_t(base, p, prm):
    if base == 10:
        return ostrokes.cubic_bezier_t(p, prm).detach()
    return ostrokes.catmull_rom_t(p, prm).detach()


def _g_cubic(p, prm):
    # strokes_sdf.h:1298-1314: grad_SDF * 2 (P - pos) dP/dparam at the polyline's t
    t = _curve_t(10, p, prm)
    A, B, C, D = ostrokes._cubic_ctrl(prm)
    P = ostrokes._cubic_point(t, A, B, C, D)
    return sum((P[k] - p[k]) ** 2 for k in range(3)) - (prm[..., 12] + t * (prm[..., 13] - prm[..., 12]))


def _g_catmull(p, prm):
    # strokes_sdf.h:1408-1445: the knot lengths t_ij = |p_i - p_j|^(1/2) enter component k of the
    # tangents only through component k of p_i - p_j, with d t_ij / d p = (p_i - p_j) / (2 (|.|^2 + 1e-8)^(3/4))
    t = _curve_t(11, p, prm)
    p0, p1, p2, p3 = ostrokes._cubic_ctrl(prm)

    def knot(u, v, k):
        d = [(u[j] - v[j]) if j == k else (u[j] - v[j]).detach() for j in range(3)]
        real = torch.sqrt(torch.sqrt(sum(x.detach() ** 2 for x in d)) + 1e-8)
        return ste(real, (sum(x * x for x in d) + 1e-8) ** 0.25)

    out = 0.0
    for k in range(3):
        t01, t12, t23 = knot(p0, p1, k), knot(p1, p2, k), knot(p2, p3, k)
        m1 = p2[k] - p1[k] + t12 * ((p1[k] - p0[k]) / t01 - (p2[k] - p0[k]) / (t01 + t12))
        m2 = p2[k] - p1[k] + t12 * ((p3[k] - p2[k]) / t23 - (p3[k] - p1[k]) / (t12 + t23))
        # :1422 dm2_dt12 has (p2 - p1) / t23 where the derivative has (p3 - p2) / t23
        m2 = m2 + (t12 - t12.detach()) * ((p2[k] - p1[k]) - (p3[k] - p2[k])).detach() / t23.detach()
        A = 2.0 * (p1[k] - p2[k]) + m1 + m2
        B = 3.0 * (p2[k] - p1[k]) - 2.0 * m1 - m2
        P = ((A * t + B) * t + m1) * t + p1[k]
        out = out + (P - p[k]) ** 2
    return out - (prm[..., 12] + t * (prm[..., 13] - prm[..., 12]))


def _quad_parts(p, prm, k1):
    """Pieces of the closed form with the reference's local slopes: t's surrogate, the pick and m(t).

    k1: the cube-root slope per unit of the root's argument (2/3 x on the parameter path, where the
    comma operator of :1119-1120 leaves x / 3 per unit of h; 1/3 x on the position path, :1063-1064)."""
    A = [prm[..., 0], prm[..., 1], prm[..., 2]]
    B = [prm[..., 3], prm[..., 4], prm[..., 5]]
    C = [prm[..., 6], prm[..., 7], prm[..., 8]]
    a = [B[i] - A[i] for i in range(3)]
    b = [A[i] - 2.0 * B[i] + C[i] for i in range(3)]
    c = [2.0 * a[i] for i in range(3)]
    d = [A[i] - p[i] for i in range(3)]
    dot = lambda u, v: u[0] * v[0] + u[1] * v[1] + u[2] * v[2]
    kk = 1.0 / dot(b, b)
    kx = kk * dot(a, b)
    ky = kk * (2.0 * dot(a, a) + dot(d, b)) / 3.0
    kz = kk * dot(d, a)
    pp = ky - kx * kx
    q = kx * (2.0 * kx * kx - 3.0 * ky) + kz
    s = q * q + 4.0 * pp * pp * pp
    # the forward's own pick (values only)
    with torch.no_grad():
        d2r, t_real = ostrokes.quad_bezier_dist2_t([x.detach() for x in p], prm.detach())
        one = s >= 0
        z0 = torch.sqrt(torch.where(one, torch.ones_like(pp), -pp) + 1e-8)
        arg = torch.clamp(torch.where(one, torch.zeros_like(q), q) /
                          (torch.where(one, -torch.ones_like(pp), pp) * z0 * 2.0 + 1e-8), -1, 1)
        v0 = torch.acos(arg) / 3.0
        ta = torch.clamp((torch.cos(v0) + torch.sin(v0) * SQ3) * z0 - kx, 0, 1)
        tb = torch.clamp((-torch.sin(v0) * SQ3 - torch.cos(v0)) * z0 - kx, 0, 1)

        def d2(t):
            w = [d[i] + (c[i] + b[i] * t) * t for i in range(3)]
            return dot(w, w)
        use_b = ~one & (d2(tb) < d2(ta))
        interior = (t_real > 0) & (t_real < 1)
    # one root (:933-958): t = cbrt(x0) + cbrt(x1) - kx, x0 = (h - q)/2, x1 = -(h + q)/2,
    # h = sqrt(s + 1e-8); the second root's q-term has the sign of -(its h-term)
    # (grad_uvy_q = -grad_uvy_h), i.e. x1 enters as -(h - q)/2
    h = torch.sqrt(torch.clamp_min(s, 0.0) + 1e-8)
    x0r = ((h - q) * 0.5).detach()
    x1r = (-(h + q) * 0.5).detach()
    t1 = slope((h - q) * 0.5, k1 * x0r) + slope(-(h - q) * 0.5, k1 * x1r) - kx
    # three roots (grad_*_category_2): both candidates differentiated as the second,
    # -(m + n) z - kx, with v's slopes from :1136-1138
    pn = torch.where(one, -torch.ones_like(pp), pp)  # the one-root rows never use this branch
    z = torch.sqrt(-pn + 1e-8)
    zd, pd, qd = z.detach(), pn.detach(), torch.where(one, torch.zeros_like(q), q).detach()
    den = pd * zd * 2.0 + 1e-8
    rs = torch.rsqrt(torch.clamp_min(1.0 - qd * qd / (pd * pd * zd * zd * 4.0) + 1e-8, 1e-300))
    gvq = -1.0 / den * rs / 3.0
    gvp = 1.0 / pd * qd / den * rs / 3.0
    gvz = 1.0 / zd * qd / den * rs / 3.0
    v = v0 + (q - q.detach()) * gvq + (pn - pd) * gvp + (z - zd) * gvz
    t2 = -(torch.cos(v) + torch.sin(v) * SQ3) * z - kx
    t_s = ste(t_real, torch.where(interior, torch.where(one, t1, t2), t_real.detach()))
    mm = [d[i] + (c[i] + b[i] * t_s) * t_s for i in range(3)]
    return dict(t=t_s, t_real=t_real, m=mm, d=d, b=b, c=c, one=one, use_b=use_b, interior=interior,
                resx=d2r, dot=dot)


def _g_quad(p, prm):
    """Surrogate whose parameter derivatives are the reference's quadratic-Bezier adjoint."""
    r1, r2 = prm[..., 9], prm[..., 10]
    Q = _quad_parts(p, prm, 2.0 / 3.0)
    t_s, t_real = Q['t'], Q['t_real']
    dist = torch.sqrt(Q['dot'](Q['m'], Q['m']) + 1e-8)
    # radius: r1 factor (t + 1) for the first three-root candidate (:1046), else 1 - t
    first = ~Q['one'] & ~Q['use_b'] & Q['interior']
    rr = r1.detach() + t_s * (r2.detach() - r1.detach())
    r1f = torch.where(first, t_real + 1.0, 1.0 - t_real).detach()
    radius = rr + (r1 * r1f - (r1 * r1f).detach()) + (r2 * t_real - (r2 * t_real).detach())
    return dist - radius


def _g_quad_xyz(p, prm):
    """Surrogate whose position derivative is the reference's quadratic-Bezier position adjoint:
    grad_ans_xyz_category_1/2 form 2m * (-I + outer(dt/dx, dm/dt)) with helper_math's row-vector
    product (v * M)_j = sum_i v_i M_ij (:1560-1566, :1584-1589), i.e. (2m . dt/dx) dm/dt rather
    than (2m . dm/dt) dt/dx; grad_ans_xyz_category_3 returns -d / |m| whichever end t is at."""
    r1, r2 = prm[..., 9].detach(), prm[..., 10].detach()
    Q = _quad_parts(p, prm, 1.0 / 3.0)
    m = [x.detach() for x in Q['m']]
    rsx = torch.rsqrt(Q['resx'] + 1e-8).detach()
    t_real = Q['t_real']
    mt = [(Q['c'][i] + 2.0 * Q['b'][i] * t_real).detach() for i in range(3)]
    # position moved along 2m by the projection of its displacement on dm/dt
    pd = [x.detach() for x in p]
    proj = sum(mt[j] * (p[j] - pd[j]) for j in range(3))
    Qm = _quad_parts([pd[i] + 2.0 * m[i] * proj for i in range(3)], prm, 1.0 / 3.0)
    d_pos = Q['dot'](m, Q['d'])  # d/dx = -m
    interior_path = rsx * (d_pos + 0.5 * (Qm['t'] - Qm['t'].detach())) - (r2 - r1) * Q['t']
    clamped_path = rsx * Q['dot']([x.detach() for x in Q['d']], Q['d'])
    return torch.where(Q['interior'], interior_path, clamped_path)


SURROGATES = {1: _g_cube, 3: _g_capped_torus, 9: _g_quad, 10: _g_cubic, 11: _g_catmull}


def _surrogate_stroke_grads(x, radius, viewdir, sp, cp, sdf_id, color_id, delta, laplace, isr, ga, gc):
    base, _, _, ss, ms = ostrokes.decode_sdf_id(sdf_id)
    nb = ostrokes.NUM_BASE_PARAMS[base]
    with torch.no_grad():
        fw = ostrokes.stroke_forward(x, radius, viewdir, sp, cp, sdf_id, color_id, delta, laplace, isr)
    spv = sp.clone().requires_grad_(True)
    cpv = cp.clone().requires_grad_(True)
    n_points = x.shape[0]
    p = ostrokes.inverse_transform(x[:, None, :], spv[None], sdf_id)
    p = [pi.expand(n_points, sp.shape[0]) for pi in p]
    prm = spv[None, :, :nb].expand(n_points, -1, -1)
    g = SURROGATES.get(base, lambda pu, pr: ostrokes.base_sdf(base, pu, pr))(p, prm)
    sdf = ste(fw['sdf'], g)
    # the transfer uses the forward's sdf scale: no derivative through the radius (H3)
    k = fw['sdf_scale'].detach()
    if laplace:
        alpha = ostrokes.laplace_cdf(-sdf * k)
    else:
        alpha = torch.clamp(-sdf * k + 0.5, 0.0, 0.9999)
    n_samples = n_points // viewdir.shape[0]
    dir_raw = [t[:, None].expand(n_points, sp.shape[0]) for t in viewdir.repeat_interleave(n_samples, 0).unbind(-1)]
    # H11: the colour field's adjoint has no position term
    col = ostrokes.color_forward(color_id, None, [pi.detach() for pi in p], dir_raw, None, cpv[None])
    loss = (ga * alpha).sum() + (gc * col).sum()
    gs, gcp = torch.autograd.grad(loss, [spv, cpv], allow_unused=True)
    return gs if gs is not None else torch.zeros_like(sp), gcp if gcp is not None else torch.zeros_like(cp)


CASES = [('ellipsoid', 'constant_rgb', False, False), ('obb', 'constant_rgb', False, False),
         ('cube', 'constant_rgb', True, False), ('cappedtorus', 'constant_rgb', False, False),
         ('ellipsoid', 'gradient_rgb', False, False), ('scapsule', 'gradient_rgb', True, False),
         ('ellipsoid', 'constant_rgb', False, True), ('sphere', 'constant_rgb', True, True),
         ('cubic_bezier', 'constant_rgb', False, False), ('catmull_rom', 'constant_rgb', True, False),
         ('quadratic_bezier', 'constant_rgb', False, False), ('quadratic_bezier', 'constant_rgb', True, False)]


@pytest.mark.parametrize('shape,color,laplace,isr', CASES)
def test_stroke_reference_adjoint_is_the_autograd_of_its_surrogate(shape, color, laplace, isr):
    from helpers import points_near, stroke_params
    if shape not in ostrokes.SHAPE_TYPES:
        pytest.skip(f'{shape} is not a registered shape')
    sdf_id = ostrokes.shape_sdf_id(shape)
    color_id = ostrokes.COLOR_TYPES[color] if isinstance(ostrokes.COLOR_TYPES, dict) else \
        list(ostrokes.COLOR_TYPES).index(color)
    sp, cp = stroke_params(shape, color, 6, seed=3)
    n_rays, n_samples = 16, 24
    x, vd = points_near(shape, sp, n_rays, n_samples, seed=5)
    x = x.reshape(-1, 3).double()
    sp, cp, vd = sp.double(), cp.double(), vd.double()
    radius = torch.full((x.shape[0],), 0.02, dtype=D64)
    delta = 1.0
    g = torch.Generator().manual_seed(9)
    ga = torch.randn(x.shape[0], sp.shape[0], generator=g, dtype=D64)
    gc = torch.randn(x.shape[0], sp.shape[0], 3, generator=g, dtype=D64)
    ref_s, ref_c, _ = _reference(ga, gc, x, radius, vd, sp, cp, sdf_id, color_id, delta, laplace, isr)
    sur_s, sur_c = _surrogate_stroke_grads(x, radius, vd, sp, cp, sdf_id, color_id, delta, laplace, isr, ga, gc)
    ok = torch.isfinite(ref_s).all(-1)
    assert ok.float().mean() > 0.9
    assert rel(ref_s[ok], sur_s[ok]) < 1e-9, (shape, rel(ref_s[ok], sur_s[ok]))
    assert rel(ref_c, sur_c) < 1e-9


def _reference(ga, gc, x, radius, vd, sp, cp, sdf_id, color_id, delta, laplace, isr):
    fw = ostrokes.stroke_forward(x, radius, vd, sp, cp, sdf_id, color_id, delta, laplace, isr)
    return ostrokes.stroke_backward_reference(ga, gc, None, x, radius, vd, fw['alpha'], sp, cp, sdf_id, color_id,
                                              delta, laplace, isr, need_grad_x=False)


@pytest.mark.parametrize('laplace', [False, True])
def test_quad_position_adjoint_is_the_autograd_of_its_surrogate(laplace):
    """dL/dx of the quadratic Bezier's reference adjoint (grad_ans_xyz_category_1/2/3)."""
    from helpers import points_near, stroke_params
    shape = 'quadratic_bezier'
    sdf_id = ostrokes.shape_sdf_id(shape)
    sp, cp = stroke_params(shape, 'constant_rgb', 6, seed=4)
    x, vd = points_near(shape, sp, 16, 24, seed=6)
    x = x.reshape(-1, 3).double()
    sp, cp, vd = sp.double(), cp.double(), vd.double()
    radius = torch.full((x.shape[0],), 0.02, dtype=D64)
    g = torch.Generator().manual_seed(10)
    ga = torch.randn(x.shape[0], sp.shape[0], generator=g, dtype=D64)
    gc = torch.zeros(x.shape[0], sp.shape[0], 3, dtype=D64)
    fw = ostrokes.stroke_forward(x, radius, vd, sp, cp, sdf_id, 0, 1.0, laplace, False)
    _, _, ref_x = ostrokes.stroke_backward_reference(ga, gc, None, x, radius, vd, fw['alpha'], sp, cp, sdf_id, 0,
                                                     1.0, laplace, False, need_grad_x=True)
    xv = x.clone().requires_grad_(True)
    n = x.shape[0]
    p = [xv[:, None, i].expand(n, sp.shape[0]) for i in range(3)]
    prm = sp[None].expand(n, -1, -1)
    sdf = ste(fw['sdf'], _g_quad_xyz(p, prm))
    k = fw['sdf_scale'].detach()
    alpha = ostrokes.laplace_cdf(-sdf * k) if laplace else torch.clamp(-sdf * k + 0.5, 0.0, 0.9999)
    gx, = torch.autograd.grad((ga * alpha).sum(), xv)
    ok = torch.isfinite(ref_x).all(-1)
    assert ok.float().mean() > 0.9
    assert float(gx[ok].abs().max()) > 0  # the points do sit in the tubes' transfer band
    assert rel(ref_x[ok], gx[ok]) < 1e-9, rel(ref_x[ok], gx[ok])

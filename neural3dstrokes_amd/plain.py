"""Plain per-op PyTorch formulation of the stroke-field hot path (the comparison baseline).

This is the reference's math written as ordinary torch tensor operations,
the way it runs without custom kernels: every stroke is evaluated at every
sample (no culling), each type of stroke is one batch of elementwise ops
(``source/strokelib/src/strokes_sdf.h`` formulas), compositing is the
closed form of the ordered "over" operator (``compositing.cu:9-59``) and the
ray march is ``source/utils/render.py:29-88``; gradients come from
torch.autograd.  Rays are processed in chunks (the unchunked per-(sample,
stroke) intermediates at 65536 rays x 128 samples x 500 strokes are tens of GB
each).  Used only as the timing / accuracy comparator of ``bench.py``; the
operators of this package never fall back to it.

Supported palette: sphere, cube, round cube, capsule, octahedron and
tetrahedron bases with any transform flags, quadratic and cubic Bezier tubes;
constant RGB colour.
"""
import torch

N_BASE = [0, 0, 1, 2, 1, 2, 1, 0, 0, 11, 14, 14]


def _nsp(sid):
    return N_BASE[sid >> 4] + 3 * (sid & 1) + 3 * ((sid >> 1) & 1) + (1 if sid & 4 else (3 if sid & 8 else 0))


def _safe_sqrt(v):
    pos = v > 0
    return torch.where(pos, torch.sqrt(torch.where(pos, v, torch.ones_like(v))), torch.zeros_like(v))


def _rot(angle):
    sx, sy, sz = torch.sin(angle[:, 0]), torch.sin(angle[:, 1]), torch.sin(angle[:, 2])
    cx, cy, cz = torch.cos(angle[:, 0]), torch.cos(angle[:, 1]), torch.cos(angle[:, 2])
    return torch.stack([torch.stack([cy * cz, sx * sy * cz - cx * sz, sx * sz + cx * sy * cz], -1),
                        torch.stack([cy * sz, cx * cz + sx * sy * sz, cx * sy * sz - sx * cz], -1),
                        torch.stack([-sy, sx * cy, cx * cy], -1)], -2)  # [S,3,3]


def _unit_pos(x, sp, sid):
    """x [N,3] -> p [N,S,3] = R^T (x - t) / s."""
    end = sp.shape[1]
    p = x[:, None, :].expand(x.shape[0], sp.shape[0], 3)
    trans = angle = scale = None
    if sid & 1:
        trans = sp[:, end - 3:end]; end -= 3
    if sid & 2:
        angle = sp[:, end - 3:end]; end -= 3
    if sid & 4:
        scale = sp[:, end - 1:end]; end -= 1
    elif sid & 8:
        scale = sp[:, end - 3:end]; end -= 3
    if trans is not None:
        p = p - trans[None]
    if angle is not None:
        p = torch.einsum('sji,nsj->nsi', _rot(angle), p)
    if scale is not None:
        p = p / scale[None]
    return p


def _sdf(base, p, prm):
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    if base == 0:
        return torch.sqrt(x * x + y * y + z * z + 1e-8) - 1.0
    if base in (1, 2):
        q = torch.abs(p) - 1.0
        qp = torch.clamp_min(q, 0.0)
        inner = torch.clamp_max(torch.amax(q, -1), 0.0)
        if base == 1:
            return _safe_sqrt((qp * qp).sum(-1)) + inner
        return torch.sqrt((qp * qp).sum(-1) + 1e-8) + inner - prm[None, :, 0]
    if base == 4:
        h = prm[None, :, 0]
        yy = y - torch.maximum(-h, torch.minimum(y, h))
        return torch.sqrt(x * x + yy * yy + z * z + 1e-8) - 1.0
    if base == 7:
        return (torch.abs(x) + torch.abs(y) + torch.abs(z) - 1.0) * 0.57735027
    if base == 8:
        return (torch.maximum(torch.abs(x + y) - z, torch.abs(x - y) + z) - 1.0) * 0.57735027
    if base == 9:
        return _quad(p, prm)
    if base == 10:
        return _cubic(p, prm)
    raise NotImplementedError(f'plain formulation does not cover base {base}')


def _quad(pos, prm):
    A, B, C = prm[None, :, 0:3], prm[None, :, 3:6], prm[None, :, 6:9]
    a = B - A
    b = A - 2.0 * B + C
    c = 2.0 * a
    d = A - pos
    dot = lambda u, v: (u * v).sum(-1)
    kk = 1.0 / dot(b, b)
    kx = kk * dot(a, b)
    ky = kk * (2.0 * dot(a, a) + dot(d, b)) / 3.0
    kz = kk * dot(d, a)
    p = ky - kx * kx
    q = kx * (2.0 * kx * kx - 3.0 * ky) + kz
    h = q * q + 4.0 * p * p * p
    one = h >= 0.0
    hh = torch.sqrt(torch.where(one, h, torch.zeros_like(h)) + 1e-8)
    cb = lambda v: torch.where(v < 0, -1.0, 1.0) * torch.where(v != 0, torch.abs(torch.where(v != 0, v, torch.ones_like(v))) ** (1.0 / 3.0), torch.zeros_like(v))
    t1 = torch.clamp(cb((hh - q) * 0.5) + cb(-(hh + q) * 0.5) - kx, 0.0, 1.0)
    pn = torch.where(one, -torch.ones_like(p), p)
    zz = torch.sqrt(-pn + 1e-8)
    arg = torch.clamp(torch.where(one, torch.zeros_like(q), q) / (pn * zz * 2.0 + 1e-8), -1.0, 1.0)
    v = torch.acos(arg) / 3.0
    m, n = torch.cos(v), torch.sin(v) * 1.732050808
    ta = torch.clamp((m + n) * zz - kx, 0.0, 1.0)
    tb = torch.clamp((-n - m) * zz - kx, 0.0, 1.0)
    d2 = lambda t: dot(d + (c + b * t[..., None]) * t[..., None], d + (c + b * t[..., None]) * t[..., None])
    da, db = d2(ta), d2(tb)
    t = torch.where(one, t1, torch.where(db < da, tb, ta))
    dd = torch.where(one, d2(t1), torch.where(db < da, db, da))
    r1, r2 = prm[None, :, 9], prm[None, :, 10]
    return torch.sqrt(dd + 1e-8) - (r1 + t * (r2 - r1))


def _bez(t, A, B, C, D):
    u = 1.0 - t
    t = t[..., None]
    u = u[..., None]
    return (u * u * u) * A + (3.0 * u * u * t) * B + (3.0 * u * t * t) * C + (t * t * t) * D


def _cubic(pos, prm):
    A, B, C, D = (prm[None, :, 3 * i:3 * i + 3] for i in range(4))
    like = pos[..., 0]
    dmin = torch.full_like(like, 1e9)
    tmin = torch.zeros_like(like)
    s0 = _bez(torch.zeros_like(like), A, B, C, D)
    for i in range(1, 11):
        ts, te = (i - 1) * 0.1, i * 0.1
        s1 = _bez(torch.full_like(like, te), A, B, C, D)
        ab = s1 - s0
        s = torch.clamp(((pos - s0) * ab).sum(-1) / (ab * ab).sum(-1), 0.0, 1.0)
        dist = _safe_sqrt(((pos - (s0 + s[..., None] * ab)) ** 2).sum(-1))
        better = dist < dmin
        dmin = torch.where(better, dist, dmin)
        tmin = torch.where(better, ts * (1.0 - s) + te * s, tmin)
        s0 = s1
    P = _bez(tmin, A, B, C, D)
    r1, r2 = prm[None, :, 12], prm[None, :, 13]
    return _safe_sqrt(((P - pos) ** 2).sum(-1)) - (r1 + tmin * (r2 - r1))


def field_chunk(origins, dirs, radii, tdist, groups, shape_params, color_params, density_params, sdf_delta=1.0,
                laplace=False, cone=False, bg=1.0):
    """rgb [R,3], acc [R] for one chunk of rays (autograd through everything)."""
    R, n1 = tdist.shape
    t = 0.5 * (tdist[:, :-1] + tdist[:, 1:])
    x = (dirs[:, None, :] * t[..., None] + origins[:, None, :]).reshape(-1, 3)
    radius = (radii[:, None] * t if cone else radii[:, None].expand_as(t)).reshape(-1, 1)
    S = shape_params.shape[0]
    alpha = torch.zeros(x.shape[0], S, dtype=x.dtype, device=x.device)
    for sid, idx in groups:
        sp = shape_params[idx, :_nsp(sid)]
        p = _unit_pos(x, sp, sid)
        sdf = _sdf(sid >> 4, p, sp[:, :N_BASE[sid >> 4]])
        sc = (2.0 if laplace else 0.5) / (sdf_delta * radius)
        if laplace:
            s = -sdf * sc
            a = torch.where(s <= 0, 0.5 * torch.exp(torch.clamp_max(s, 0.0)), 1.0 - 0.5 * torch.exp(-torch.clamp_min(s, 0.0)))
        else:
            a = torch.clamp(-sdf * sc + 0.5, 0.0, 0.9999)
        alpha = alpha.index_copy(1, idx, a)
    # ordered over: stroke S-1 in front
    af = torch.flip(alpha, [1])
    T_ex = torch.cumprod(torch.cat([torch.ones_like(af[:, :1]), 1.0 - af[:, :-1]], 1), 1)
    w = af * T_ex
    dens = (w * torch.flip(density_params, [0])[None]).sum(1)
    col = w @ torch.flip(color_params, [0])
    Tn = T_ex[:, -1] * (1.0 - af[:, -1])
    col = torch.clamp(col / (1.00000095367431640625 - Tn)[:, None], 0.0, 1.0)
    dens = dens.reshape(R, -1)
    col = col.reshape(R, -1, 3)
    delta = (tdist[:, 1:] - tdist[:, :-1]) * torch.norm(dirs, dim=-1, keepdim=True)
    sd = dens * delta
    trans = torch.exp(-torch.cat([torch.zeros_like(sd[:, :1]), torch.cumsum(sd[:, :-1], -1)], -1))
    wr = (1.0 - torch.exp(-sd)) * trans
    acc = wr.sum(-1)
    rgb = (wr[..., None] * col).sum(-2) + (1.0 - acc[:, None]).clamp_min(0.0) * bg
    return rgb, acc


def fwd_bwd(origins, dirs, radii, tdist, stroke_ids, shape_params, color_params, density_params, g_rgb, g_acc,
            chunk=512, **kw):
    """One chunked forward + backward; returns (rgb, acc, grad_shape, grad_color, grad_density)."""
    ids = stroke_ids.tolist()
    dev = shape_params.device
    groups = [(sid, torch.tensor([k for k in range(len(ids)) if ids[k] == sid], device=dev)) for sid in sorted(set(ids))]
    sp = shape_params.detach().requires_grad_(True)
    cp = color_params.detach().requires_grad_(True)
    dp = density_params.detach().requires_grad_(True)
    rgbs, accs = [], []
    for r0 in range(0, tdist.shape[0], chunk):
        sl = slice(r0, r0 + chunk)
        rgb, acc = field_chunk(origins[sl], dirs[sl], radii[sl], tdist[sl], groups, sp, cp, dp, **kw)
        ((rgb * g_rgb[sl]).sum() + (acc * g_acc[sl]).sum()).backward()
        rgbs.append(rgb.detach())
        accs.append(acc.detach())
    return torch.cat(rgbs), torch.cat(accs), sp.grad, cp.grad, dp.grad

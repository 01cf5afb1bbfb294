"""Seeded synthetic stroke scenes and camera rays for the benchmark workloads.

Nothing here is loaded from disk: every tensor comes from one
``torch.Generator`` seed so a workload is reproducible on any machine.

Stroke parameter layout follows the reference registry
(``source/strokelib/strokes.py:40-106``): ``[base params..., scale (1 or 3),
euler xyz (3), translation (3)]``; Bezier curves carry only base params.
Rays follow ``source/models.py:116-191``: stratified single-jitter intervals
(``source/utils/stepfun.py:280-317``) on [near, far] with the identity ray warp.
"""
import math

import torch

BASE_ID = {'sphere': 0, 'cube': 1, 'round_cube': 2, 'capped_torus': 3, 'capsule': 4, 'line': 5,
           'triprism': 6, 'octahedron': 7, 'tetrahedron': 8, 'quadratic_bezier': 9, 'cubic_bezier': 10,
           'catmull_rom': 11}
N_BASE = [0, 0, 1, 2, 1, 2, 1, 0, 0, 11, 14, 14]
CURVES = (9, 10, 11)


def sdf_id_of(base, translation=True, rotation=True, singlescale=False, multiscale=True):
    """``base << 4 | T | R<<1 | SS<<2 | MS<<3`` (strokes.py:95-106)."""
    if base in CURVES:
        return base << 4
    return (base << 4) | (1 if translation else 0) | (2 if rotation else 0) | (4 if singlescale else 0) | \
        (8 if multiscale else 0)


def n_shape_params(sdf_id):
    base = sdf_id >> 4
    return N_BASE[base] + 3 * (sdf_id & 1) + 3 * ((sdf_id >> 1) & 1) + (1 if sdf_id & 4 else (3 if sdf_id & 8 else 0))


def _u(gen, *shape, lo=0.0, hi=1.0):
    return torch.rand(*shape, generator=gen, dtype=torch.float64) * (hi - lo) + lo


def _logu(gen, *shape, lo, hi):
    return torch.exp(_u(gen, *shape, lo=math.log(lo), hi=math.log(hi)))


def random_euler(n, gen):
    """Uniform random rotations (Shoemake quaternions) as euler angles of R = Rz Ry Rx."""
    u1, u2, u3 = _u(gen, n), _u(gen, n), _u(gen, n)
    qx = torch.sqrt(1 - u1) * torch.sin(2 * math.pi * u2)
    qy = torch.sqrt(1 - u1) * torch.cos(2 * math.pi * u2)
    qz = torch.sqrt(u1) * torch.sin(2 * math.pi * u3)
    qw = torch.sqrt(u1) * torch.cos(2 * math.pi * u3)
    r00 = 1 - 2 * (qy * qy + qz * qz)
    r10 = 2 * (qx * qy + qz * qw)
    r20 = 2 * (qx * qz - qy * qw)
    r21 = 2 * (qy * qz + qx * qw)
    r22 = 1 - 2 * (qx * qx + qy * qy)
    y = -torch.asin(torch.clamp(r20, -1.0, 1.0))
    x = torch.atan2(r21, r22)
    z = torch.atan2(r10, r00)
    return torch.stack([x, y, z], -1)


def make_strokes(types, gen, scale_range=(0.05, 0.3), radius_range=(0.01, 0.05), density_range=(1.0, 100.0),
                 color_id=0, singlescale=False):
    """Random strokes of the given base types (list of base ids, in draw order).

    Returns dict(stroke_ids int32 [S], shape_params f32 [S,Pmax], color_params f32 [S,Q],
    density_params f32 [S]).
    """
    S = len(types)
    ids = [sdf_id_of(b, multiscale=not singlescale, singlescale=singlescale) for b in types]
    pmax = max(n_shape_params(i) for i in ids)
    sp = torch.zeros(S, pmax, dtype=torch.float64)
    for k, (b, sid) in enumerate(zip(types, ids)):
        row = []
        if b == 9:
            row = list(_u(gen, 9, lo=-1.0, hi=1.0)) + list(_u(gen, 2, lo=radius_range[0], hi=radius_range[1]))
        elif b in (10, 11):
            row = list(_u(gen, 12, lo=-1.0, hi=1.0)) + list(_u(gen, 2, lo=radius_range[0], hi=radius_range[1]))
        else:
            if b == 2:
                row += list(0.8 * _u(gen, 1))
            elif b == 3:
                row += [_u(gen, 1)[0] * math.pi, _u(gen, 1)[0]]
            elif b == 4:
                row += list(_u(gen, 1) + 0.25)
            elif b == 5:
                row += [_u(gen, 1)[0] + 0.25, _u(gen, 1)[0] - 0.5]
            elif b == 6:
                row += list(_u(gen, 1))
            if singlescale:
                row += list(_logu(gen, 1, lo=scale_range[0], hi=scale_range[1]))
            else:
                row += list(_logu(gen, 3, lo=scale_range[0], hi=scale_range[1]))
            row += list(random_euler(1, gen)[0])
            row += list(_u(gen, 3, lo=-1.0, hi=1.0))
        sp[k, :len(row)] = torch.tensor([float(v) for v in row], dtype=torch.float64)
    if color_id == 0:
        cp = _u(gen, S, 3)
    elif color_id == 1:
        cp = torch.cat([_u(gen, S, 6, lo=-0.5, hi=0.5), _u(gen, S, 6)], -1)
    else:
        cp = torch.randn(S, 12 if color_id == 2 else 27, generator=gen, dtype=torch.float64)
    dens = _logu(gen, S, lo=density_range[0], hi=density_range[1])
    return dict(stroke_ids=torch.tensor(ids, dtype=torch.int32), shape_params=sp.float(),
                color_params=cp.float(), density_params=dens.float())


def stratified_tdist(n_rays, n_samples, near, far, gen):
    """sample_intervals(rand=True, single_jitter=True) on one uniform bin, stepfun.py:243-317."""
    eps = torch.finfo(torch.float32).eps
    u_max = eps + (1 - eps) / n_samples
    max_jitter = (1 - u_max) / (n_samples - 1) - eps
    u = torch.linspace(0, 1 - u_max, n_samples, dtype=torch.float64)[None, :] + \
        _u(gen, n_rays, 1) * max_jitter
    mid = (u[:, 1:] + u[:, :-1]) / 2
    first = torch.clamp_min(2 * u[:, :1] - mid[:, :1], 0.0)
    last = torch.clamp_max(2 * u[:, -1:] - mid[:, -1:], 1.0)
    s = torch.cat([first, mid, last], -1)
    return (s * far + (1 - s) * near).float()


def make_rays(n_rays, n_samples, gen, n_cams=64, cam_radius=4.0, fov_deg=40.0, near=2.0, far=6.0,
              radius=0.01):
    """Cameras uniform on a sphere looking at the origin; unit directions uniform in a cone.

    Returns dict(origins, directions, viewdirs [R,3], radii [R], tdist [R, n+1]) as float32.
    """
    z = _u(gen, n_cams, lo=-1.0, hi=1.0)
    phi = _u(gen, n_cams, lo=0.0, hi=2 * math.pi)
    rr = torch.sqrt(1 - z * z)
    centres = torch.stack([rr * torch.cos(phi), rr * torch.sin(phi), z], -1) * cam_radius
    fwd = -centres / cam_radius
    helper = torch.where(torch.abs(fwd[:, 2:3]) < 0.9, torch.tensor([[0.0, 0.0, 1.0]], dtype=torch.float64),
                         torch.tensor([[1.0, 0.0, 0.0]], dtype=torch.float64))
    right = torch.linalg.cross(fwd, helper)
    right = right / right.norm(dim=-1, keepdim=True)
    up = torch.linalg.cross(right, fwd)
    cam = torch.arange(n_rays) * n_cams // n_rays
    half = math.radians(fov_deg) / 2
    cos_t = _u(gen, n_rays, lo=math.cos(half), hi=1.0)
    sin_t = torch.sqrt(1 - cos_t * cos_t)
    psi = _u(gen, n_rays, lo=0.0, hi=2 * math.pi)
    d = fwd[cam] * cos_t[:, None] + (right[cam] * torch.cos(psi)[:, None] + up[cam] * torch.sin(psi)[:, None]) * \
        sin_t[:, None]
    d = d / d.norm(dim=-1, keepdim=True)
    tdist = stratified_tdist(n_rays, n_samples, near, far, gen)
    return dict(origins=centres[cam].float(), directions=d.float(), viewdirs=d.float(),
                radii=torch.full((n_rays,), radius, dtype=torch.float32), tdist=tdist)


SOLIDS_ELLIPSOID_OBB_CAPSULE = [0, 1, 4]          # ellipsoid, obb, scapsule
SOLIDS_MIXED = [0, 1, 2, 7, 8]    # sphere, cube, round-cube, octahedron, tetrahedron (per-axis scale)


def palette(n_strokes, gen, kind):
    """Seeded type assignment (a permutation of the type multiset) in draw order."""
    if kind == 'solids':
        types = [SOLIDS_ELLIPSOID_OBB_CAPSULE[int(i)] for i in torch.randint(0, 3, (n_strokes,), generator=gen)]
        return types
    n_cubic = int(round(0.4 * n_strokes))
    n_quad = int(round(0.2 * n_strokes))
    n_solid = n_strokes - n_cubic - n_quad
    solids = [SOLIDS_MIXED[int(i)] for i in torch.randint(0, 5, (n_solid,), generator=gen)]
    types = [10] * n_cubic + [9] * n_quad + solids
    perm = torch.randperm(n_strokes, generator=gen)
    return [types[int(i)] for i in perm]


WORKLOADS = {
    'solids50_4096x64': dict(n_strokes=50, n_rays=4096, n_samples=64, palette='solids'),
    'palette500_65536x128': dict(n_strokes=500, n_rays=65536, n_samples=128, palette='mixed'),
    'palette1000_640000x192': dict(n_strokes=1000, n_rays=640000, n_samples=192, palette='mixed'),
}


def make_workload(name, seed=0, n_rays=None, n_samples=None, n_strokes=None):
    """Stroke scene + rays for a named workload, optionally resized (for small parity cases)."""
    spec = dict(WORKLOADS[name])
    if n_rays is not None:
        spec['n_rays'] = n_rays
    if n_samples is not None:
        spec['n_samples'] = n_samples
    if n_strokes is not None:
        spec['n_strokes'] = n_strokes
    gen = torch.Generator().manual_seed(seed)
    types = palette(spec['n_strokes'], gen, spec['palette'])
    scene = make_strokes(types, gen)
    rays = make_rays(spec['n_rays'], spec['n_samples'], gen)
    scene.update(rays)
    scene['types'] = types
    return scene


def make_multiscene(n_scenes=32, seed=0, stroke_range=(20, 1000), ray_range=(1024, 65536), n_samples=128,
                    shape_width=14, color_id=0):
    """Independent scenes with log-uniform stroke and ray counts, each drawn like the
    500-stroke mixed palette (40% cubic, 20% quadratic Bezier, 40% solids) with its own
    cameras; shape params zero-padded to ``shape_width`` so the scenes pack together.

    Returns (list of per-scene dicts, ray counts, stroke counts).
    """
    gen = torch.Generator().manual_seed(seed)
    s_counts = [int(round(v)) for v in _logu(gen, n_scenes, lo=stroke_range[0], hi=stroke_range[1]).tolist()]
    r_counts = [int(round(v)) for v in _logu(gen, n_scenes, lo=ray_range[0], hi=ray_range[1]).tolist()]
    scenes = []
    for S, R in zip(s_counts, r_counts):
        types = palette(S, gen, 'mixed')
        sc = make_strokes(types, gen, color_id=color_id)
        sp = sc['shape_params']
        if sp.shape[1] < shape_width:
            sc['shape_params'] = torch.cat([sp, sp.new_zeros(S, shape_width - sp.shape[1])], 1)
        sc.update(make_rays(R, n_samples, gen))
        sc['types'] = types
        scenes.append(sc)
    return scenes, r_counts, s_counts

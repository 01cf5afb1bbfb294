"""Oracle restatement of the multiresolution hash-grid encoder (test infrastructure only).

Follows ``source/gridencoder/grid.py:96-170`` (level offsets and resolutions)
and ``source/gridencoder/src/gridencoder.cu``: ``fast_hash`` (:50-63),
``get_grid_index`` (:66-84), ``kernel_grid`` (:87-245, forward + dy/dx),
``kernel_grid_backward`` (:248-340), ``kernel_input_backward`` (:343-369) and
``kernel_grad_tv`` (:506-610).  Integer index arithmetic is done in uint32
with wrap-around exactly as the CUDA code; the interpolation runs in float64
(the oracle) or float32 (emulation; ``x*scale+0.5`` is a single fused
multiply-add as compiled by nvcc, reproduced by rounding the exact float64
product-sum once).

Pinned by ``tests/test_oracle_grid.py``: offsets/resolutions against the
closed-form table for the 16-level 2^19 configuration; a query exactly on a
lattice vertex returns that table row; trilinear weights sum to one; the dense
index equals ``x + y·R + z·R²``; the hash equals the Instant-NGP prime-XOR hash
computed independently with Python integers; backward is the adjoint of the
forward (<g, J e> = <Jᵀ g, e>) and matches finite differences for inputs.
"""
import numpy as np

PRIMES = [1, 2654435761, 805459861, 3674653429, 2097192037, 1434869437, 2165219737]


def grid_layout(input_dim=3, num_levels=16, level_dim=2, per_level_scale=2.0, base_resolution=16,
                log2_hashmap_size=19, desired_resolution=None, align_corners=False):
    """GridEncoder.__init__ bookkeeping, grid.py:104-144. Returns (offsets int32, resolutions, per_level_scale)."""
    if desired_resolution is not None:
        per_level_scale = np.exp2(np.log2(desired_resolution / base_resolution) / (num_levels - 1))
    max_params = 2 ** log2_hashmap_size
    offsets, resolutions = [], []
    offset = 0
    for i in range(num_levels):
        res = int(np.ceil(base_resolution * per_level_scale ** i))
        res = res if align_corners else res + 1
        n = min(max_params, res ** input_dim)
        n = int(np.ceil(n / 8) * 8)
        resolutions.append(res)
        offsets.append(offset)
        offset += n
    offsets.append(offset)
    return np.array(offsets, dtype=np.int32), resolutions, per_level_scale


def level_scale_f32(level, S, H):
    """``exp2f(level * S) * H - 1.0f`` in float32, gridencoder.cu:138."""
    ls = np.float32(np.float32(level) * np.float32(S))
    e = np.float32(np.exp2(np.float64(ls)))
    return np.float32(np.float32(e * np.float32(H)) - np.float32(1.0))


def fast_hash(pos_grid):
    """gridencoder.cu:50-63: XOR of coordinate * prime, uint32 wrap."""
    res = np.zeros(pos_grid.shape[:-1], dtype=np.uint32)
    for i in range(pos_grid.shape[-1]):
        res ^= (pos_grid[..., i].astype(np.uint64) * np.uint64(PRIMES[i]) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return res


def grid_index(gridtype, align_corners, hashmap_size, resolution, pos_grid):
    """gridencoder.cu:66-84, returns the row index (times C is applied by the caller)."""
    D = pos_grid.shape[-1]
    stride = 1
    index = np.zeros(pos_grid.shape[:-1], dtype=np.uint64)
    for d in range(D):
        if stride > hashmap_size:
            break
        index = (index + pos_grid[..., d].astype(np.uint64) * np.uint64(stride)) & np.uint64(0xFFFFFFFF)
        # uint32 product, as in the CUDA code: for resolution+1 = 65537 the stride
        # wraps to 131073 <= hashmap_size and the level keeps a (wrapped) dense index
        stride = (stride * (resolution if align_corners else resolution + 1)) & 0xFFFFFFFF
    if gridtype == 0 and stride > hashmap_size:
        index = fast_hash(pos_grid).astype(np.uint64)
    return (index % np.uint64(hashmap_size)).astype(np.int64)


def _level_coords(inputs, level, offsets, S, H, align_corners, interp, f32):
    scale = level_scale_f32(level, S, H)
    resolution = int(np.ceil(np.float64(scale))) + 1
    x = inputs.astype(np.float64)
    if f32:
        pos = (x * np.float64(scale) + (0.0 if align_corners else 0.5)).astype(np.float32).astype(np.float64)
    else:
        pos = x * np.float64(scale) + (0.0 if align_corners else 0.5)
    pg = np.floor(pos).astype(np.int64)
    frac = pos - pg
    if f32:
        frac = frac.astype(np.float32).astype(np.float64)
    if interp == 1:
        deriv = 6 * frac * (1.0 - frac)
        frac = frac * frac * (3.0 - 2.0 * frac)
    else:
        deriv = np.ones_like(frac)
    return scale, resolution, pg.astype(np.uint32), frac, deriv


def grid_forward(inputs, embeddings, offsets, per_level_scale, base_resolution, gridtype=0,
                 align_corners=False, interp=0, calc_dy_dx=False, f32=False):
    """kernel_grid + permute, grid.py:47-57.  inputs [B,D] float, embeddings [n,C].

    Returns outputs [B, L*C] (float64 unless f32) and dy_dx [B, L*D*C] or None.
    """
    B, D = inputs.shape
    L = len(offsets) - 1
    C = embeddings.shape[1]
    S = np.float32(np.log2(per_level_scale))
    H = base_resolution
    emb = embeddings.astype(np.float64)
    oob = np.any((inputs < 0) | (inputs > 1), axis=1)
    out = np.zeros((B, L, C))
    dydx = np.zeros((B, L, D, C)) if calc_dy_dx else None
    for l in range(L):
        hs = int(offsets[l + 1] - offsets[l])
        scale, res, pg, frac, deriv = _level_coords(inputs, l, offsets, S, H, align_corners, interp, f32)
        table = emb[offsets[l]:offsets[l + 1]]
        acc = np.zeros((B, C))
        for corner in range(1 << D):
            w = np.ones(B)
            loc = pg.copy()
            for d in range(D):
                if corner & (1 << d):
                    w = w * frac[:, d]
                    loc[:, d] = loc[:, d] + np.uint32(1)
                else:
                    w = w * (1.0 - frac[:, d])
            idx = grid_index(gridtype, align_corners, hs, res, loc)
            acc += w[:, None] * table[idx]
        out[:, l, :] = acc
        if calc_dy_dx:
            for gd in range(D):
                rg = np.zeros((B, C))
                for corner in range(1 << (D - 1)):
                    w = np.full(B, np.float64(scale))
                    loc = pg.copy()
                    for nd in range(D - 1):
                        d = nd + 1 if nd >= gd else nd
                        if corner & (1 << nd):
                            w = w * frac[:, d]
                            loc[:, d] = loc[:, d] + np.uint32(1)
                        else:
                            w = w * (1.0 - frac[:, d])
                    loc[:, gd] = pg[:, gd]
                    il = grid_index(gridtype, align_corners, hs, res, loc)
                    loc[:, gd] = pg[:, gd] + np.uint32(1)
                    ir = grid_index(gridtype, align_corners, hs, res, loc)
                    rg += (w * deriv[:, gd])[:, None] * (table[ir] - table[il])
                dydx[:, l, gd, :] = rg
    out[oob] = 0.0
    if calc_dy_dx:
        dydx[oob] = 0.0
        dydx = dydx.reshape(B, L * D * C)
    return out.reshape(B, L * C), dydx


def grid_backward(grad, inputs, embeddings, offsets, per_level_scale, base_resolution, gridtype=0,
                  align_corners=False, interp=0, dy_dx=None, f32=False):
    """kernel_grid_backward (+ kernel_input_backward), grid.py:65-89.

    grad [B, L*C].  Returns (grad_embeddings [n,C] float64, grad_inputs [B,D] or None).
    """
    B, D = inputs.shape
    L = len(offsets) - 1
    C = embeddings.shape[1]
    S = np.float32(np.log2(per_level_scale))
    H = base_resolution
    g = grad.reshape(B, L, C).astype(np.float64)
    ok = ~np.any((inputs < 0) | (inputs > 1), axis=1)
    gemb = np.zeros((embeddings.shape[0], C))
    for l in range(L):
        hs = int(offsets[l + 1] - offsets[l])
        _, res, pg, frac, _ = _level_coords(inputs, l, offsets, S, H, align_corners, interp, f32)
        for corner in range(1 << D):
            w = np.ones(B)
            loc = pg.copy()
            for d in range(D):
                if corner & (1 << d):
                    w = w * frac[:, d]
                    loc[:, d] = loc[:, d] + np.uint32(1)
                else:
                    w = w * (1.0 - frac[:, d])
            idx = grid_index(gridtype, align_corners, hs, res, loc) + int(offsets[l])
            contrib = (w[:, None] * g[:, l, :]) * ok[:, None]
            for ch in range(C):
                np.add.at(gemb[:, ch], idx, contrib[:, ch])
    gin = None
    if dy_dx is not None:
        dd = dy_dx.reshape(B, L, D, C).astype(np.float64)
        gin = np.einsum('blc,bldc->bd', g, dd)
    return gemb, gin


def grad_total_variation(inputs, embeddings, grad, offsets, weight, per_level_scale, base_resolution,
                         gridtype=0, align_corners=False):
    """kernel_grad_tv, gridencoder.cu:506-610: adds the TV gradient into ``grad`` (returned copy)."""
    B, D = inputs.shape
    L = len(offsets) - 1
    C = embeddings.shape[1]
    S = np.float32(np.log2(per_level_scale))
    H = base_resolution
    emb = embeddings.astype(np.float64)
    out = grad.astype(np.float64).copy()
    ok = ~np.any((inputs < 0) | (inputs > 1), axis=1)
    w = weight / (2 * D)
    for l in range(L):
        hs = int(offsets[l + 1] - offsets[l])
        _, res, pg, _, _ = _level_coords(inputs, l, offsets, S, H, align_corners, 0, False)
        table = emb[offsets[l]:offsets[l + 1]]
        idx = grid_index(gridtype, align_corners, hs, res, pg)
        results = np.zeros((B, C))
        idelta = np.zeros((B, C))
        for d in range(D):
            cur = pg[:, d].copy()
            right = cur < res
            loc = pg.copy()
            loc[:, d] = cur + np.uint32(1)
            ir = grid_index(gridtype, align_corners, hs, res, loc)
            gv = table[idx] - table[ir]
            results += gv * right[:, None]
            idelta += gv * gv * right[:, None]
            left = cur > 0
            loc = pg.copy()
            loc[:, d] = cur - np.uint32(1)
            il = grid_index(gridtype, align_corners, hs, res, loc)
            gv = table[idx] - table[il]
            results += gv * left[:, None]
            idelta += gv * gv * left[:, None]
        val = w * results / np.sqrt(idelta + 1e-9)
        val = val * ok[:, None]
        for ch in range(C):
            np.add.at(out[:, ch], idx + int(offsets[l]), val[:, ch])
    return out

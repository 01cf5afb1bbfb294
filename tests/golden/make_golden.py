"""Writes tests/golden/field_small.npz: seeded float64 fixtures of the whole stroke-field path.

Two small scenes (the 50-stroke solid palette and the mixed curve/solid
palette, each reduced to 24 rays x 32 samples and few strokes) with their
inputs, upstream gradients, the oracle's rgb / acc and exact parameter
gradients.  The CPU suite checks that the oracle still reproduces them; the GPU
suite checks the HIP renderer against them.  Run from the repository root:
``python tests/golden/make_golden.py``.
"""
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from neural3dstrokes_amd import scenes  # noqa: E402
from oracle import field as F  # noqa: E402

KEYS = ('origins', 'directions', 'viewdirs', 'radii', 'tdist', 'shape_params', 'color_params', 'density_params')


def case(name, seed, n_strokes, radius):
    sc = scenes.make_workload(name, seed=seed, n_rays=24, n_samples=32, n_strokes=n_strokes)
    sc['radii'] = torch.full_like(sc['radii'], radius)
    g = torch.Generator().manual_seed(seed + 1)
    g_rgb = torch.randn(24, 3, generator=g, dtype=torch.float64)
    g_acc = torch.randn(24, generator=g, dtype=torch.float64)
    d = {k: sc[k].double() for k in KEYS}
    out, gs, gc, gd = F.render_field_grads(g_rgb, g_acc, d['origins'], d['directions'], d['viewdirs'], d['radii'],
                                           d['tdist'], sc['stroke_ids'], d['shape_params'], d['color_params'],
                                           d['density_params'])
    res = {k: sc[k].numpy() for k in KEYS}  # float32 inputs as the GPU receives them
    res.update(stroke_ids=sc['stroke_ids'].numpy(), g_rgb=g_rgb.numpy(), g_acc=g_acc.numpy(),
               rgb=out['rgb'].detach().numpy(), acc=out['acc'].detach().numpy(), grad_shape=gs.numpy(), grad_color=gc.numpy(),
               grad_density=gd.numpy())
    return {f'{name}_{k}': v for k, v in res.items()}


def main():
    data = {}
    data.update(case('solids50_4096x64', 11, 12, 0.05))
    data.update(case('palette500_65536x128', 12, 16, 0.05))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'field_small.npz')
    np.savez_compressed(path, **data)
    print(f'wrote {path}: ' + ', '.join(f'{k} {v.shape}' for k, v in data.items() if k.endswith(('rgb', 'acc'))))
    for c in ('solids50_4096x64', 'palette500_65536x128'):
        print(c, 'acc > 0.01 on', int((data[f'{c}_acc'] > 0.01).sum()), 'rays')


if __name__ == '__main__':
    main()

"""Autotuner for the render kernels' launch shapes.

Tunes, by timing the real workload on the device, the parameters exposed by
``n3ds_render_tuning``:

* ``flush_pairs``: pending (sample, stroke) pairs that trigger an evaluation
  window (how many strokes' pairs are staged per window);
* ``bwd_oversub``: backward persistent blocks per block resident at once;
* ``batch_oversub``: a scene batch's backward blocks per resident block;
* ``cull_group``: 64-stroke chunks culled by one wavefront of the culling pass
  (a scene batch's plan depends on it: replan after changing it);
* ``bwd_queue``: the adjoint queues its live (sample, stroke) pairs per SDF
  family across rays and evaluates full groups of one family (1), or evaluates
  each window's pairs as they come (0);
* ``cull_grid``: the culling pass of one scene tests each ray against the
  strokes binned in the grid cells its segment crosses (1), or against every
  stroke (0); the hits are the same;
* ``bwd_split``: with the forward's kept hits, the adjoint runs as a per-ray
  kernel that lists its live pairs' adjoint records per SDF family and a kernel
  over those records (1), or as one kernel (0);
* ``split_cap``: records per ray and family the split adjoint's lists hold (a
  fuller list makes the one-kernel adjoint take every ray).

Results do not depend on them beyond float summation order.  ``tune`` runs a
caller-supplied step over a small grid (coordinate descent, one parameter at a
time), keeps the fastest setting and caches it under the caller's key for the
process.
"""
import torch

from . import _lib

PARAMS = {'flush_pairs': 0, 'bwd_oversub': 1, 'batch_oversub': 2, 'cull_group': 3, 'bwd_queue': 4, 'cull_grid': 5,
          'bwd_split': 6, 'split_cap': 7}
SPACE = {'flush_pairs': (32, 48, 64, 96, 128, 192, 256), 'bwd_oversub': (1, 2, 4), 'batch_oversub': (2, 4, 8, 16),
         'cull_group': (1, 2, 4, 8, 12, 16), 'bwd_queue': (0, 1), 'cull_grid': (0, 1), 'bwd_split': (0, 1),
         'split_cap': (16, 32, 64)}
_cache = {}


def get(name):
    return int(_lib.lib().n3ds_render_tuning(PARAMS[name], -1))


def set(name, value):  # noqa: A001 - mirrors get
    old = int(_lib.lib().n3ds_render_tuning(PARAMS[name], int(value)))
    if old < 0:
        _lib.check(1)
    return old


def current():
    return {k: get(k) for k in PARAMS}


def _time(step, reps, warmup):
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        step()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(step, names=('flush_pairs', 'bwd_oversub'), key=None, reps=3, warmup=1, space=None):
    """Coordinate descent over ``names``; returns (best settings, {setting: ms}).

    ``step`` runs one timed unit of the workload (e.g. a fused fwd+bwd).  With a
    ``key`` seen before, the cached settings are applied without timing.
    """
    if key is not None and key in _cache:
        for k, v in _cache[key].items():
            set(k, v)
        return dict(_cache[key]), {}
    space = dict(SPACE, **(space or {}))
    best = current()
    timings = {}
    best_ms = _time(step, reps, warmup)
    timings[tuple(sorted(best.items()))] = best_ms
    for name in names:
        for v in space[name]:
            if v == best[name]:
                continue
            set(name, v)
            trial = dict(best, **{name: v})
            ms = _time(step, reps, warmup)
            timings[tuple(sorted(trial.items()))] = ms
            if ms < best_ms:
                best_ms, best = ms, trial
            set(name, best[name])
    for k, v in best.items():
        set(k, v)
    if key is not None:
        _cache[key] = dict(best)
    return best, {str(dict(k)): v for k, v in timings.items()}

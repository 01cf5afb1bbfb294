"""Many independent stroke scenes trained together, one launch per stage.

The reference trains one scene per process (``train.py:127-144``: forward
through ``StrokeField``, loss, ``clip_gradients``, ``optimizer.step()``,
``clip_params``).  Here a batch of scenes with uneven stroke and ray counts
shares packed arrays: scene ``i`` owns rays ``[r_i, r_i + R_i)`` and strokes
``[k_i, k_i + S_i)`` of every per-ray / per-stroke tensor, and

* ``SceneBatch`` holds the scene table the kernels read (host and device
  copies) with the backward's block plan (blocks per scene in proportion to
  rays x strokes);
* ``render_strokes(..., scenes=batch)`` renders every ray against its own
  scene's strokes in one forward launch and one adjoint launch per class of
  scenes (scenes whose LDS gradient accumulators allow the same wavefronts
  per block), each block working on one scene;
* ``MultiSceneTrainer`` runs the whole training step (forward, photometric
  loss, backward, fused Adam with the per-column clamp) for all scenes, and
  can capture it once as a HIP graph and replay it every step.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .optim import Adam
from .render import render_strokes

SCENE_REC = 10


class SceneBatch:
    """Scene table of a packed batch (int32, host and device copies): per scene first
    ray, rays, first stroke, strokes, the backward plan (first block, blocks,
    partial-row offset, wavefronts per block) and the first culling task, then the
    scene of every backward block."""

    def __init__(self, ray_counts, stroke_counts, n_samples, shape_width, color_width, color_id=0, device='cuda'):
        ray_counts = [int(r) for r in ray_counts]
        stroke_counts = [int(s) for s in stroke_counts]
        if len(ray_counts) != len(stroke_counts) or not ray_counts:
            raise RuntimeError('one ray count and one stroke count per scene')
        if min(ray_counts) < 0 or min(stroke_counts) < 0:
            raise RuntimeError('negative ray / stroke count')
        n = len(ray_counts)
        rec = np.zeros((n, SCENE_REC), dtype=np.int32)
        rec[:, 1] = ray_counts
        rec[:, 3] = stroke_counts
        rec[1:, 0] = np.cumsum(ray_counts)[:-1]
        rec[1:, 2] = np.cumsum(stroke_counts)[:-1]
        self.n_scenes = n
        self.n_samples, self.shape_width, self.color_width, self.color_id = n_samples, shape_width, color_width, color_id
        self.n_rays = int(sum(ray_counts))
        self.n_strokes = int(sum(stroke_counts))
        self.s_max = int(max(stroke_counts))
        self._rec = rec
        self.device = device
        self.replan()

    def replan(self):
        """(Re)computes the backward plan for the current device and tuning."""
        rec, n = self._rec.copy(), self.n_scenes
        n_samples, shape_width, color_width, color_id = (self.n_samples, self.shape_width, self.color_width,
                                                         self.color_id)
        plan = _lib.lib().n3ds_render_batch_plan
        args = (n_samples, shape_width, color_width, color_id)
        nb = int(plan(rec.ctypes.data_as(ctypes.c_void_p), n, *args, None, 0))
        if nb < 0:
            _lib.check(-nb)
        # scene records followed by the scene of every backward block
        self.table = np.zeros(n * SCENE_REC + max(nb, 1), dtype=np.int32)
        self.table[:n * SCENE_REC] = rec.reshape(-1)
        got = int(plan(self.host_ptr(), n, *args, self.table[n * SCENE_REC:].ctypes.data_as(ctypes.c_void_p),
                       max(nb, 1)))
        if got != nb:
            _lib.check(-got if got < 0 else 1)
        self.n_blocks = nb
        self.records = self.table[:n * SCENE_REC].reshape(n, SCENE_REC)
        self.device_table = torch.from_numpy(self.table.copy()).to(self.device)

    def host_ptr(self):
        return self.table.ctypes.data_as(ctypes.c_void_p)

    def c_args(self):
        return (self.host_ptr(), _lib.ptr(self.device_table), self.n_scenes)

    def rays(self, i):
        r = self.records[i]
        return slice(int(r[0]), int(r[0] + r[1]))

    def strokes(self, i):
        r = self.records[i]
        return slice(int(r[2]), int(r[2] + r[3]))

    def check(self, n_rays, n_samples, n_strokes, shape_width, color_width, color_id):
        got = (n_rays, n_samples, n_strokes, shape_width, color_width, color_id)
        want = (self.n_rays, self.n_samples, self.n_strokes, self.shape_width, self.color_width, self.color_id)
        if got != want:
            raise RuntimeError(f'arrays (rays, samples, strokes, shape width, colour width, colour id) = {got} do not '
                               f'match the scene batch {want}')


def pack(scenes):
    """Concatenate per-scene dicts (stroke_ids, shape/color/density params, rays) in scene order."""
    keys = scenes[0].keys()
    out = {}
    for k in keys:
        if torch.is_tensor(scenes[0][k]):
            out[k] = torch.cat([s[k] for s in scenes], 0)
    return out


class MultiSceneTrainer:
    """One photometric training step for every scene of a batch.

    Loss: for each scene, the mean over its rays of ``|rgb - target|^2``
    (summed over scenes, so each scene's gradients are those of its own
    loss).  The optimizer is the fused ``Adam`` with the stroke-parameter
    ranges (``clip_params``) and density >= 0.  ``capture()`` records the step
    as a HIP graph; ``step()`` then replays it.
    """

    def __init__(self, batch, rays, stroke_ids, shape_params, color_params, density_params, targets, lr=1e-3,
                 shape_ranges=None, color_ranges=None, render_kw=None):
        self.batch = batch
        self.rays = {k: rays[k].contiguous() for k in ('origins', 'directions', 'viewdirs', 'radii', 'tdist')}
        self.stroke_ids = stroke_ids
        self.shape = shape_params.detach().clone().requires_grad_(True)
        self.color = color_params.detach().clone().requires_grad_(True)
        self.density = density_params.detach().clone().requires_grad_(True)
        self.targets = targets
        w = torch.empty(batch.n_rays, device=targets.device)
        for i in range(batch.n_scenes):
            sl = batch.rays(i)
            w[sl] = 1.0 / max(1, sl.stop - sl.start)
        self.ray_weight = w[:, None]
        self.lr = torch.full((1,), float(lr), device=targets.device)
        groups = [dict(params=[self.shape], ranges=shape_ranges), dict(params=[self.color], ranges=color_ranges),
                  dict(params=[self.density], min_value=0.0)]
        self.opt = Adam(groups, lr=self.lr, capturable=True)
        self.render_kw = dict(render_kw or {})
        self.graph = None
        self.loss = None

    def _forward_backward_update(self):
        r = self.rays
        rgb, _, _ = render_strokes(r['origins'], r['directions'], r['viewdirs'], r['radii'], r['tdist'],
                                   self.stroke_ids, self.shape, self.color, self.density, scenes=self.batch,
                                   **self.render_kw)
        loss = ((rgb - self.targets) ** 2 * self.ray_weight).sum()
        loss.backward()
        self.opt.step()
        for p in (self.shape, self.color, self.density):
            p.grad = None
        return loss.detach()

    def eager_step(self):
        self.loss = self._forward_backward_update()
        return self.loss

    def capture(self, warmup=2):
        """Run ``warmup`` eager steps (optimizer state, stroke tables), then record one step."""
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                self.loss = self._forward_backward_update()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.loss = self._forward_backward_update()
        torch.cuda.synchronize()

    def step(self):
        if self.graph is None:
            return self.eager_step()
        self.graph.replay()
        return self.loss

"""GPU parity of the batched multi-scene path: one launch per stage for many scenes
must give each scene exactly what rendering that scene alone gives, and the
graph-captured training step must replay the eager step."""
import pytest
import torch

from neural3dstrokes_amd import scenes as W

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda:0')
KEYS = ('origins', 'directions', 'viewdirs', 'radii', 'tdist')


def _rel(a, b):
    return float((a - b).norm() / b.norm().clamp_min(1e-30))


def _batch_case(n_scenes, seed, n_samples, strokes, rays, extra_empty=True):
    sc, rc, kc = W.make_multiscene(n_scenes=n_scenes, seed=seed, stroke_range=strokes, ray_range=rays,
                                   n_samples=n_samples)
    if extra_empty:
        # a scene with rays but no strokes, and one with strokes but no rays
        g = torch.Generator().manual_seed(seed + 99)
        r = W.make_rays(37, n_samples, g)
        empty = dict(stroke_ids=torch.zeros(0, dtype=torch.int32), shape_params=torch.zeros(0, 14),
                     color_params=torch.zeros(0, 3), density_params=torch.zeros(0))
        empty.update(r)
        s2 = W.make_strokes(W.palette(25, g, 'mixed'), g)
        s2['shape_params'] = torch.cat([s2['shape_params'], torch.zeros(25, 14 - s2['shape_params'].shape[1])], 1)
        s2.update({k: v[:0] for k, v in W.make_rays(1, n_samples, g).items()})
        sc = sc[:1] + [empty] + sc[1:2] + [s2] + sc[2:]
        rc = rc[:1] + [37] + rc[1:2] + [0] + rc[2:]
        kc = kc[:1] + [0] + kc[1:2] + [25] + kc[2:]
    return sc, rc, kc


@pytest.mark.parametrize('n_samples,store', [(128, 'auto'), (64, False), (192, 'auto')])
def test_batch_matches_per_scene(n_samples, store):
    from neural3dstrokes_amd.multiscene import SceneBatch, pack
    from neural3dstrokes_amd.render import render_strokes
    sc, rc, kc = _batch_case(5, 11, n_samples, (20, 300), (64, 700))
    packed = {k: v.to(DEV) for k, v in pack(sc).items()}
    batch = SceneBatch(rc, kc, n_samples, 14, 3, 0, device=DEV)
    print(f'scenes rays={rc} strokes={kc} blocks={batch.n_blocks}', flush=True)
    leaves = [packed[k].clone().requires_grad_(True) for k in ('shape_params', 'color_params', 'density_params')]
    out = render_strokes(*[packed[k] for k in KEYS], packed['stroke_ids'], *leaves, scenes=batch, store=store)
    g = torch.Generator().manual_seed(5)
    ups = [torch.randn(o.shape, generator=g).to(DEV) for o in out]
    grads = torch.autograd.grad(out, leaves, ups)
    torch.cuda.synchronize()
    for i in range(batch.n_scenes):
        rs, ks = batch.rays(i), batch.strokes(i)
        if rs.stop == rs.start:
            assert all(float(gr[ks].abs().max()) == 0.0 for gr in grads if ks.stop > ks.start)
            continue
        if ks.stop == ks.start:
            assert torch.equal(out[0][rs], torch.ones_like(out[0][rs])) and float(out[1][rs].detach().abs().max()) == 0.0
            continue
        lv = [packed[k][ks].clone().requires_grad_(True) for k in ('shape_params', 'color_params', 'density_params')]
        o1 = render_strokes(*[packed[k][rs].contiguous() for k in KEYS], packed['stroke_ids'][ks].contiguous(), *lv,
                            store=store)
        g1 = torch.autograd.grad(o1, lv, [u[rs] for u in ups])
        for a, b in zip(out, o1):
            assert torch.equal(a[rs], b), f'scene {i}: batched forward differs from the scene alone'
        for a, b in zip(grads, g1):
            assert _rel(a[ks], b) <= 1e-5, f'scene {i}: gradient rel err {_rel(a[ks], b):.3e}'


def test_batch_rejects_mismatched_arrays():
    from neural3dstrokes_amd.multiscene import SceneBatch, pack
    from neural3dstrokes_amd.render import render_strokes
    sc, rc, kc = _batch_case(2, 3, 64, (20, 40), (16, 32), extra_empty=False)
    packed = {k: v.to(DEV) for k, v in pack(sc).items()}
    batch = SceneBatch(rc[:1] + [rc[1] + 1], kc, 64, 14, 3, 0, device=DEV)
    with pytest.raises(RuntimeError):
        render_strokes(*[packed[k] for k in KEYS], packed['stroke_ids'], packed['shape_params'],
                       packed['color_params'], packed['density_params'], scenes=batch)


def _trainer(seed, n_scenes=4):
    from neural3dstrokes_amd.multiscene import MultiSceneTrainer, SceneBatch, pack
    sc, rc, kc = W.make_multiscene(n_scenes=n_scenes, seed=seed, stroke_range=(20, 200), ray_range=(256, 2048),
                                   n_samples=128)
    packed = {k: v.to(DEV) for k, v in pack(sc).items()}
    batch = SceneBatch(rc, kc, 128, 14, 3, 0, device=DEV)
    g = torch.Generator().manual_seed(seed + 1)
    targets = torch.rand(batch.n_rays, 3, generator=g).to(DEV)
    return MultiSceneTrainer(batch, packed, packed['stroke_ids'], packed['shape_params'], packed['color_params'],
                             packed['density_params'], targets, lr=1e-3), packed, batch, targets


def test_multiscene_graph_step_replays_eager_steps():
    graph_t, _, _, _ = _trainer(21)
    eager_t, _, _, _ = _trainer(21)
    graph_t.capture(warmup=2)
    for _ in range(2):
        eager_t.eager_step()
    losses = []
    for step in range(4):
        lg = float(graph_t.step())
        le = float(eager_t.eager_step())
        losses.append(le)
        assert abs(lg - le) <= 1e-5 * abs(le) + 1e-7, f'step {step}: loss {lg} vs {le}'
    torch.cuda.synchronize()
    assert float(graph_t.opt.state[graph_t.shape]['step']) == 6
    for a, b in ((graph_t.shape, eager_t.shape), (graph_t.color, eager_t.color), (graph_t.density, eager_t.density)):
        d = (a - b).detach().abs()
        # Adam moves a parameter by at most ~lr per step whatever its gradient; a gradient that
        # is zero up to rounding may take either sign, so bound the rare outliers by that
        assert float(d.max()) <= 6 * 2e-3
        assert float((d > 1e-5 + 1e-4 * b.abs()).float().mean()) <= 0.01
    assert losses[-1] < losses[0]


def test_multiscene_step_gradients_match_per_scene_loop():
    """The gradients one batched step hands to Adam are, scene by scene, those of the
    per-scene launch loop (the loss is a sum of per-scene losses)."""
    from neural3dstrokes_amd.render import render_strokes
    t, packed, batch, targets = _trainer(33)
    r = t.rays
    rgb, _, _ = render_strokes(r['origins'], r['directions'], r['viewdirs'], r['radii'], r['tdist'], t.stroke_ids,
                               t.shape, t.color, t.density, scenes=batch)
    loss = ((rgb - targets) ** 2 * t.ray_weight).sum()
    gb = torch.autograd.grad(loss, [t.shape, t.color, t.density])
    for i in range(batch.n_scenes):
        rs, ks = batch.rays(i), batch.strokes(i)
        lv = [x[ks].detach().clone().requires_grad_(True) for x in (t.shape, t.color, t.density)]
        o = render_strokes(*[packed[k][rs].contiguous() for k in KEYS], packed['stroke_ids'][ks].contiguous(), *lv)
        li = ((o[0] - targets[rs]) ** 2).sum() / (rs.stop - rs.start)
        gi = torch.autograd.grad(li, lv)
        for a, b in zip(gb, gi):
            assert _rel(a[ks], b) <= 1e-5


def test_batch_plan_follows_cull_group():
    """A batch planned with several chunks per culling task renders as the scenes alone."""
    from neural3dstrokes_amd import autotune
    from neural3dstrokes_amd.multiscene import SceneBatch, pack
    from neural3dstrokes_amd.render import render_strokes
    old = autotune.current()
    try:
        autotune.set('cull_group', 3)
        sc, rc, kc = _batch_case(3, 17, 128, (60, 300), (64, 300), extra_empty=False)
        packed = {k: v.to(DEV) for k, v in pack(sc).items()}
        batch = SceneBatch(rc, kc, 128, 14, 3, 0, device=DEV)
        out = render_strokes(*[packed[k] for k in KEYS], packed['stroke_ids'], packed['shape_params'],
                             packed['color_params'], packed['density_params'], scenes=batch)
        for i in range(batch.n_scenes):
            rs, ks = batch.rays(i), batch.strokes(i)
            o1 = render_strokes(*[packed[k][rs].contiguous() for k in KEYS], packed['stroke_ids'][ks].contiguous(),
                                packed['shape_params'][ks], packed['color_params'][ks], packed['density_params'][ks])
            for a, b in zip(out, o1):
                assert torch.equal(a[rs], b)
    finally:
        for k, v in old.items():
            autotune.set(k, v)


@pytest.mark.parametrize('color_id', [1, 3])
def test_batch_color_types_match_per_scene(color_id):
    """gradient_rgb (kept colour planes) and SH colour in a batch equal the scenes alone."""
    from neural3dstrokes_amd.multiscene import SceneBatch, pack
    from neural3dstrokes_amd.render import COLOR_WIDTH, render_strokes
    sc, rc, kc = W.make_multiscene(n_scenes=3, seed=29, stroke_range=(20, 150), ray_range=(32, 160),
                                   n_samples=128, color_id=color_id)
    packed = {k: v.to(DEV) for k, v in pack(sc).items()}
    batch = SceneBatch(rc, kc, 128, 14, COLOR_WIDTH[color_id], color_id, device=DEV)
    leaves = [packed[k].clone().requires_grad_(True) for k in ('shape_params', 'color_params', 'density_params')]
    out = render_strokes(*[packed[k] for k in KEYS], packed['stroke_ids'], *leaves, color_id=color_id, scenes=batch)
    g = torch.Generator().manual_seed(6)
    ups = [torch.randn(o.shape, generator=g).to(DEV) for o in out]
    grads = torch.autograd.grad(out, leaves, ups)
    for i in range(batch.n_scenes):
        rs, ks = batch.rays(i), batch.strokes(i)
        lv = [packed[k][ks].clone().requires_grad_(True) for k in ('shape_params', 'color_params', 'density_params')]
        o1 = render_strokes(*[packed[k][rs].contiguous() for k in KEYS], packed['stroke_ids'][ks].contiguous(), *lv,
                            color_id=color_id)
        g1 = torch.autograd.grad(o1, lv, [u[rs] for u in ups])
        for a, b in zip(out, o1):
            assert torch.equal(a[rs].detach(), b.detach())
        for a, b in zip(grads, g1):
            assert _rel(a[ks], b) <= 1e-5


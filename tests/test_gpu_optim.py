"""GPU parity of the fused optimizer stage (neural3dstrokes_amd.optim.Adam) against
torch's own clip_grad_norm_ / clip_grad_value_ / nan_to_num_ / torch.optim.Adam /
clamp_ sequence (training.py:67-77, 119-136; models.py:576-584), and against the
fp64 oracle restatement."""
import math

import pytest
import torch

from oracle import optim as O

pytestmark = pytest.mark.gpu

RANGES = [(0.01, 0.5), (None, 0.0), (-1.0, None), (None, None), (-0.2, 0.2)]


def _setup(seed, dev):
    g = torch.Generator().manual_seed(seed)
    shape = torch.rand(257, 5, generator=g) * 0.6
    color = torch.randn(257, 3, generator=g)
    dens = torch.rand(257, generator=g)
    return [t.to(dev) for t in (shape, color, dens)]


def _grads(step, dev, bad):
    g = torch.Generator().manual_seed(100 + step)
    gs = [torch.randn(257, 5, generator=g) * 3, torch.randn(257, 3, generator=g), torch.randn(257, generator=g) * 0.1]
    if bad:
        gs[0][4, 1] = float('nan')
        gs[1][7, 2] = float('inf')
        gs[2][9] = -float('inf')
    return [x.to(dev) for x in gs]


def _torch_step(params, grads, opt, max_norm, max_val):
    for p, gr in zip(params, grads):
        p.grad = gr.clone()
    if max_norm > 0:
        torch.nn.utils.clip_grad_norm_(params, max_norm)
    if max_val > 0:
        torch.nn.utils.clip_grad_value_(params, max_val)
    for p in params:
        p.grad.nan_to_num_()
    opt.step()
    with torch.no_grad():
        for i, (lo, hi) in enumerate(RANGES):
            if lo is not None or hi is not None:
                params[0][:, i].clamp_(lo, hi)
        params[2].clamp_(min=0.0)


@pytest.mark.parametrize('max_norm,max_val,bad', [(0.0, 0.0, False), (0.0, 0.0, True), (1.5, 0.0, False),
                                                  (0.0, 0.4, True), (1.5, 0.4, False)])
def test_fused_adam_matches_torch(max_norm, max_val, bad):
    from neural3dstrokes_amd.optim import Adam
    dev = torch.device('cuda:0')
    ref = [t.clone().requires_grad_(True) for t in _setup(0, dev)]
    ours = [t.clone().requires_grad_(True) for t in _setup(0, dev)]
    topt = torch.optim.Adam(ref, lr=5e-3, betas=(0.9, 0.99), eps=1e-15)
    fopt = Adam([dict(params=[ours[0]], ranges=RANGES), dict(params=[ours[1]]), dict(params=[ours[2]], min_value=0.0)],
                lr=5e-3, betas=(0.9, 0.99), eps=1e-15, grad_max_norm=max_norm, grad_max_val=max_val)
    for step in range(1, 7):
        lr = 5e-3 * (0.8 ** step)
        for o in (topt, fopt):
            for grp in o.param_groups:
                grp['lr'] = lr
        grads = _grads(step, dev, bad and step % 2 == 1)
        _torch_step(ref, grads, topt, max_norm, max_val)
        for p, gr in zip(ours, grads):
            p.grad = gr.clone()
        fopt.step()
        for a, b in zip(ours, ref):
            torch.testing.assert_close(a.grad, b.grad, rtol=1e-5, atol=1e-7, equal_nan=True)
            torch.testing.assert_close(a.detach(), b.detach(), rtol=1e-5, atol=1e-6, equal_nan=True)
            torch.testing.assert_close(fopt.state[a]['exp_avg'], topt.state[b]['exp_avg'], rtol=1e-5, atol=1e-7)
            torch.testing.assert_close(fopt.state[a]['exp_avg_sq'], topt.state[b]['exp_avg_sq'], rtol=1e-5,
                                       atol=1e-9)
        assert ours[0][:, 0].min() >= 0.01 and ours[0][:, 1].max() <= 0.0 and ours[2].min() >= 0.0


def test_fused_adam_matches_fp64_oracle():
    from neural3dstrokes_amd.optim import Adam
    dev = torch.device('cuda:0')
    p0 = _setup(1, dev)[0]
    ours = p0.clone().requires_grad_(True)
    opt = Adam([dict(params=[ours], ranges=RANGES)], lr=1e-2, grad_max_norm=2.0, grad_max_val=0.5)
    p, m, v = p0.double().cpu(), torch.zeros(257, 5, dtype=torch.float64), torch.zeros(257, 5, dtype=torch.float64)
    for step in range(1, 5):
        gr = _grads(step, dev, False)[0]
        ours.grad = gr.clone()
        opt.step()
        gd = O.clip_gradients([gr.double().cpu()], max_norm=2.0, max_val=0.5)[0]
        p, m, v = O.adam_step(p, gd, m, v, step, 1e-2)
        p = O.clip_params(p, ranges=RANGES)
    torch.testing.assert_close(ours.detach().cpu().double(), p, rtol=1e-5, atol=1e-6)


def test_capturable_adam_graph_replays():
    """A captured step replays with the device step count and a device learning rate
    (set between replays), matching eager torch.optim.Adam step by step."""
    from neural3dstrokes_amd.optim import Adam
    dev = torch.device('cuda:0')
    ref = [t.clone().requires_grad_(True) for t in _setup(2, dev)]
    ours = [t.clone().requires_grad_(True) for t in _setup(2, dev)]
    lr_t = torch.full((1,), 4e-3, device=dev)
    topt = torch.optim.Adam(ref, lr=4e-3, betas=(0.9, 0.99), eps=1e-15)
    fopt = Adam([dict(params=[ours[0]], ranges=RANGES), dict(params=[ours[1]]), dict(params=[ours[2]], min_value=0.0)],
                lr=lr_t, grad_max_norm=3.0, capturable=True)
    static = [torch.zeros_like(p) for p in ours]
    for p, s in zip(ours, static):
        p.grad = s
    # eager first step (state, bounds and accumulators allocated outside the capture)
    grads = _grads(1, dev, False)
    for s, gr in zip(static, grads):
        s.copy_(gr)
    _torch_step(ref, grads, topt, 3.0, 0.0)
    fopt.step()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    snap = [(p.detach().clone(), {k: v.clone() for k, v in fopt.state[p].items()}) for p in ours]
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            fopt.step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    with torch.no_grad():  # capture does not run the kernels; restore anyway in case it did
        for p, (pv, st) in zip(ours, snap):
            p.copy_(pv)
            for k, v in st.items():
                fopt.state[p][k].copy_(v)
    for step in range(2, 7):
        lr = 4e-3 * math.exp(-0.3 * step)
        lr_t.fill_(lr)
        for grp in topt.param_groups:
            grp['lr'] = lr
        grads = _grads(step, dev, step == 4)
        for s, gr in zip(static, grads):
            s.copy_(gr)
        _torch_step(ref, grads, topt, 3.0, 0.0)
        graph.replay()
        torch.cuda.synchronize()
        assert float(fopt.state[ours[0]]['step']) == step
        for a, b in zip(ours, ref):
            torch.testing.assert_close(a.detach(), b.detach(), rtol=1e-5, atol=1e-6)


def test_adam_rejects_bad_arguments():
    from neural3dstrokes_amd.optim import Adam
    p = torch.zeros(4, 3, device='cuda:0', requires_grad=True)
    p.grad = torch.ones_like(p)
    with pytest.raises(RuntimeError):
        Adam([dict(params=[p], ranges=[(0, 1)] * 2)]).step()
    q = torch.zeros(4, 3, requires_grad=True)
    q.grad = torch.ones_like(q)
    with pytest.raises(RuntimeError):
        Adam([q]).step()

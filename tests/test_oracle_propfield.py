"""Pins of the proposal-field oracle (oracle/propfield.py)."""
import numpy as np
import torch

from neural3dstrokes_amd.gridencoder import GridEncoder
from oracle import propfield as oprop

D = torch.float64


def test_contract_closed_form():
    g = torch.Generator().manual_seed(0)
    x = torch.randn(2000, 3, generator=g, dtype=D) * 3
    z = oprop.contract(x)
    r = x.norm(dim=-1)
    inside = r <= 1
    assert torch.equal(z[inside], x[inside])
    # outside: radial, |z| = 2 - 1/|x|
    out = ~inside
    assert torch.allclose(z[out].norm(dim=-1), 2 - 1 / r[out], atol=1e-12)
    cos = (z[out] * x[out]).sum(-1) / (z[out].norm(dim=-1) * r[out])
    assert torch.allclose(cos, torch.ones_like(cos), atol=1e-12)
    # continuous at the unit sphere, and |z| < 2 everywhere
    u = x / x.norm(dim=-1, keepdim=True)
    assert torch.allclose(oprop.contract(u * (1 + 1e-9)), u, atol=1e-8)
    assert float(oprop.contract(x * 1e6).norm(dim=-1).max()) < 2.0


def _case(seed=0, C=1, L=6, n=64):
    torch.manual_seed(seed)
    enc = GridEncoder(3, L, C, base_resolution=8, log2_hashmap_size=10, desired_resolution=64)
    emb = (torch.rand(enc.embeddings.shape, dtype=D) - 0.5) * 0.4
    g = torch.Generator().manual_seed(seed + 1)
    coords = torch.randn(n, 3, generator=g, dtype=D) * 1.5
    lin0, lin2 = torch.nn.Linear(L * C, 64).double(), torch.nn.Linear(64, 1).double()
    gd = torch.randn(n, generator=g, dtype=D)
    return enc, emb, coords, lin0, lin2, gd


def test_head_equals_torch_modules():
    enc, emb, coords, lin0, lin2, gd = _case()
    feats, _, _ = oprop.features(coords, emb.numpy(), enc.offsets.numpy(), enc.per_level_scale, enc.base_resolution)
    want = torch.nn.functional.softplus(lin2(torch.relu(lin0(torch.from_numpy(feats))))[:, 0] - 1.0)
    got = oprop.density_and_grads(coords, emb.numpy(), enc.offsets.numpy(), enc.per_level_scale, enc.base_resolution,
                                  lin0.weight.detach(), lin0.bias.detach(), lin2.weight.detach(), lin2.bias.detach(),
                                  -1.0, gd)
    assert torch.allclose(got['density'], want.detach(), atol=1e-14)


def test_gradients_match_central_differences():
    enc, emb, coords, lin0, lin2, gd = _case(seed=3, C=2, L=5, n=48)
    offs = enc.offsets.numpy()
    w1, b1, w2, b2 = (t.detach().clone() for t in (lin0.weight, lin0.bias, lin2.weight, lin2.bias))

    def loss(emb_, w1_):
        r = oprop.density_and_grads(coords, emb_, offs, enc.per_level_scale, enc.base_resolution, w1_, b1, w2, b2,
                                    -1.0, gd)
        return float((r['density'] * gd).sum())

    ref = oprop.density_and_grads(coords, emb.numpy(), offs, enc.per_level_scale, enc.base_resolution, w1, b1, w2,
                                  b2, -1.0, gd)
    h = 1e-6
    g = torch.Generator().manual_seed(5)
    for _ in range(6):
        j, i = int(torch.randint(0, 64, (1,), generator=g)), int(torch.randint(0, 10, (1,), generator=g))
        wp, wm = w1.clone(), w1.clone()
        wp[j, i] += h
        wm[j, i] -= h
        fd = (loss(emb.numpy(), wp) - loss(emb.numpy(), wm)) / (2 * h)
        assert abs(fd - float(ref['grad_w1'][j, i])) < 1e-6 * max(1.0, abs(fd))
    # table rows the points touch
    rows = np.nonzero(np.abs(ref['grad_embeddings'].numpy()).sum(1) > 1e-8)[0]
    for r in rows[:: max(1, len(rows) // 6)][:6]:
        ep, em = emb.numpy().copy(), emb.numpy().copy()
        ep[r, 0] += h
        em[r, 0] -= h
        fd = (loss(ep, w1) - loss(em, w1)) / (2 * h)
        assert abs(fd - float(ref['grad_embeddings'][r, 0])) < 1e-6 * max(1.0, abs(fd))

"""GPU guide network (hand-written MFMA forward / backward / Adam) against
torch FP32 autograd and torch.optim.Adam; device vMF mixture sampling and
density against quadrature and a chi-square test."""
import math

import numpy as np
import pytest

from lm_amd import guide as G
from oracle import guide_ref as ref

pytestmark = pytest.mark.gpu

BOX_LO, BOX_HI = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)


def random_records(n, rng, concentrated=False):
    pos = rng.uniform(-1, 1, size=(n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    if concentrated:
        # target: incident light around a position-dependent direction
        t = np.stack([np.sin(3 * pos[:, 0]), np.cos(3 * pos[:, 0]), 0.5 * pos[:, 1]], 1)
        t /= np.linalg.norm(t, axis=1, keepdims=True)
        wi = t + 0.15 * rng.normal(size=(n, 3))
    else:
        wi = rng.normal(size=(n, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    pdf = np.full((n, 1), 1 / (4 * math.pi)) if concentrated else rng.uniform(0.1, 2.0, size=(n, 1))
    L = np.repeat(rng.uniform(0.0, 3.0, size=(n, 1)), 3, axis=1) if not concentrated else np.ones((n, 3))
    rec = np.concatenate([pos, nrm, wo, wi, pdf, L], axis=1).astype(np.float32)
    return rec


def test_forward_matches_torch(device):
    import torch
    rng = np.random.default_rng(0)
    g = G.Guide(BOX_LO, BOX_HI, seed=3)
    p = g.get_params()
    # larger random weights than the init so that all lobes are exercised
    p = (p * 3.0).astype(np.float32)
    g.set_params(p)
    n = 2000
    rec = random_records(n, rng)
    q = torch.from_numpy(rec[:, :9].copy()).to(device)
    mix = g.infer(q).cpu().numpy()
    layers = ref.unpack(g.get_params())
    r = torch.from_numpy(rec).double().float()
    x = ref.encode(r[:, 0:3], r[:, 3:6], r[:, 6:9], BOX_LO, BOX_HI)
    y = ref.forward(layers, x)
    pi, kappa, mu = ref.decode(y)
    # identical BF16 rounding points; FP32 sums in another order can move a
    # hidden activation by one BF16 ulp
    dpi = np.abs(mix[:, :8] - pi.numpy())
    dk = np.abs(mix[:, 8:16] - kappa.numpy()) / np.maximum(kappa.numpy(), 1e-3)
    cos = (mix[:, 16:].reshape(-1, 8, 3) * mu.numpy()).sum(-1)
    print(f'max |dpi| {dpi.max():.2e} mean {dpi.mean():.2e}; max rel dkappa {dk.max():.2e}; min cos mu {cos.min():.6f}')
    assert dpi.max() < 1e-2 and dpi.mean() < 1e-4
    assert dk.max() < 3e-2 and dk.mean() < 1e-3
    assert cos.min() > 0.999


def test_gradient_matches_autograd(device):
    import torch
    rng = np.random.default_rng(1)
    g = G.Guide(BOX_LO, BOX_HI, seed=5)
    p0 = g.get_params()
    n = 4096
    rec = random_records(n, rng)
    loss, grad = g.train(torch.from_numpy(rec).to(device), lr=0.0, want_grad=True)
    layers = ref.unpack(p0)
    params = []
    for W, b in layers:
        W.requires_grad_(True)
        b.requires_grad_(True)
        params += [W, b]
    lt = ref.loss_fn(layers, torch.from_numpy(rec), BOX_LO, BOX_HI)
    lt.backward()
    print(f'loss gpu {loss:.6f} torch {lt.item():.6f}')
    assert abs(loss - lt.item()) < 1e-3 * max(1.0, abs(lt.item()))
    got = G.unpack_params(grad)
    for li, ((Wg, bg), (W, b)) in enumerate(zip(got, layers)):
        for a, e, name in ((Wg, W.grad.numpy(), 'W'), (bg, b.grad.numpy(), 'b')):
            a = np.asarray(a, dtype=np.float64).ravel()
            e = np.asarray(e, dtype=np.float64).ravel()
            cos = a @ e / (np.linalg.norm(a) * np.linalg.norm(e) + 1e-30)
            rel = np.linalg.norm(a - e) / (np.linalg.norm(e) + 1e-30)
            print(f'layer {li + 1} {name}: cos {cos:.6f} rel {rel:.3e}')
            assert cos > 0.999 and rel < 3e-2, (li, name, cos, rel)


def test_gradient_ragged_batch_after_a_step(device):
    """A batch that ends inside a 128-row tile, right after another step:
    the gradient is that batch's alone (the buffer is cleared between steps)
    and matches autograd; the step count advances once per step."""
    import torch
    rng = np.random.default_rng(4)
    g = G.Guide(BOX_LO, BOX_HI, seed=11)
    p0 = g.get_params()
    g.train(torch.from_numpy(random_records(2048, rng)).to(device), lr=0.0)
    rec = random_records(1000, rng)
    loss, grad = g.train(torch.from_numpy(rec).to(device), lr=0.0, want_grad=True)
    assert g.stats()[0] == 2
    assert np.array_equal(g.get_params(), p0)
    layers = ref.unpack(p0)
    for W, b in layers:
        W.requires_grad_(True)
        b.requires_grad_(True)
    lt = ref.loss_fn(layers, torch.from_numpy(rec), BOX_LO, BOX_HI)
    lt.backward()
    assert abs(loss - lt.item()) < 1e-3 * max(1.0, abs(lt.item()))
    got = G.unpack_params(grad)
    for li, ((Wg, bg), (W, b)) in enumerate(zip(got, layers)):
        for a, e in ((Wg, W.grad.numpy()), (bg, b.grad.numpy())):
            a = np.asarray(a, dtype=np.float64).ravel()
            e = np.asarray(e, dtype=np.float64).ravel()
            cos = a @ e / (np.linalg.norm(a) * np.linalg.norm(e) + 1e-30)
            rel = np.linalg.norm(a - e) / (np.linalg.norm(e) + 1e-30)
            assert cos > 0.999 and rel < 3e-2, (li, cos, rel)


def test_adam_step_matches_torch(device):
    import torch
    rng = np.random.default_rng(2)
    g = G.Guide(BOX_LO, BOX_HI, seed=7)
    p0 = g.get_params()
    rec = torch.from_numpy(random_records(2048, rng)).to(device)
    _, grad = g.train(rec, lr=1e-3, want_grad=True)
    p1 = g.get_params()
    t = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([t], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    t.grad = torch.tensor(grad.copy())
    opt.step()
    assert np.allclose(p1, t.detach().numpy(), rtol=1e-5, atol=1e-7)


def test_training_learns_a_direction_field(device):
    """200 Adam steps on a position-dependent incident direction: the loss
    drops and the learned density at the target direction rises well above
    the uniform 1/(4 pi)."""
    import torch
    rng = np.random.default_rng(3)
    g = G.Guide(BOX_LO, BOX_HI, seed=9)
    losses = []
    for it in range(200):
        rec = torch.from_numpy(random_records(1 << 14, rng, concentrated=True)).to(device)
        loss, _ = g.train(rec, lr=3e-3)
        losses.append(loss)
    print('loss', losses[0], '->', losses[-1])
    assert losses[-1] < losses[0] - 1.0
    test = random_records(4096, rng, concentrated=True)
    q = torch.from_numpy(test[:, :9].copy()).to(device)
    mix = g.infer(q)
    pos = test[:, :3]
    t = np.stack([np.sin(3 * pos[:, 0]), np.cos(3 * pos[:, 0]), 0.5 * pos[:, 1]], 1)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    pdf = G.pdf(mix, torch.from_numpy(t.astype(np.float32)).to(device)).cpu().numpy()
    print('median pdf at target', np.median(pdf))
    assert np.median(pdf) > 10 * (1 / (4 * math.pi))


def test_mixture_sampling_matches_pdf(device):
    """Device vMF-mixture sampling: the returned pdf equals the mixture pdf
    of the sample, the pdf integrates to 1 over the sphere (quadrature), and
    sample counts per sphere cell follow the pdf (chi-square)."""
    import torch
    rng = np.random.default_rng(4)
    pi = np.array([0.3, 0.2, 0.1, 0.1, 0.1, 0.1, 0.05, 0.05])
    kappa = np.array([1.0, 10.0, 100.0, 0.1, 3.0, 30.0, 1000.0, 0.5])
    mu = rng.normal(size=(8, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    one = np.concatenate([pi, kappa, mu.ravel()]).astype(np.float32)
    n = 1_000_000
    mix = torch.from_numpy(np.tile(one, (n, 1))).to(device)
    u = torch.from_numpy(rng.random((n, 3)).astype(np.float32)).to(device)
    dirs, pdf = G.sample(mix, u)
    pdf2 = G.pdf(mix, dirs)
    assert torch.allclose(pdf, pdf2, rtol=1e-5)
    d = dirs.cpu().numpy()
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    # quadrature over (cos theta, phi) cells
    nz, nph = 400, 800
    z = -1 + (np.arange(nz) + 0.5) * 2 / nz
    ph = (np.arange(nph) + 0.5) * 2 * np.pi / nph
    Z, PH = np.meshgrid(z, ph, indexing='ij')
    s = np.sqrt(1 - Z ** 2)
    q = np.stack([s * np.cos(PH), s * np.sin(PH), Z], -1).reshape(-1, 3).astype(np.float32)
    mq = torch.from_numpy(np.tile(one, (q.shape[0], 1))).to(device)
    pq = G.pdf(mq, torch.from_numpy(q).to(device)).cpu().numpy().astype(np.float64)
    cell = (2 / nz) * (2 * np.pi / nph)
    total = pq.sum() * cell
    print('pdf integral', total)
    assert abs(total - 1) < 2e-2   # the kappa=1000 lobe is narrower than a cell
    # chi-square on coarse (z, phi) bins against the exact bin masses
    bz, bp = 16, 32
    iz = np.clip(((d[:, 2] + 1) / 2 * bz).astype(int), 0, bz - 1)
    ip = np.clip((np.mod(np.arctan2(d[:, 1], d[:, 0]), 2 * np.pi) / (2 * np.pi) * bp).astype(int), 0, bp - 1)
    hist = np.bincount(iz * bp + ip, minlength=bz * bp).astype(np.float64)
    qz = np.clip(((q[:, 2] + 1) / 2 * bz).astype(int), 0, bz - 1)
    qp = np.clip((np.mod(np.arctan2(q[:, 1], q[:, 0]), 2 * np.pi) / (2 * np.pi) * bp).astype(int), 0, bp - 1)
    mass = np.bincount(qz * bp + qp, weights=pq * cell, minlength=bz * bp)
    exp = mass / mass.sum() * n
    m = exp > 50
    chi2 = ((hist[m] - exp[m]) ** 2 / exp[m]).sum()
    from scipy.stats import chi2 as c2
    pval = c2.sf(chi2, m.sum() - 1)
    print('chi2', chi2, 'dof', m.sum() - 1, 'p', pval)
    assert pval > 1e-4

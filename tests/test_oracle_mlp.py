"""Pins of the cooling-surrogate network's CPU oracle (oracle/mlp.py), CPU only.

The FP64 gradient is checked against central finite differences and against
torch autograd on the CPU (a library routine, not the reference); the BF16
rounding against torch's bfloat16 conversion; Adam against torch.optim.Adam.
A wrong sign, a transposed weight or a dropped bias fails one of them.
"""
import numpy as np
import pytest

from oracle import mlp


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (n, 4)), rng.normal(0, 1, n)


def test_layout():
    assert mlp.NPARAM == 33793
    p = np.arange(mlp.NPARAM, dtype=np.float64)
    assert np.array_equal(mlp.pack(mlp.unpack(p)), p)
    b = mlp.unpack(p)
    assert b["W1"].shape == (128, 4) and b["W1"][1, 0] == 4  # [out][in], row-major
    assert b["b4"][0] == mlp.NPARAM - 1


def test_gradient_vs_finite_differences():
    p = mlp.init(3).astype(np.float64)
    x, t = _batch(8, 4)
    g, loss = mlp.gradient(p, x, t)
    rng = np.random.default_rng(5)
    sl = mlp.block_slices()
    idx = np.concatenate([rng.choice(np.arange(s.start, s.stop), min(6, s.stop - s.start), replace=False)
                          for s in sl.values()])
    eps = 1e-6
    for i in idx:
        pp = p.copy(); pp[i] += eps
        pm = p.copy(); pm[i] -= eps
        lp = np.mean((mlp.forward(pp, x)[0] - t) ** 2)
        lm = np.mean((mlp.forward(pm, x)[0] - t) ** 2)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
    assert loss == pytest.approx(np.mean((mlp.forward(p, x)[0] - t) ** 2), rel=1e-15)


def test_gradient_vs_torch_autograd():
    torch = pytest.importorskip("torch")
    p = mlp.init(7).astype(np.float64)
    x, t = _batch(50, 8)
    g, loss = mlp.gradient(p, x, t)
    b = {k: torch.tensor(v.copy(), dtype=torch.float64, requires_grad=True) for k, v in mlp.unpack(p).items()}
    xt = torch.tensor(x)
    h = torch.nn.functional.relu(xt @ b["W1"].T + b["b1"])
    h = torch.nn.functional.relu(h @ b["W2"].T + b["b2"])
    h = torch.nn.functional.relu(h @ b["W3"].T + b["b3"])
    y = h @ b["W4"] + b["b4"][0]
    lt = torch.mean((y - torch.tensor(t)) ** 2)
    lt.backward()
    gt = mlp.pack({k: v.grad.numpy() for k, v in b.items()})
    assert float(lt.detach()) == pytest.approx(loss, rel=1e-13)
    assert np.max(np.abs(gt - g)) <= 1e-12 * np.max(np.abs(gt))


def test_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(11)
    a = np.concatenate([rng.normal(0, 3, 100000), rng.uniform(-1e-30, 1e-30, 1000),
                        [0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 65504.0, 3.0e38]]).astype(np.float32)
    ours = mlp.bf16(a)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    assert mlp.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0  # tie to even
    assert mlp.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == np.float32(1.0 + 2.0 ** -6)


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    p0 = rng.normal(0, 1, 64)
    w = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([w], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(64), np.zeros(64)
    for step in range(1, 6):
        g = rng.normal(0, 1, 64)
        opt.zero_grad()
        w.grad = torch.tensor(g)
        opt.step()
        p, m, v = mlp.adam(p, g, m, v, step, 3e-3)
    assert np.max(np.abs(w.detach().numpy() - p)) < 1e-12


def test_emulated_forward_is_rounding_only():
    p = mlp.init(9)
    x, t = _batch(256, 10)
    y64, _ = mlp.forward(p, x)
    yem, _ = mlp.forward(p, x, emulate=True)
    scale = np.sqrt(np.mean(y64 ** 2))
    assert np.sqrt(np.mean((yem - y64) ** 2)) < 2e-2 * scale
    g64, _ = mlp.gradient(p, x, t)
    gem, _ = mlp.gradient(p, x, t, emulate=True)
    for k, s in mlp.block_slices().items():
        assert np.linalg.norm(gem[s] - g64[s]) < 5e-2 * np.linalg.norm(g64[s]), k

"""The CPU oracle of the surrogate network (oracle/surrogate_ref.py) pinned by independent
checkers: torch CPU autograd (float64), central finite differences, and BF16 rounding cases."""
import numpy as np
import pytest

from oracle import surrogate_ref as sr


def _case(n=96, seed=3):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=0.3, size=sr.NPARAM)
    f = np.stack([rng.uniform(0.8, 4.0, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                  rng.uniform(-1, 1, n)], axis=1)
    t = rng.normal(scale=2.0, size=n)
    t[:5] = 50.0                                      # some targets above the clip
    return p, f, t


def test_forward_and_gradient_match_torch_autograd():
    torch = pytest.importorskip("torch")
    p, f, t = _case()
    clip = 4.0
    loss, g = sr.loss_and_grad(p, f, t, clip)
    P = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    W1 = P[0:256].reshape(64, 4); b1 = P[256:320]
    W2 = P[320:4416].reshape(64, 64); b2 = P[4416:4480]
    W3 = P[4480:4544]; b3 = P[4544]
    x = torch.tensor(f, dtype=torch.float64).clone()
    x[:, 0] = 0.5 * x[:, 0] - 1.0
    h1 = torch.nn.functional.silu(x @ W1.T + b1)
    h2 = torch.nn.functional.silu(h1 @ W2.T + b2)
    y = h2 @ W3 + b3
    tt = torch.clamp(torch.tensor(t, dtype=torch.float64), max=clip)
    L = torch.mean((y - tt) ** 2)
    L.backward()
    np.testing.assert_allclose(sr.forward(p, f), y.detach().numpy(), rtol=1e-12, atol=1e-12)
    assert loss == pytest.approx(L.item(), rel=1e-12)
    np.testing.assert_allclose(g, P.grad.numpy(), rtol=1e-9, atol=1e-12)


def test_gradient_matches_finite_differences():
    p, f, t = _case(n=40, seed=8)
    clip = 10.0
    _, g = sr.loss_and_grad(p, f, t, clip)
    rng = np.random.default_rng(1)
    idx = np.concatenate([[0, 255, 256, 319, 320, 4415, 4416, 4479, 4480, 4543, 4544],
                          rng.integers(0, sr.NPARAM, 20)])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (sr.loss_and_grad(pp, f, t, clip)[0] - sr.loss_and_grad(pm, f, t, clip)[0]) / (2 * h)
        assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-8), k


def test_bf16_rounding_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7                                   # BF16 spacing at 1
    vals = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 4, -2.5, 3.0e38, 1e-40])
    got = sr.bf16_round(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                              # tie -> even (1.0)
    assert got[2] == one + 2 * ulp                    # tie -> even (1 + 2 ulp)
    assert got[3] == 1.0
    assert got[4] == -2.5
    assert np.isfinite(got[5]) or got[5] == np.inf
    assert abs(got[6]) < 1e-38

"""Pins of the CPU network restatement (oracle/mlp.py) and of the host BF16
row encoder of the C ABI; no GPU needed."""
import numpy as np
import pytest
import torch

from oracle import mlp as om
from padenti_amd import proposal


def _small_problem(seed=3, n=6):
    rng = np.random.default_rng(seed)
    theta = om.init_params(seed, scale=1.5)
    x = rng.uniform(-1, 1, size=(n, om.IN))
    t = rng.uniform(0, 1, size=n)
    return theta, x, t


def test_bf16_rounding_known_cases():
    one = 1.0
    cases = {
        one: one,
        1.0 + 2 ** -8: 1.0,                   # halfway, ties to even (down)
        1.0 + 3 * 2 ** -8: 1.0 + 2 ** -6,     # halfway, ties to even (up)
        1.0 + 2 ** -8 + 2 ** -20: 1.0 + 2 ** -7,
        -2.5: -2.5,
        3.0e38: (1 + 98 / 128) * 2.0 ** 127,   # 1.7632 * 2^127: 7-bit fraction 97.69 -> 98
    }
    for x, want in cases.items():
        assert om.bf16(np.array([x]))[0] == want, x


def test_gradient_matches_finite_differences():
    theta, x, t = _small_problem()
    _, g = om.loss_and_grad(theta, x, t)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(om.NPARAM, 300, replace=False),
                          np.arange(om.NPARAM - 257, om.NPARAM)])   # w4, b4 in full
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        lp = om.loss_and_grad(tp, x, t)[0]
        lm = om.loss_and_grad(tm, x, t)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_forward_matches_torch():
    theta, x, _ = _small_problem(n=17)
    p = om.unpack(theta)
    with torch.no_grad():
        a = torch.tensor(x, dtype=torch.float64)
        for W, b in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
            a = torch.nn.functional.gelu(torch.nn.functional.linear(a, torch.tensor(p[W]), torch.tensor(p[b])),
                                         approximate="tanh")
        y = (a @ torch.tensor(p["w4"]) + p["b4"][0]).numpy()
    np.testing.assert_allclose(om.forward(theta, x)[0], y, rtol=1e-12, atol=1e-12)


def test_adam_matches_torch():
    rng = np.random.default_rng(5)
    theta = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(3)]
    w = torch.tensor(theta.copy(), requires_grad=True)
    opt = torch.optim.Adam([w], lr=3e-3, betas=(0.8, 0.99), eps=1e-7)
    th, m, v = theta.copy(), np.zeros(50), np.zeros(50)
    for k, g in enumerate(grads, start=1):
        w.grad = torch.tensor(g)
        opt.step()
        th, m, v = om.adam_step(th, m, v, g, k, lr=3e-3, beta1=0.8, beta2=0.99, eps=1e-7)
        np.testing.assert_allclose(th, w.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_bf16_mode_only_rounds():
    theta, x, t = _small_problem(n=32)
    y0 = om.forward(theta, x)[0]
    y1 = om.forward(theta, x, bf16_mode=True)[0]
    assert not np.array_equal(y0, y1)
    assert np.max(np.abs(y0 - y1)) < 0.05 * (np.max(np.abs(y0)) + 0.1)


def test_host_row_encoder_matches_oracle_rounding():
    rng = np.random.default_rng(9)
    x = rng.normal(size=(40, om.IN)).astype(np.float32)
    x[0, :4] = [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -0.0, 65504.0]
    rows = proposal.encode_rows(x)
    assert rows.shape == (40, 32) and not rows[:, om.IN:].any()
    np.testing.assert_array_equal(proposal.bf16_to_f32(rows[:, :om.IN]), om.bf16(x).astype(np.float32))


@pytest.mark.parametrize("name", ["W1", "b4"])
def test_param_layout(name):
    th = np.arange(om.NPARAM, dtype=np.float64)
    p = om.unpack(th)
    offs = {"W1": proposal.OFF_W1, "b4": proposal.OFF_B4}
    assert p[name].reshape(-1)[0] == offs[name]
    assert om.NPARAM == proposal.NPARAM

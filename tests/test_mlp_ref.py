"""Pins of the CPU MLP reference (oracle/mlp_ref.py) that checks the HIP inference and training
kernels of the learned restructure scheduler.

* bf16 rounding against torch's float32 -> bfloat16 conversion (round to nearest even);
* gradients against central finite differences of the exact float64 network;
* the bf16 forward equals the exact network when every weight, input and activation is
  representable in bf16 (so rounding is applied where the contract says and nowhere else);
* one Adam step against torch.optim.Adam.
"""
import numpy as np
import pytest
import torch

import mlp_ref


def test_bf16_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=4000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 4000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.005859375, 65504.0, 3.0e38, -3.0e38, 1e-40,
                                  np.inf, -np.inf], np.float32)]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    got = mlp_ref.bf16(x)
    assert np.array_equal(got, ref)
    assert np.isnan(mlp_ref.bf16(np.array([np.nan], np.float32))[0])


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = mlp_ref.init_params(3).astype(np.float64)
    p[mlp_ref.NIN * mlp_ref.NH:mlp_ref.NIN * mlp_ref.NH + mlp_ref.NH] = 0.05   # b1 > 0: fewer kinks
    x = rng.normal(size=(40, mlp_ref.NIN))
    t = rng.normal(size=40)
    _, g = mlp_ref.loss_and_grad(p, x, t, quant=False)
    idx = rng.choice(mlp_ref.NPARAM, 300, replace=False)
    eps = 1e-6
    num = np.empty(idx.size)
    for k, i in enumerate(idx):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = mlp_ref.loss_and_grad(pp, x, t, quant=False)
        lm, _ = mlp_ref.loss_and_grad(pm, x, t, quant=False)
        num[k] = (lp - lm) / (2 * eps)
    err = np.linalg.norm(num - g[idx]) / np.linalg.norm(num)
    assert err < 1e-5, err


def test_quantised_forward_is_exact_on_representable_values():
    rng = np.random.default_rng(2)
    # weights and inputs are small multiples of 1/8: every product and partial sum below is
    # exact in bf16 / fp32, so the bf16 network must equal the float64 one bit for bit
    d = {"W1": rng.integers(-4, 5, (mlp_ref.NH, mlp_ref.NIN)) / 8.0, "b1": rng.integers(-4, 5, mlp_ref.NH) / 8.0,
         "W2": rng.integers(-2, 3, (mlp_ref.NH, mlp_ref.NH)) / 8.0, "b2": rng.integers(-4, 5, mlp_ref.NH) / 8.0,
         "W3": rng.integers(-4, 5, mlp_ref.NH) / 8.0, "b3": np.array([0.25])}
    p = mlp_ref.pack(d)
    x = rng.integers(-4, 5, (64, mlp_ref.NIN)) / 4.0
    yq = mlp_ref.forward(p, x, quant=True)
    y = mlp_ref.forward(p, x, quant=False)
    # h1 values are multiples of 1/32 below 64, h1 @ W2 then needs <= 8 significant bits only if
    # small: check the premise and the equality
    _, c = mlp_ref.forward(p, x, keep=True, quant=False)
    assert np.array_equal(mlp_ref.bf16(np.maximum(c["z1"], 0)), np.maximum(c["z1"], 0))
    assert np.array_equal(yq, y)


def test_adam_step_matches_torch():
    rng = np.random.default_rng(4)
    p0 = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(3)]
    w = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([w], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for k, g in enumerate(grads, start=1):
        opt.zero_grad()
        w.grad = torch.tensor(g)
        opt.step()
        p, m, v = mlp_ref.adam_step(p, m, v, g, k, 1e-2)
    np.testing.assert_allclose(p, w.detach().numpy(), rtol=0, atol=1e-12)


@pytest.mark.parametrize("quant", [False, True])
def test_loss_is_mean_squared_error(quant):
    rng = np.random.default_rng(5)
    p = mlp_ref.init_params(5)
    x = rng.normal(size=(17, mlp_ref.NIN))
    t = rng.normal(size=17)
    loss, _ = mlp_ref.loss_and_grad(p, x, t, quant=quant)
    assert loss == pytest.approx(np.mean((mlp_ref.forward(p, x, quant=quant) - t) ** 2), rel=1e-12)

"""CPU pins of the learned-proposal oracle (oracle/mlp.py) by independent checks: the Adam step against
torch.optim.Adam (Kingma & Ba's update as torch implements it), the spectral features against a hand-computed
two-bin case (bin means, the interpolation between occupied bins, the end-bin extension, the N and C inputs),
and the forward pass against torch.nn layers of the same weights."""
from __future__ import annotations

import math

import numpy as np
import torch

from oracle import mlp as omlp

EDGES = np.logspace(-2.0, 3.0, omlp.NBIN + 1)  # the feature bins: 32 log-spaced eigenvalue bins over 1e-2 .. 1e3


def test_adam_step_matches_torch_optim_adam():
    rng = np.random.default_rng(0)
    n = 257
    theta0 = rng.normal(size=n)
    grads = [rng.normal(size=n) * s for s in (1.0, 0.3, 2.0, 0.01)]
    for lr, b1, b2, eps in ((1e-3, 0.9, 0.999, 1e-8), (3e-2, 0.8, 0.99, 1e-6)):
        p = torch.nn.Parameter(torch.from_numpy(theta0.copy()))
        opt = torch.optim.Adam([p], lr=lr, betas=(b1, b2), eps=eps)
        th, m, v = theta0.copy(), np.zeros(n), np.zeros(n)
        for t, g in enumerate(grads, start=1):
            p.grad = torch.from_numpy(g.copy())
            opt.step()
            th, m, v = omlp.adam_step(th, g, m, v, t, lr=lr, b1=b1, b2=b2, eps=eps)
            np.testing.assert_allclose(th, p.detach().numpy(), rtol=1e-12, atol=1e-15)
        st = opt.state[p]
        np.testing.assert_allclose(m, st["exp_avg"].numpy(), rtol=1e-13)
        np.testing.assert_allclose(v, st["exp_avg_sq"].numpy(), rtol=1e-13)


def test_features_hand_computed_two_bins():
    # two eigenvalues in bin 5 and two in bin 9 (edges 10^(-2 + 5 b / 32)); everything else empty
    centre = lambda b: 10.0 ** (-2.0 + 5.0 * (b + 0.5) / 32.0)  # noqa: E731
    s = np.array([centre(5), centre(5) * 1.01, centre(9), centre(9) * 1.01])
    bins = omlp.bins_of(s, EDGES)
    np.testing.assert_array_equal(bins, [5, 5, 9, 9])
    y = np.array([[1.0, 3.0], [1.0, 1.0], [2.0, 0.0], [2.0, 2.0]])  # two columns
    got = omlp.features(y, bins, n=1000.0, c=4.0)
    # column 0: y^2 = 1, 1, 4, 4 -> mean 2.5; bin 5 mean 1, bin 9 mean 4
    # column 1: y^2 = 9, 1, 0, 4 -> mean 3.5; bin 5 mean 5, bin 9 mean 2
    want = np.zeros((2, 34))
    for col, (m_all, m5, m9) in enumerate(((2.5, 1.0, 4.0), (3.5, 5.0, 2.0))):
        f5, f9 = math.log(m5 / m_all), math.log(m9 / m_all)
        for b in range(32):
            if b <= 5:
                want[col, b] = f5  # below the first occupied bin: its value
            elif b >= 9:
                want[col, b] = f9  # above the last occupied bin: its value
            else:
                x = (b - 5) / 4.0  # linear in the bin index between bins 5 and 9
                want[col, b] = (1 - x) * f5 + x * f9
        want[col, 32] = math.log(1000.0) / 10.0
        want[col, 33] = 0.4
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)


def test_forward_matches_torch_layers():
    rng = np.random.default_rng(3)
    theta = rng.normal(scale=0.2, size=omlp.n_params())
    x = rng.normal(size=(17, omlp.N_IN))
    W1, b1, W2, b2, w3, b3 = omlp.unpack(theta)
    net = torch.nn.Sequential(torch.nn.Linear(omlp.N_IN, omlp.H), torch.nn.Tanh(), torch.nn.Linear(omlp.H, omlp.H),
                              torch.nn.Tanh(), torch.nn.Linear(omlp.H, 1)).double()
    with torch.no_grad():
        for layer, (W, b) in zip((net[0], net[2], net[4]), ((W1, b1), (W2, b2), (w3, b3))):
            layer.weight.copy_(torch.from_numpy(W))
            layer.bias.copy_(torch.from_numpy(b))
        want = net(torch.from_numpy(x))[:, 0].numpy()
    got, _ = omlp.forward(theta, x)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)

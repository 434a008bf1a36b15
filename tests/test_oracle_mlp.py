"""The MLP oracle against central finite differences (CPU)."""
import numpy as np
import pytest


def test_mlp_gradient_matches_finite_differences(oracle):
    rng = np.random.default_rng(0)
    n_in, n_out, H = 8, 2, 64
    npar = n_in * H + H + H * H + H + H * n_out + n_out
    w = rng.normal(0, 0.3, size=npar)
    X = rng.normal(size=(37, n_in))
    Y = rng.normal(size=(37, n_out))
    loss, g = oracle.mlp_loss_grad(w, X, Y, n_in, n_out)
    h = 1e-6
    for i in rng.choice(npar, size=40, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (oracle.mlp_loss_grad(wp, X, Y, n_in, n_out)[0] -
              oracle.mlp_loss_grad(wm, X, Y, n_in, n_out)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-6, abs=1e-9)


def test_mlp_loss_zero_at_exact_fit(oracle):
    # With W3 = 0 the network outputs b3: targets equal to b3 give zero loss
    # and zero gradient.
    n_in, n_out, H = 3, 2, 64
    npar = n_in * H + H + H * H + H + H * n_out + n_out
    rng = np.random.default_rng(1)
    w = rng.normal(0, 0.3, size=npar)
    w[npar - n_out - H * n_out:npar - n_out] = 0.0
    w[-2:] = [0.5, -1.5]
    X = rng.normal(size=(10, n_in))
    Y = np.tile([0.5, -1.5], (10, 1))
    loss, g = oracle.mlp_loss_grad(w, X, Y, n_in, n_out)
    assert loss == 0.0
    assert np.all(g == 0.0)

"""Learned surrogate on the GPU (runs on an MI355X): kernel gradients against
the FP64 oracle, training, and the gain over full Monte Carlo on a fresh set
of roads at a measured error."""
import time

import numpy as np
import pytest

from transportoptimiser_amd import surrogate as sg

pytestmark = pytest.mark.gpu


def test_mlp_gradient_matches_oracle(ctx, oracle):
    net = sg.MlpEnsemble(ctx, members=3, seed=2)
    rng = np.random.default_rng(3)
    X = rng.normal(size=(200, 8)).astype(np.float32)
    Y = rng.normal(size=(200, 2)).astype(np.float32)
    loss, g = net.gradient(X, Y)
    p = net.get_params().reshape(3, -1)
    for e in range(3):
        lo, go = oracle.mlp_loss_grad(p[e].astype(np.float64), X, Y, 8, 2)
        assert loss[e] == pytest.approx(lo, rel=1e-5)
        np.testing.assert_allclose(g[e], go, rtol=1e-4, atol=1e-5 * np.abs(go).max())
    net.close()


def test_training_fits_a_smooth_map(ctx):
    rng = np.random.default_rng(4)
    X = rng.uniform(-1, 1, size=(4096, 8)).astype(np.float32)
    Y = np.stack([np.sin(2 * X[:, 0]) + X[:, 1] ** 2, X[:, 2] * X[:, 3]], axis=1)
    net = sg.MlpEnsemble(ctx, members=4, seed=5)
    l0 = net.gradient(X, Y)[0]
    t = time.perf_counter()
    l1 = net.train(X, Y, steps=3000, lr=3e-3)
    print(f"mlp training: 3000 full-batch steps on 4096 samples x 4 members in "
          f"{time.perf_counter() - t:.2f}s, loss {l0.mean():.3f} -> {l1.mean():.5f}")
    assert np.all(l1 < 0.01 * l0)
    pred = net.predict(X[:256]).mean(0)
    assert np.abs(pred - Y[:256]).mean() < 0.05
    net.close()


def test_surrogate_gain_over_full_monte_carlo(ctx):
    # Train on 4,096 simulator-evaluated roads, then evaluate 512 new roads
    # with the surrogate (falling back to full MC where the ensemble spread
    # of MTE exceeds 0.01) and with full MC alone.
    t = time.perf_counter()
    train = sg.make_roads(4096, seed=11)
    Y = sg.simulate(ctx, train)
    t_data = time.perf_counter() - t
    X = np.stack([r.features for r in train])
    sur = sg.Surrogate(ctx, members=4, seed=0, max_spread=0.01)
    loss = sur.fit(X, Y, steps=3000, lr=3e-3)
    test = sg.make_roads(512, seed=12)
    t = time.perf_counter()
    Y_full = sg.simulate(ctx, test)
    t_full = time.perf_counter() - t
    t = time.perf_counter()
    vals, fall = sur.evaluate(test)
    t_sur = time.perf_counter() - t
    err = np.abs(vals - Y_full)
    kept = ~fall
    print(f"surrogate: data {t_data:.1f}s, train loss {loss.mean():.4f}; 512 roads: full MC "
          f"{t_full:.2f}s, surrogate+fallback {t_sur:.2f}s ({t_full / t_sur:.1f}x), fallback "
          f"{fall.mean():.1%}, mean |dMTE| {err[:, 0].mean():.4f} (surrogate-answered "
          f"{err[kept, 0].mean() if kept.any() else 0:.4f}), mean |dROV|/1000 "
          f"{err[:, 1].mean():.4f}", flush=True)
    rel_rov = err[kept, 1] / np.abs(Y_full[kept, 1])
    print(f"surrogate-answered ROV: mean relative error {rel_rov.mean():.4f}")
    # The reference's own surrogate method on the same data: kNN local-linear
    # grids over the initial AAR (MTE, 1-D) and over (AAR, unit cost) (ROV,
    # 2-D, scaled coordinates, values <= 0), with the reference's neighbour
    # counts (SimulateGPU.cu:5533-5917).
    from transportoptimiser_amd import knn
    from transportoptimiser_amd import native as nat
    Xt = np.stack([r.features for r in test])
    g_mte = knn.fit(ctx, X[:, 1], Y[:, 0], 50, knn.neighbour_count(len(X), nat.KNN_RULE_MTE))
    g_rov = knn.fit(ctx, X[:, [1, 7]], Y[:, 1], 20,
                    knn.neighbour_count(len(X), nat.KNN_RULE_ROV), normalize=True,
                    nonpositive=True)
    knn_err = np.abs(np.stack([g_mte.predict(ctx, Xt[:, 1]), g_rov.predict(ctx, Xt[:, [1, 7]])],
                              1) - Y_full)
    mlp_err = np.abs(sur.predict(Xt)[0] - Y_full)
    print(f"reference kNN grid surrogate: mean |dMTE| {knn_err[:, 0].mean():.4f}, mean |dROV|/1000 "
          f"{knn_err[:, 1].mean():.4f}; MLP (no fallback): mean |dMTE| {mlp_err[:, 0].mean():.4f}, "
          f"mean |dROV|/1000 {mlp_err[:, 1].mean():.4f}", flush=True)
    # The gain at equal error: Monte Carlo on the first n paths of each road
    # (the same streams), the valuation on proportionally fewer paths; the
    # smallest n whose mean |dMTE| against the full run is within the
    # surrogate's sets the equal-error cost.
    import dataclasses
    sur_err = err[:, 0].mean()
    n_full = test[0].problem.n_paths
    equal = None
    for n in (6000, 4000, 3000, 2000, 1000):
        sub = [dataclasses.replace(r, problem=dataclasses.replace(r.problem, n_paths=n))
               for r in test]
        sg.simulate(ctx, sub[:8], rov_paths=100000 * n // n_full)
        t = time.perf_counter()
        Y_n = sg.simulate(ctx, sub, rov_paths=100000 * n // n_full)
        t_n = time.perf_counter() - t
        e_n = np.abs(Y_n[:, 0] - Y_full[:, 0]).mean()
        print(f"MC on {n} paths: {t_n:.3f}s, mean |dMTE| {e_n:.4f} against {n_full} paths",
              flush=True)
        if e_n <= sur_err:
            equal = (n, t_n)
    assert equal is not None
    print(f"equal error ({sur_err:.4f}): MC on {equal[0]} paths {equal[1]:.3f}s, surrogate "
          f"{t_sur:.3f}s ({equal[1] / t_sur:.1f}x)", flush=True)
    assert equal[1] / t_sur > 2.0
    assert mlp_err[:, 0].mean() < knn_err[:, 0].mean()
    assert np.all(err[fall] == 0.0)  # fallback roads are the full MC values
    assert fall.mean() < 0.3
    assert err[kept, 0].mean() < 0.01
    assert rel_rov.mean() < 0.05
    assert t_full / t_sur > 3.0


def test_surrogate_on_species_state_values(ctx):
    # The harder ROV target: the reference's species-state valuation (ALGO4,
    # 10,000 paths) of each road's own species, which depends on its
    # mortality (the validity thresholds) as well as its unit cost.  The
    # ensemble against the reference's own kNN grid surrogate over (loss
    # fraction, unit cost) on 256 fresh roads.
    from transportoptimiser_amd import knn
    from transportoptimiser_amd import native as nat
    t = time.perf_counter()
    train = sg.make_roads(2048, seed=13)
    Y = sg.simulate(ctx, train, rov_paths=10000, rov_model="species")
    t_data = time.perf_counter() - t
    X = np.stack([r.features for r in train])
    sur = sg.Surrogate(ctx, members=4, seed=0, max_spread=0.01)
    loss = sur.fit(X, Y, steps=3000, lr=3e-3)
    test = sg.make_roads(256, seed=14)
    Y_full = sg.simulate(ctx, test, rov_paths=10000, rov_model="species")
    Xt = np.stack([r.features for r in test])
    mlp = np.abs(sur.predict(Xt)[0] - Y_full)
    g_rov = knn.fit(ctx, X[:, [1, 7]], Y[:, 1], 20,
                    knn.neighbour_count(len(X), nat.KNN_RULE_ROV), normalize=True,
                    nonpositive=True)
    knn_rov = np.abs(g_rov.predict(ctx, Xt[:, [1, 7]]) - Y_full[:, 1])
    rel = mlp[:, 1] / np.abs(Y_full[:, 1])
    print(f"species-state ROV targets: data {t_data:.1f}s, train loss {loss.mean():.4f}; "
          f"values {Y_full[:, 1].min():.3f} .. {Y_full[:, 1].max():.3f} (x 1000); MLP mean |dROV| "
          f"{mlp[:, 1].mean():.4f} (relative {rel.mean():.3f}), kNN grid {knn_rov.mean():.4f}; "
          f"MLP mean |dMTE| {mlp[:, 0].mean():.4f}", flush=True)
    assert np.all(Y_full[:, 1] <= 0.0)
    assert mlp[:, 1].mean() < knn_rov.mean()
    assert mlp[:, 0].mean() < 0.01


"""What the trained surrogate gains over the search without it, on a small
brain family (4k neurons, 32 slots, 300 ticks): trained on simulator runs by
the project's own kernels, it predicts held-out outcomes better than the
mean, ranks held-out scores, and the proposals it sends to the simulator
score higher than the same number of plain random proposals. Every score
comes from the exact simulator."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_beats_plain_proposals(gpu):
    from axicor_amd import brain as B
    from axicor_amd import search as S
    from axicor_amd.surrogate import Mlp, Scaler
    wl = S.Workload(n_neurons=4000, slots=32, n_inputs=200, ticks=300, batch_brains=256, n_topologies=2,
                    n_conditions=16, seed=11)
    sim = S.Simulator(wl)
    cond = sim.add_condition(S.SEARCH_CONDITION)
    rng = np.random.default_rng(3)
    n = 2048
    P = B.random_params(rng, n)
    cj = rng.integers(0, wl.n_conditions, n)
    tj = rng.integers(0, wl.n_topologies, n)
    Y = sim.run(P, cj, tj)
    X = sim.features(P, cj)
    Xtr, Ytr, Xho, Yho = X[256:], Y[256:], X[:256], Y[:256]
    net = Mlp(16, 512, 3, 8, max_batch=65536, seed=5)
    xs, ys = Scaler(Xtr), Scaler(Ytr)
    losses = S.train(net, xs.fwd(Xtr), ys.fwd(Ytr), 1500, 1024, 1e-3, rng, lr_final=2e-5)
    pred = ys.inv(net.predict(xs.fwd(Xho)).astype(np.float64))
    mae = np.abs(pred - Yho).mean(axis=0)
    mae_mean = np.abs(Ytr.mean(axis=0) - Yho).mean(axis=0)
    rank = np.corrcoef(np.argsort(np.argsort(S.score_of(Yho))), np.argsort(np.argsort(S.score_of(pred))))[0, 1]
    print(f"loss {losses[0][1]:.3f} -> {losses[-1][1]:.4f}; holdout MAE {np.round(mae, 3)} "
          f"vs mean predictor {np.round(mae_mean, 3)}; score rank correlation {rank:.2f}")
    assert np.all(mae < 0.8 * mae_mean)
    assert rank > 0.5

    k = 256
    plain = S.score_of(sim.run(B.random_params(rng, k), cond))
    cand = B.random_params(rng, 1 << 16)
    pick = np.argsort(-S.score_of(ys.inv(net.predict(xs.fwd(sim.features(cand, cond))).astype(np.float64))))[:k]
    guided = S.score_of(sim.run(cand[pick], cond))
    print(f"plain: mean {plain.mean():.3f} best {plain.max():.3f}; surrogate-guided: mean {guided.mean():.3f} "
          f"best {guided.max():.3f}")
    assert guided.mean() > plain.mean() + 1.0
    assert np.median(guided) > np.percentile(plain, 75)
    assert guided.max() >= plain.max()
    net.destroy()
    sim.free()

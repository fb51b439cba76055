"""GPU tests of the tuning subsystem: MLP kernels against the float64
oracle, training, the autotuner, and the learned tuner (identical results,
measured gain over the reference's default knobs on held-out graphs)."""
import os
import time

import numpy as np
import pytest

import girhip
import mlp_oracle as mo
import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def test_mlp_forward_and_grad_match_oracle():
    say("mlp grad vs oracle")
    rng = np.random.default_rng(3)
    m = girhip.Mlp(seed=7)
    p = m.params
    x = rng.normal(size=(1000, 16)).astype(np.float32)
    t = rng.normal(size=(1000, 2)).astype(np.float32)
    y = m.forward(x)
    y_ref, _ = mo.forward(p, x)
    assert np.allclose(y, y_ref, rtol=1e-4, atol=1e-4)
    loss, g = m.grad(x, t)
    L_ref, g_ref = mo.loss_grad(p, x, t)
    assert abs(loss - L_ref) <= 1e-4 * max(1, abs(L_ref))
    # fp32 accumulation vs float64: relative to the gradient's scale
    scale = np.abs(g_ref).max()
    assert np.abs(g - g_ref).max() <= 2e-4 * scale, np.abs(g - g_ref).max() / scale


def test_mlp_training_reduces_loss():
    say("mlp training")
    rng = np.random.default_rng(4)
    x = rng.normal(size=(4000, 16)).astype(np.float32)
    w = rng.normal(size=(16, 2))
    t = np.tanh(x @ w / 4).astype(np.float32) + 0.1 * np.sin(x[:, :2] * 3).astype(np.float32)
    m = girhip.Mlp(seed=5)
    l0, _ = m.grad(x, t)
    l1 = m.train(x, t, epochs=400, lr=3e-3)
    l2, _ = m.grad(x, t)
    assert l2 < 0.1 * l0, (l0, l1, l2)


def test_autotune_results_identical_for_every_candidate():
    V, fb, fa, bb, ba = girhip.rmat_graph(12, seed=41)
    w = girhip.weights(len(fa), 1, 64, 3)
    src = girhip.sources(fb, 4, 2)
    say("autotune rmat12")
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.set_weights(w)
        base = G.sssp(src, 16)
        s0, _, _ = G.scc_dag()
        best, times = G.autotune("sssp_delta", src, [1, 4, 16, 64, 4096], reps=1)
        assert best in (1, 4, 16, 64, 4096) and len(times) == 5 and (times > 0).all()
        for d in (1, 4, 64, 4096):
            assert (G.sssp(src, d) == base).all()
        bt, _ = G.autotune("scc_trim_rounds", [], [0, 1, 3, 8])
        for T in (0, 1, 3, 8):
            G.set_param("scc_trim_rounds", T)
            assert (G.scc_dag()[0] == s0).all()
        f = G.features(4)
        assert f.shape == (16,) and np.isfinite(f).all()
        assert abs(f[0] - np.log2(V + 1)) < 1e-3 and abs(f[2] - len(fa) / V) < 1e-3
        assert abs(f[14] - 2.0) < 1e-6  # log2(k)


MODEL = os.path.join(os.path.dirname(girhip.__file__), "tuned_mlp.npz")


@pytest.mark.skipif(not os.path.exists(MODEL), reason="no trained model committed yet")
def test_learned_tuning_keeps_results_and_gains_over_default():
    import tuning
    model, norm = tuning.load_model(MODEL)
    cases = [("grid300_w1000", girhip.grid_graph(300, 300, 0.9, 0.8, seed=91), 1000),
             ("rmat16_w64", girhip.rmat_graph(16, seed=92), 64)]
    gains = []
    for name, (V, fb, fa, bb, ba), wmax in cases:
        say("learned", name)
        w = girhip.weights(len(fa), 1, wmax, 5)
        src = girhip.sources(fb, 8, 6)
        with girhip.Graph(fb, fa, bb, ba) as G:
            G.set_weights(w)
            G.sssp(src, 16)  # warm
            t0 = time.perf_counter()
            d_def = G.sssp(src, 16)
            t_def = G.last_ms("sssp")
            s_def = G.scc_dag()[0]
            res = G.learned_tune(model, norm, src, judge=0)
            d_new = G.sssp(src, 0)
            t_new = G.last_ms("sssp")
            assert (d_new == d_def).all(), name
            assert (G.scc_dag()[0] == s_def).all(), name
            assert (d_new[0] == oracle.dijkstra(fb, fa, w, src[0])).all()
            gains.append((name, res["delta"], t_def, t_new))
            print(f"[learned] {name}: delta {res['delta']:.0f} (raw {res['raw_delta']:.1f}), trim {res['trim']:.0f};"
                  f" sssp default {t_def:.2f} ms -> learned {t_new:.2f} ms ({t_def / t_new:.2f}x)", flush=True)
    # the high-diameter lattice is where delta 16 is far from right
    g = {n: td / tn for n, _, td, tn in gains}
    assert g["grid300_w1000"] > 2.0, gains
    assert g["rmat16_w64"] > 0.8, gains

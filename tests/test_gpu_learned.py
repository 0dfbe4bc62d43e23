"""GPU tests of the learned candidate filter: device gradients equal the FP64
oracle's, training lowers the loss, the trained filter keeps >= 99.9% of the
exact edges and needs far fewer exact evaluations than the untrained one."""
import numpy as np
import pytest

import mdet_amd as M
from mdet_amd import learned as L
from mdet_amd import workload
from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _edge_set(f, t):
    return set(zip(f.tolist(), t.tolist()))


def test_mlp_gradient_matches_fp64_oracle(ctx):
    w = workload.make(ctx, "c1_2k", n=400, seed=3)
    tnf = w.tnf.numpy()
    mean, istd = L.input_norm(tnf)
    mlp = M.Mlp(ctx, seed=5)
    mlp.set_input_norm(mean, istd)
    rng = np.random.default_rng(1)
    pi = rng.integers(0, 400, 96).astype(np.uint32)
    pj = rng.integers(0, 400, 96).astype(np.uint32)
    y = (rng.random(96) < 0.4).astype(np.float32)
    margin = 1e3  # every negative inside the margin: no relu kinks near the data
    d_pi, d_pj, d_y = ctx.to_device(pi), ctx.to_device(pj), ctx.to_device(y)
    loss, g = mlp.grad(w.tnf, d_pi, d_pj, d_y, 96, margin)
    p = mlp.params()
    L_ref, g_ref = MO.loss_and_grad(p.astype(np.float64), mean, istd, tnf, pi, pj, y, margin, round_weights=True)
    print(f"loss gpu {loss:.6f} oracle {L_ref:.6f}", flush=True)
    assert abs(loss - L_ref) <= 1e-4 * abs(L_ref)
    off = 0
    for name, shp in MO.SHAPES:
        k = int(np.prod(shp))
        a, b = g[off:off + k].astype(np.float64), g_ref[off:off + k]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"  {name}: rel err {rel:.2e}", flush=True)
        assert rel < 1e-4, name
        off += k
    mlp.close()


def test_learned_filter_trains_and_keeps_recall(ctx):
    w = workload.make(ctx, "c1_2k", seed=11)
    n = w.n
    f0, t0, s0, st0 = ctx.tnf_edges(w.tnf, w.logsize, n, 0.95, 200, M.BF16_REFINE)
    exact = _edge_set(f0, t0)
    pi, pj, y, pos = L.training_set(ctx, w.tnf, w.logsize, n, 0.95, 200, n_pairs=200_000, seed=1)
    mean, istd = L.input_norm(w.tnf.numpy())
    results = {}
    for name, steps in (("untrained", 0), ("trained", 600)):
        mlp = M.Mlp(ctx, seed=7)
        mlp.set_input_norm(mean, istd)
        curve, dt = L.train(ctx, mlp, w.tnf, pi, pj, y, steps, batch=512, lr=2e-3, margin=4.0, log_every=100)
        emb = L.embeddings(ctx, mlp, w.tnf, n)
        r2 = L.calibrate_radius2(emb, pos["i"], pos["j"], quantile=0.9995)
        f, t, s, st = ctx.tnf_edges(w.tnf, w.logsize, n, 0.95, 200, M.LEARNED, mlp=mlp, radius2=r2)
        got = _edge_set(f, t)
        recall = len(got & exact) / len(exact)
        results[name] = (recall, st.survivors)
        print(f"{name}: curve {curve} ({dt:.2f} s), radius2 {r2:.4g}, recall {recall:.5f}, "
              f"survivors {st.survivors} (exhaustive {st0.survivors})", flush=True)
        assert got <= exact  # proposals are decided exactly: no false edges
        if steps:
            assert curve[-1][1] < 0.5 * curve[0][1]
        mlp.close()
    assert results["trained"][0] >= 0.999
    # what training buys: the untrained network needs far more exact evaluations
    assert results["untrained"][1] >= 2 * results["trained"][1]

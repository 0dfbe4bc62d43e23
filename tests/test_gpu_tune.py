"""Launch shapes of the pair stage: every prefilter column span, per-item or
persistent prefilter blocks (one item queue or one per XCD) and refine grid
returns exactly the same edge list, and the autotuner keeps one of its
candidates."""
import numpy as np
import pytest

import mdet_amd as M
from mdet_amd import learned as L
from mdet_amd import workload

pytestmark = pytest.mark.gpu

DEFAULT = dict(seg_cols=0, persistent=1, refine_blocks=64, item_order=1)


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3]))


def test_every_gram_launch_shape_gives_the_same_edges(ctx):
    w = workload.make(ctx, "c2_50k", seed=4)
    try:
        for mode in (M.BF16_REFINE, M.F16_REFINE):
            ctx.tune_set(mode, **DEFAULT)
            base = ctx.tnf_edges(w.tnf, w.logsize, w.n, 0.95, 200, mode)
            print(f"mode {mode}: {len(base[0])} edges, {base[3].survivors} survivors", flush=True)
            assert len(base[0]) > 10000
            shapes = [dict(seg_cols=s, persistent=p, refine_blocks=32, item_order=o) for s in (256, 512, 1024, 4096, 8192)
                      for p, o in ((0, 0), (1, 0), (1, 1))] + \
                     [dict(seg_cols=0, persistent=0, refine_blocks=r, item_order=0) for r in (1, 8, 64)]
            for sh in shapes:
                ctx.tune_set(mode, **sh)
                got = ctx.tnf_edges(w.tnf, w.logsize, w.n, 0.95, 200, mode)
                print(f"  {sh}: prefilter {got[3].ms_prefilter:.3f} ms, refine {got[3].ms_refine:.3f} ms", flush=True)
                assert got[3].survivors == base[3].survivors
                assert _same(got, base), sh
    finally:
        for mode in (M.BF16_REFINE, M.F16_REFINE):
            ctx.tune_set(mode, **DEFAULT)
        for a in (w.tnf, w.logsize, w.counts):
            a.free()


def test_autotune_keeps_a_candidate_and_the_edges(ctx):
    w = workload.make(ctx, "c2_50k", seed=4)
    try:
        ctx.tune_set(M.BF16_REFINE, **DEFAULT)
        base = ctx.tnf_edges(w.tnf, w.logsize, w.n, 0.95, 200, M.BF16_REFINE)
        t = ctx.autotune(w.tnf, w.logsize, w.n, 0.95, 200, M.BF16_REFINE, reps=1)
        print(f"autotuned: {t}", flush=True)
        assert t["seg_cols"] in (512, 1024, 2048, 4096) and t["persistent"] in (0, 1) and t["item_order"] in (0, 1)
        assert t["refine_blocks"] in (8, 16, 32, 64)
        assert t["ms_prefilter"] > 0 and t["ms_refine"] > 0
        assert ctx.tune_get(M.BF16_REFINE) == t
        assert _same(ctx.tnf_edges(w.tnf, w.logsize, w.n, 0.95, 200, M.BF16_REFINE), base)
        # the other modes keep their own shapes
        assert ctx.tune_get(M.F16_REFINE)["ms_prefilter"] == 0
    finally:
        ctx.tune_set(M.BF16_REFINE, **DEFAULT)
        for a in (w.tnf, w.logsize, w.counts):
            a.free()


def test_learned_launch_shapes_give_the_same_edges(ctx):
    w = workload.make(ctx, "c1_2k", seed=5)
    n = w.n
    mean, istd = L.input_norm(w.tnf.numpy())
    mlp = M.Mlp(ctx, seed=3)
    mlp.set_input_norm(mean, istd)
    emb = L.embeddings(ctx, mlp, w.tnf, n)
    q = np.random.default_rng(0).integers(0, n, (20000, 2))
    r2 = float(np.quantile(((emb[q[:, 0]] - emb[q[:, 1]]) ** 2).sum(1), 0.05))
    try:
        ctx.tune_set(M.LEARNED, **DEFAULT)
        kw = dict(mlp=mlp, radius2=r2, n_probe=3, n_clusters=9)
        base = ctx.tnf_edges(w.tnf, w.logsize, n, 0.95, 200, M.LEARNED, **kw)
        for sh in (dict(seg_cols=1024, persistent=0, refine_blocks=32, item_order=0),
                   dict(seg_cols=0, persistent=1, refine_blocks=4, item_order=1)):
            ctx.tune_set(M.LEARNED, **sh)
            assert _same(ctx.tnf_edges(w.tnf, w.logsize, n, 0.95, 200, M.LEARNED, **kw), base), sh
        t = ctx.autotune(w.tnf, w.logsize, n, 0.95, 200, M.LEARNED, reps=1, **kw)
        print(f"learned autotuned: {t}", flush=True)
        assert t["persistent"] in (0, 1) and t["refine_blocks"] in (8, 16, 32, 64)
        assert _same(ctx.tnf_edges(w.tnf, w.logsize, n, 0.95, 200, M.LEARNED, **kw), base)
    finally:
        ctx.tune_set(M.LEARNED, **DEFAULT)
        mlp.close()
        for a in (w.tnf, w.logsize, w.counts):
            a.free()


def test_tune_set_rejects_bad_shapes(ctx):
    ctx.tune_set(M.BF16_REFINE, **DEFAULT)
    for bad in (dict(seg_cols=100), dict(seg_cols=1000), dict(seg_cols=1 << 17), dict(persistent=2),
                dict(refine_blocks=0), dict(refine_blocks=65), dict(item_order=2)):
        with pytest.raises(M.MdetError, match="mdet_tune_set"):
            ctx.tune_set(M.BF16_REFINE, **{**DEFAULT, **bad})
    with pytest.raises(M.MdetError, match="precision"):
        ctx.tune_set(7, **DEFAULT)
    assert ctx.tune_get(M.BF16_REFINE) == {**DEFAULT, "ms_prefilter": 0.0, "ms_refine": 0.0}

"""GPU parity of the pair stage: the FP64 tie-breaker equals the oracle bit
for bit, the FP32 evaluator stays far inside its tie band, and every precision
mode returns exactly the oracle's edge set."""
import math

import numpy as np
import pytest

import mdet_amd as M
from mdet_amd import workload
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EPS_PRE = 2e-3  # tie band of the FP32 evaluator (tnf_pairs.hip)
MODES = [M.EXACT_F32, M.BF16_REFINE, M.F16_REFINE]


def _mixed_lengths_tnf(ctx, n=600, seed=31):
    """Contigs across the whole length domain [1500, 500000] from 6 genomes."""
    cfg = M.gen_config(seed=seed, n_genomes=6, len_median=9000, len_sigma=1.6, len_min=1500, len_max=500000)
    lens, gid = M.gen_lengths(cfg, n)
    seqs = [M.gen_ascii(cfg, i, int(lens[i])) for i in range(n)]
    tnf = O.normalize_f32(np.stack([O.count_seq(s) for s in seqs]))
    ls = O.logsize(lens)
    return tnf, ls, gid


def test_preprob_fp64_bitexact_fp32_within_band(ctx):
    tnf, ls, gid = _mixed_lengths_tnf(ctx)
    n = tnf.shape[0]
    rng = np.random.default_rng(1)
    same = [(i, j) for i in range(n) for j in range(i + 1, n) if gid[i] == gid[j]]
    pi = np.array([p[0] for p in same] + list(rng.integers(0, n, 20000)), np.uint32)
    pj = np.array([p[1] for p in same] + list(rng.integers(0, n, 20000)), np.uint32)
    print(f"{len(pi)} pairs", flush=True)
    dt, dl = ctx.to_device(tnf), ctx.to_device(ls)
    p32, p64 = ctx.tnf_preprob(dt, dl, n, pi, pj)
    want = O.preprob_pairs(tnf, ls, pi, pj)
    assert np.array_equal(p64.view(np.uint64), want.view(np.uint64)), \
        f"{np.count_nonzero(p64 != want)} FP64 mismatches"
    err = np.abs(p32.astype(np.float64) - want)
    print(f"max |pre32 - pre64| = {err.max():.3e} over {len(pi)} pairs", flush=True)
    assert err.max() < EPS_PRE / 8
    dt.free()
    dl.free()


def _edge_set(f, t):
    return set(zip(f.tolist(), t.tolist()))


@pytest.mark.parametrize("cutoff,K", [(0.95, 200), (0.95, 5), (0.8, 30), (0.99, 0)])
def test_edges_all_modes_equal_oracle_2k(ctx, cutoff, K):
    w = workload.make(ctx, "c1_2k", seed=5)
    tnf, ls = w.tnf.numpy(), w.logsize.numpy()
    f0, t0, s0 = O.edges(tnf, ls, cutoff, K)
    want = _edge_set(f0, t0)
    print(f"2k contigs cutoff={cutoff} K={K}: oracle {len(want)} edges", flush=True)
    assert len(want) > 100
    for mode in MODES:
        f, t, s, st = ctx.tnf_edges(w.tnf, w.logsize, w.n, cutoff, K, mode)
        print(f"  mode {mode}: {len(f)} edges, stats {st.as_dict()}", flush=True)
        assert np.all(f < t)
        key = f.astype(np.uint64) << 32 | t
        assert np.all(np.diff(key.astype(np.int64)) > 0)  # sorted by (from, to), no duplicates
        assert _edge_set(f, t) == want
        # similarities: FP32 against the oracle's FP64
        order = np.lexsort((t0, f0))
        assert np.max(np.abs(s.astype(np.float64) - s0[order])) < 2e-6
    for a in (w.tnf, w.logsize, w.counts):
        a.free()


def test_edges_rows_of_50k_match_oracle(ctx):
    """50k contigs (1.25e9 pairs): every mode gives the same full edge list,
    and the edges of rows [0, 120) equal the oracle's (all 50k columns)."""
    w = workload.make(ctx, "c2_50k", seed=2)
    tnf, ls = w.tnf.numpy(), w.logsize.numpy()
    print(f"50k contigs, {w.n_bases} bases, count {w.count_ms:.1f} ms", flush=True)
    res = {}
    for mode in MODES:
        f, t, s, st = ctx.tnf_edges(w.tnf, w.logsize, w.n, 0.95, 200, mode)
        print(f"  mode {mode}: {len(f)} edges, stats {st.as_dict()}", flush=True)
        res[mode] = (f, t, s)
    for mode in MODES[1:]:
        assert np.array_equal(res[mode][0], res[0][0]) and np.array_equal(res[mode][1], res[0][1])
    f0, t0, s0 = O.edges(tnf, ls, 0.95, 200, rows=(0, 120))
    f, t, s = res[M.BF16_REFINE]
    m = f < 120
    assert _edge_set(f[m], t[m]) == _edge_set(f0, t0)
    for a in (w.tnf, w.logsize, w.counts):
        a.free()


def test_edges_reject_short_contigs_and_bad_cutoff(ctx):
    tnf = np.full((4, 136), 1 / math.sqrt(136), np.float32)
    ls = np.array([math.log10(1499), 4.0, 4.0, 4.0])
    dt, dl = ctx.to_device(tnf), ctx.to_device(ls)
    with pytest.raises(M.MdetError, match="1500"):
        ctx.tnf_edges(dt, dl, 4, 0.95, 10)
    with pytest.raises(M.MdetError, match="cutoff"):
        ctx.tnf_edges(dt, dl, 4, 1.0, 10)
    dt.free()
    dl.free()

"""Training and calibration of the learned candidate filter.

The exact path labels the data: one exhaustive edge-list call returns every
pair its prefilter passed to exact evaluation, flagged when preProb clears
the threshold (Ctx.last_pairs).  Positives are those pairs; negatives are
4 per positive, half "hard" (evaluated but below threshold: close in TNF
space) and half uniform random pairs (labelled against the positive set).
The network trains on the device (mdet_mlp_train_step); the embedding radius
is calibrated on the positives so that a chosen fraction of them fall inside.
"""
from __future__ import annotations

import time

import numpy as np

from . import BF16_REFINE, Ctx, DevArray, Mlp


def _pair_keys(rec: np.ndarray, n: int, seed: int):
    """Unordered pair key min * n + max and a splitmix64 hash of it."""
    i, j = rec["i"].astype(np.uint64), rec["j"].astype(np.uint64)
    key = np.minimum(i, j) * np.uint64(n) + np.maximum(i, j)
    z = key + np.uint64((0x9E3779B97F4A7C15 * (seed + 1)) & 0xFFFFFFFFFFFFFFFF)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return key, z ^ (z >> np.uint64(31))


def _sample_pairs(rec: np.ndarray, m: int, n: int, seed: int, rng) -> np.ndarray:
    """m records of rec chosen independently of rec's order (which follows the
    device's internal survivor order): the m smallest pair hashes, in key
    order; with replacement from the key-sorted records if m > len(rec)."""
    key, h = _pair_keys(rec, n, seed)
    if m >= len(rec):
        srt = rec[np.argsort(key, kind="stable")]
        return srt if m == len(rec) else srt[rng.choice(len(srt), m, replace=True)]
    take = np.argpartition(h, m - 1)[:m] if m else np.zeros(0, np.int64)
    return rec[take[np.argsort(key[take], kind="stable")]]


def training_set(ctx: Ctx, tnf: DevArray, logsize: DevArray, n: int, cutoff: float, max_edges: int,
                 n_pairs: int, seed: int = 0):
    """(pi, pj, y) with y in {0,1}, positives : negatives = 1 : 4."""
    ctx.tnf_edges(tnf, logsize, n, cutoff, max_edges, BF16_REFINE, fetch=False)
    rec = ctx.last_pairs()
    rng = np.random.default_rng(seed)
    is_pos = (rec["flags"] & 2) != 0
    pos, hard = rec[is_pos], rec[~is_pos]
    n_pos = min(len(pos), max(1, n_pairs // 5))
    P = _sample_pairs(pos, n_pos, n, seed, rng)
    n_hard = 2 * n_pos if len(hard) else 0
    H = _sample_pairs(hard, n_hard, n, seed + 1, rng) if n_hard else hard[:0]
    n_rand = 4 * n_pos - n_hard
    ri = rng.integers(0, n, n_rand).astype(np.uint64)
    rj = rng.integers(0, n, n_rand).astype(np.uint64)
    keep = ri != rj
    ri, rj = ri[keep], rj[keep]
    lo, hi = np.minimum(ri, rj), np.maximum(ri, rj)
    pkey = np.minimum(pos["i"], pos["j"]).astype(np.uint64) * n + np.maximum(pos["i"], pos["j"]).astype(np.uint64)
    ry = np.isin(lo * n + hi, pkey).astype(np.float32)
    pi = np.concatenate([P["i"], H["i"], ri.astype(np.uint32)])
    pj = np.concatenate([P["j"], H["j"], rj.astype(np.uint32)])
    y = np.concatenate([np.ones(len(P), np.float32), np.zeros(len(H), np.float32), ry])
    perm = rng.permutation(len(pi))
    return pi[perm].astype(np.uint32), pj[perm].astype(np.uint32), y[perm], pos


def input_norm(tnf_host: np.ndarray):
    x = np.where(np.isnan(tnf_host), 0.0, tnf_host).astype(np.float64)
    mean = x.mean(0)
    std = x.std(0)
    return mean.astype(np.float32), (1.0 / (std + 1e-6)).astype(np.float32)


def train(ctx: Ctx, mlp: Mlp, tnf: DevArray, pi: np.ndarray, pj: np.ndarray, y: np.ndarray, steps: int,
          batch: int = 512, lr: float = 1e-3, margin: float = 4.0, log_every: int = 0):
    """Adam steps over the uploaded training set (cycled); returns the loss curve
    sampled every log_every steps and the wall time."""
    d_pi, d_pj, d_y = ctx.to_device(pi), ctx.to_device(pj), ctx.to_device(y)
    m = len(pi)
    curve = []
    t0 = time.perf_counter()
    for s in range(steps):
        off = (s * batch) % max(1, m - batch + 1)
        want = bool(log_every) and (s % log_every == 0 or s == steps - 1)
        loss = mlp.train_step(tnf, d_pi, d_pj, d_y, min(batch, m), lr, margin, want_loss=want, offset=off)
        if want:
            curve.append((s, loss))
    ctx.sync()
    dt = time.perf_counter() - t0
    for a in (d_pi, d_pj, d_y):
        a.free()
    return curve, dt


def embeddings(ctx: Ctx, mlp: Mlp, tnf: DevArray, n: int) -> np.ndarray:
    emb = ctx.empty((n, 32), np.float32)
    mlp.embed(tnf, n, emb)
    out = emb.numpy()
    emb.free()
    return out


def knn_candidate_recall(emb: np.ndarray, exact_keys: np.ndarray, k: int, tile: int = 0,
                         rows: int = 2048, device: str = "cuda") -> dict:
    """Candidate pairs and edge recall of the per-row top-k form of the filter, for comparison with
    the cluster index: every contig keeps its k nearest embedding neighbours within each column
    tile of `tile` contigs (0: all columns), and a pair is a candidate if either endpoint kept the
    other.  Brute force through torch (on the GPU, torch has to be initialised before the library's
    context in the same process): a measurement of that alternative, not a product path.
    exact_keys = min * n + max of the exhaustive edge list."""
    import torch

    dev = torch.device(device)
    n = emb.shape[0]
    e = torch.from_numpy(np.ascontiguousarray(emb, dtype=np.float32)).to(dev)
    nrm = (e * e).sum(1)
    tile = tile or n
    keys = []
    for r0 in range(0, n, rows):
        r1 = min(n, r0 + rows)
        rid = torch.arange(r0, r1, device=dev)
        picks = []
        for c0 in range(0, n, tile):
            c1 = min(n, c0 + tile)
            d2 = nrm[r0:r1, None] + nrm[None, c0:c1] - 2.0 * (e[r0:r1] @ e[c0:c1].T)
            if c0 < r1 and r0 < c1:  # a contig is not its own neighbour
                d2 = d2.masked_fill(rid[:, None] == torch.arange(c0, c1, device=dev)[None, :], float("inf"))
            picks.append(torch.topk(d2, min(k, c1 - c0 - 1), dim=1, largest=False).indices + c0)
        idx = torch.cat(picks, 1)
        row = rid[:, None].expand_as(idx)
        keys.append((torch.minimum(row, idx) * n + torch.maximum(row, idx)).reshape(-1))
    cand = torch.unique(torch.cat(keys))
    del keys
    ex = torch.from_numpy(exact_keys.astype(np.int64)).to(dev)
    recall = float(torch.isin(ex, cand).float().mean()) if ex.numel() else 1.0
    return {"k": k, "tile": tile, "candidate_pairs": int(cand.numel()), "edge_recall": recall}


def calibrate_radius2(emb: np.ndarray, pi: np.ndarray, pj: np.ndarray, quantile: float = 0.9999,
                      slack: float = 1.05) -> float:
    """Squared embedding radius holding the given fraction of the positive pairs."""
    d = emb[pi.astype(np.int64)] - emb[pj.astype(np.int64)]
    D = (d * d).sum(1)
    return float(np.quantile(D, quantile) * slack)

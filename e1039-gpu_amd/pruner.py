"""Learned pruning of YZ station-3 units: weights, reference scoring, recall.

A YZ unit is one (u3, v3) window-list pair of a tracklet; the plain chain
evaluates every (u2, v2) pair under every unit (the i_uv3 loop of
reconstruction_kernels.cuh:1320-1800).  The device scores each unit with a
16-64-64-1 ReLU perceptron (bf16 operands on the matrix cores, fp32 sums) and
skips units scoring below a threshold (or_session_set_pruner).  This module
holds the weight layout, a torch fp32 restatement of the device arithmetic
(bf16 roundings included), the training routine and the recall measure used to
pick the threshold.  Features: or_session_set_unit_dump in include/e1039_or.h.
"""
from __future__ import annotations

import json
import os

import numpy as np

NF, NH = 16, 64
NW = NH * NF + NH + NH * NH + NH + NH + 1   # 5313 = OR_PRUNE_WEIGHTS
FEATURE_NAMES = ["x0/100", "10 tx", "st3 is D3m", "u3 hits - 1", "v3 hits - 1", "y_u3a/50", "y_u3b/50", "y_v3a/50",
                 "y_v3b/50", "ey_u3a", "ey_u3b", "ey_v3a", "ey_v3b", "(yu - yv)/10", "|yu - yv|/10", "(y_u3a - y_u3b)/10"]


def pack(w1, b1, w2, b2, w3, b3) -> np.ndarray:
    """Flat weight vector in the C ABI order (row-major, out x in)."""
    parts = [np.asarray(w1, np.float32).reshape(NH, NF), np.asarray(b1, np.float32).reshape(NH),
             np.asarray(w2, np.float32).reshape(NH, NH), np.asarray(b2, np.float32).reshape(NH),
             np.asarray(w3, np.float32).reshape(NH), np.asarray([b3], np.float32).reshape(1)]
    flat = np.concatenate([p.ravel() for p in parts]).astype(np.float32)
    assert flat.size == NW
    return flat


def unpack(flat: np.ndarray):
    f = np.asarray(flat, np.float32).ravel()
    if f.size != NW:
        raise ValueError(f"pruner weights: {f.size} values, expected {NW}")
    o = 0
    w1 = f[o:o + NH * NF].reshape(NH, NF); o += NH * NF
    b1 = f[o:o + NH]; o += NH
    w2 = f[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = f[o:o + NH]; o += NH
    w3 = f[o:o + NH]; o += NH
    return w1, b1, w2, b2, w3, float(f[o])


def bf16_round(x):
    """fp32 -> bf16 -> fp32, round to nearest even (the device conversion), torch tensors."""
    import torch
    u = x.contiguous().view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    u = torch.where(u >= 2**31, u - 2**32, u).to(torch.int32)
    return u.view(torch.float32)


def score_reference(flat: np.ndarray, feats):
    """torch fp32 restatement of the device scoring: bf16 weights, features and
    layer-1 activations, fp32 sums and output layer.  feats: (n, 16) tensor."""
    import torch
    w1, b1, w2, b2, w3, b3 = (torch.as_tensor(np.asarray(a, np.float32)) if not isinstance(a, float) else a
                              for a in unpack(flat))
    dev = feats.device
    w1, b1, w2, b2, w3 = (t.to(dev) for t in (w1, b1, w2, b2, w3))
    x = bf16_round(feats.float())
    h1 = torch.relu(x @ bf16_round(w1).T + b1)
    h2 = torch.relu(bf16_round(h1) @ bf16_round(w2).T + b2)
    return h2 @ w3 + b3


def save(path: str, flat: np.ndarray, threshold: float, meta: dict):
    np.savez(path, weights=np.asarray(flat, np.float32), threshold=np.float32(threshold),
             meta=np.array(json.dumps(meta)))


def load(path: str | None = None):
    """(flat weights, threshold, metadata) of a saved pruner; default: the shipped one."""
    if path is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "pruner_u3v3.npz")
    with np.load(path, allow_pickle=False) as z:
        flat = z["weights"].astype(np.float32)
        thr = float(z["threshold"])
        meta = json.loads(str(z["meta"]))
    if flat.size != NW:
        raise ValueError(f"{path}: {flat.size} weights, expected {NW}")
    return flat, thr, meta


def track_rows(res) -> np.ndarray:
    """Every output track as one opaque row: event index and the 96-byte record.
    Two chains output the same track when these rows are equal."""
    nt = res.n_tracks.astype(np.int64)
    valid = np.arange(res.tracks.shape[1])[None, :] < nt[:, None]
    ev = np.broadcast_to(np.arange(len(nt), dtype=np.int64)[:, None], valid.shape)[valid]
    rec = np.ascontiguousarray(res.tracks[valid]).view(np.uint8).reshape(-1, res.tracks.dtype.itemsize)
    rows = np.concatenate([ev.view(np.uint8).reshape(-1, 8), rec], axis=1)
    return np.ascontiguousarray(rows).view(np.dtype((np.void, rows.shape[1]))).ravel()


def recall(plain, pruned) -> tuple[float, int, int]:
    """Fraction of the plain chain's tracks that the pruned chain outputs identically."""
    a, b = track_rows(plain), track_rows(pruned)
    if a.size == 0:
        return 1.0, 0, 0
    found = int(np.intersect1d(a, b).size)
    return found / a.size, found, int(a.size)


def train(feats: np.ndarray, labels: np.ndarray, epochs: int = 12, seed: int = 7, device: str = "cpu",
          batch: int = 16384, log=print) -> np.ndarray:
    """Fits the perceptron (class-weighted logistic loss) and returns flat weights
    for raw features (the input standardisation is folded into layer 1)."""
    import torch
    g = torch.Generator().manual_seed(seed)
    torch.manual_seed(seed)
    x = torch.as_tensor(feats, dtype=torch.float32)
    y = torch.as_tensor(labels, dtype=torch.float32)
    mu, sd = x.mean(0), x.std(0) + 1e-3
    xn = ((x - mu) / sd).to(device)
    y = y.to(device)
    net = torch.nn.Sequential(torch.nn.Linear(NF, NH), torch.nn.ReLU(), torch.nn.Linear(NH, NH), torch.nn.ReLU(),
                              torch.nn.Linear(NH, 1)).to(device)
    opt = torch.optim.Adam(net.parameters(), lr=3e-3)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=epochs)
    pos = float(y.mean())
    pw = torch.tensor(min(100.0, (1 - pos) / max(pos, 1e-6)), device=device)
    n = xn.shape[0]
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(device)
        tot = 0.0
        for i in range(0, n, batch):
            b = perm[i:i + batch]
            loss = torch.nn.functional.binary_cross_entropy_with_logits(net(xn[b]).squeeze(1), y[b], pos_weight=pw)
            opt.zero_grad()
            loss.backward()
            opt.step()
            tot += float(loss) * b.numel()
        sched.step()
        log(f"epoch {ep}: loss {tot / n:.4f}")
    l1, l2, l3 = net[0], net[2], net[4]
    with torch.no_grad():
        w1 = (l1.weight.cpu() / sd).numpy()
        b1 = (l1.bias.cpu() - (l1.weight.cpu() * (mu / sd)).sum(1)).numpy()
        return pack(w1, b1, l2.weight.cpu().numpy(), l2.bias.cpu().numpy(), l3.weight.cpu().numpy().ravel(),
                    float(l3.bias.cpu()[0]))

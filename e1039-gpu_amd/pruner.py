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
THRESHOLD_OFF = -3.0e38   # absolute score threshold that keeps every unit (the margin decides)
NW = NH * NF + NH + NH * NH + NH + NH + 1   # 5313 = OR_PRUNE_WEIGHTS
FEATURE_NAMES = ["10 tx", "y3/50", "(yU3 - yV3)/10", "|yU3 - yV3|/10", "(yU3a - yU3b)/10", "(yV3a - yV3b)/10",
                 "U3 hits - 1", "V3 hits - 1", "U2 nearest distance/10", "V2 nearest distance/10",
                 "representative log(1 + chi2)/4", "representative stage reached (0-2)", "(nearest U2 - nearest V2)/10",
                 "P1 Y nearest distance/10", "P2 Y nearest distance/10", "representative |y0|/50"]


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


def score_reference(flat: np.ndarray, feats, chunk: int = 1 << 22):
    """torch fp32 restatement of the device scoring: bf16 weights, features and
    layer-1 activations, fp32 sums and output layer.  feats: (n, 16) tensor;
    evaluated in chunks of rows (keeps every intermediate tensor far below 2^31
    elements)."""
    import torch
    w1, b1, w2, b2, w3, b3 = (torch.as_tensor(np.asarray(a, np.float32)) if not isinstance(a, float) else a
                              for a in unpack(flat))
    dev = feats.device
    w1, b1, w2, b2, w3 = (t.to(dev) for t in (w1, b1, w2, b2, w3))
    w1r, w2r = bf16_round(w1).T, bf16_round(w2).T
    out = []
    for a in range(0, feats.shape[0], chunk):
        x = bf16_round(feats[a:a + chunk].float())
        h1 = torch.relu(x @ w1r + b1)
        h2 = torch.relu(bf16_round(h1) @ w2r + b2)
        out.append(h2 @ w3 + b3)
    return torch.cat(out) if out else torch.zeros(0, device=dev)


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


def train_rank(feats, group, winner, epochs: int = 12, seed: int = 7, device: str = "cuda", batch_rows: int = 1 << 22,
               tune_epochs: int = 4, tune_temp: float = 0.3, log=print) -> np.ndarray:
    """Fits the perceptron as a ranking of the units of each tracklet: a softmax
    over the units of a group (a tracklet) with the group's YZ winner as the
    target (listwise cross-entropy), then tune_epochs more at softmax
    temperature tune_temp, where the loss tends to the winner's distance below
    the best unit of its group (what the keep margin has to cover).  feats (n,
    16), group (n,) non-decreasing group ids (rows of a group contiguous),
    winner (n,) bool, at most one per group; groups without a winner do not
    contribute.  Returns flat weights for raw features (the input
    standardisation folded into layer 1)."""
    import torch
    torch.manual_seed(seed)
    x = torch.as_tensor(feats, device=device)
    g = torch.as_tensor(group, dtype=torch.int64, device=device)
    w = torch.as_tensor(winner, dtype=torch.bool, device=device)
    n = x.shape[0]
    mu = torch.zeros(NF, dtype=torch.float64, device=device)
    sq = torch.zeros(NF, dtype=torch.float64, device=device)
    for a in range(0, n, batch_rows):
        xb = x[a:a + batch_rows].double()
        mu += xb.sum(0)
        sq += (xb * xb).sum(0)
    mu = mu / n
    sd = (torch.sqrt(torch.clamp(sq / n - mu * mu, min=0)) + 1e-3).float()
    mu = mu.float()
    # batches of whole groups
    starts = torch.nonzero(torch.cat([torch.ones(1, dtype=torch.bool, device=device), g[1:] != g[:-1]])).squeeze(1)
    bounds = torch.cat([starts, torch.tensor([n], device=device)]).cpu().numpy()
    cuts = [0]
    for b in bounds[1:]:
        if b - cuts[-1] >= batch_rows:
            cuts.append(int(b))
    if cuts[-1] != n:
        cuts.append(n)
    net = torch.nn.Sequential(torch.nn.Linear(NF, NH), torch.nn.ReLU(), torch.nn.Linear(NH, NH), torch.nn.ReLU(),
                              torch.nn.Linear(NH, 1)).to(device)
    order = np.arange(len(cuts) - 1)
    rng = np.random.default_rng(seed)

    def run(nep, temp, lr):
        opt = torch.optim.Adam(net.parameters(), lr=lr)
        sched = torch.optim.lr_scheduler.OneCycleLR(opt, max_lr=lr, total_steps=max(nep * (len(cuts) - 1), 1))
        for ep in range(nep):
            rng.shuffle(order)
            tot, cnt = 0.0, 0
            for it, k in enumerate(order):
                if it % 50 == 49:
                    log(f"  epoch {ep}: {it + 1} of {len(order)} batches")
                a, b = cuts[k], cuts[k + 1]
                wb = w[a:b]
                gb = g[a:b] - g[a]
                ng = int(gb[-1]) + 1
                s = net((x[a:b].float() - mu) / sd).squeeze(1) / temp
                top = s.max().detach()            # one shift for the batch (no per-group reduction)
                z = torch.zeros(ng, device=device).index_add(0, gb, torch.exp(s - top))
                lse = torch.log(z) + top
                loss = temp * (lse[gb[wb]] - s[wb]).mean()
                opt.zero_grad()
                loss.backward()
                opt.step()
                sched.step()
                tot += float(loss.detach()) * int(wb.sum())
                cnt += int(wb.sum())
            log(f"epoch {ep} (temperature {temp}): loss {tot / max(cnt, 1):.4f}")

    run(epochs, 1.0, 3e-3)
    if tune_epochs > 0:
        run(tune_epochs, tune_temp, 1e-3)
    l1, l2, l3 = net[0], net[2], net[4]
    with torch.no_grad():
        w1 = (l1.weight.cpu() / sd.cpu()).numpy()
        b1 = (l1.bias.cpu() - (l1.weight.cpu() * (mu.cpu() / sd.cpu())).sum(1)).numpy()
        return pack(w1, b1, l2.weight.cpu().numpy(), l2.bias.cpu().numpy(), l3.weight.cpu().numpy().ravel(),
                    float(l3.bias.cpu()[0]))


def margin_gaps(scores, group, winner):
    """Per group with a winner: best score of the group minus the winner's score
    (>= 0; the margin a unit must be within for the winner to be kept), and the
    best score of every group.  Computed on the host."""
    import torch
    s = torch.as_tensor(scores).float().cpu()
    g = torch.as_tensor(group, dtype=torch.int64).cpu()
    w = torch.as_tensor(winner, dtype=torch.bool).cpu()
    ng = int(g.max()) + 1 if g.numel() else 0
    m = torch.full((ng,), -3.0e38).scatter_reduce(0, g, s, "amax", include_self=True)
    return (m[g[w]] - s[w]), m

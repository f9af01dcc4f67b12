"""Oracle of the learned candidate ranker (test infrastructure only).

numpy restatement of the ranker's mathematics: the candidate grid, the point
features, the MLP forward pass, the masked softmax cross-entropy and its
gradient, and one Adam step. The ranker has no counterpart in the reference; this
module is pinned by the mathematics instead: the backward pass against central
finite differences of the forward pass in float64, the grid against brute force
(every polygon that contains a point is listed in the point's cell). Only
``tests/`` and ``bench.py`` import it, as the checker.

``bf16=True`` rounds operands and stored activations to bfloat16 (round to
nearest even) where the GPU kernels store bf16, and accumulates in float64, so
the GPU results (fp32 accumulation on MFMA) agree within a small tolerance.
"""

from __future__ import annotations

import numpy as np

CAND = 16
EMB = 64
FOUR = 32
IN = 96
HID = 128
FREQS = 8


def layout(cells: int) -> dict:
    """Element offsets of the flat parameter array."""
    o = {"e": 0, "w1": cells * EMB}
    o["b1"] = o["w1"] + HID * IN
    o["w2"] = o["b1"] + HID
    o["b2"] = o["w2"] + HID * HID
    o["w3"] = o["b2"] + HID
    o["b3"] = o["w3"] + CAND * HID
    o["total"] = o["b3"] + CAND
    return o


def unpack(p: np.ndarray, cells: int) -> dict:
    o = layout(cells)
    return {"E": p[o["e"]:o["w1"]].reshape(cells, EMB),
            "W1": p[o["w1"]:o["b1"]].reshape(HID, IN), "b1": p[o["b1"]:o["w2"]],
            "W2": p[o["w2"]:o["b2"]].reshape(HID, HID), "b2": p[o["b2"]:o["w3"]],
            "W3": p[o["w3"]:o["b3"]].reshape(CAND, HID), "b3": p[o["b3"]:o["total"]]}


def pack(d: dict, cells: int) -> np.ndarray:
    o = layout(cells)
    p = np.zeros(o["total"], np.float64)
    for k, (a, b) in {"E": ("e", "w1"), "W1": ("w1", "b1"), "b1": ("b1", "w2"),
                      "W2": ("w2", "b2"), "b2": ("b2", "w3"), "W3": ("w3", "b3"),
                      "b3": ("b3", "total")}.items():
        p[o[a]:o[b]] = np.asarray(d[k]).reshape(-1)
    return p


def bf16_round(a) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    f = np.ascontiguousarray(np.asarray(a, np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32).reshape(f.shape)


def bf16_bits(a) -> np.ndarray:
    return (bf16_round(a).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(h) -> np.ndarray:
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


# ---------------------------------------------------------------------------
# Grid
# ---------------------------------------------------------------------------
def grid_coord(v, v0: float, cell: float, dim: int) -> np.ndarray:
    """floor((v - v0) * (1 / cell)) clamped to [0, dim - 1], NaN -> 0."""
    t = np.floor((np.asarray(v, np.float64) - v0) * (1.0 / cell))
    t = np.where(t >= 0.0, t, 0.0)
    return np.minimum(t, dim - 1).astype(np.int64)


def point_cells(x, y, x0, y0, cell, dim) -> np.ndarray:
    return grid_coord(y, y0, cell, dim) * dim + grid_coord(x, x0, cell, dim)


def local_uv(x, y, x0, y0, cell, dim):
    ix = grid_coord(x, x0, cell, dim)
    iy = grid_coord(y, y0, cell, dim)
    u = (np.asarray(x, np.float64) - (x0 + ix.astype(np.float64) * cell)) * (1.0 / cell)
    v = (np.asarray(y, np.float64) - (y0 + iy.astype(np.float64) * cell)) * (1.0 / cell)
    return u.astype(np.float32), v.astype(np.float32)


def fourier(u, v) -> np.ndarray:
    """[n][32]: for k = 1..8: sin(pi k u), cos(pi k u), sin(pi k v), cos(pi k v)."""
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    cols = []
    for k in range(FREQS):
        w = np.float32(np.pi) * np.float32(k + 1)
        wu = (w * u.astype(np.float32)).astype(np.float64)
        wv = (w * v.astype(np.float32)).astype(np.float64)
        cols += [np.sin(wu), np.cos(wu), np.sin(wv), np.cos(wv)]
    return np.stack(cols, 1)


# ---------------------------------------------------------------------------
# Forward / backward
# ---------------------------------------------------------------------------
def _r(a, bf16):
    return bf16_round(a).astype(np.float64) if bf16 else np.asarray(a, np.float64)


def forward(params: np.ndarray, cells: int, X: np.ndarray, bf16: bool = True):
    """X: [n][96] features (already bf16 when bf16). Returns dict of activations."""
    P = unpack(np.asarray(params, np.float64), cells)
    W1, W2, W3 = (_r(P[k], bf16) for k in ("W1", "W2", "W3"))
    Xf = np.asarray(X, np.float64)
    h1 = _r(np.maximum(Xf @ W1.T + P["b1"], 0.0), bf16)
    h2 = _r(np.maximum(h1 @ W2.T + P["b2"], 0.0), bf16)
    z = h2 @ W3.T + P["b3"]
    return {"X": Xf, "H1": h1, "H2": h2, "Z": z, "W1": W1, "W2": W2, "W3": W3}


def features(params, cells_n, x, y, x0, y0, cell, dim, bf16=True):
    P = unpack(np.asarray(params, np.float64), cells_n)
    c = point_cells(x, y, x0, y0, cell, dim)
    u, v = local_uv(x, y, x0, y0, cell, dim)
    E = _r(P["E"], bf16)
    X = np.concatenate([E[c], fourier(u, v)], 1)
    return (_r(X, bf16) if bf16 else X), c


def loss_and_grad(params, cells_n, X, point_cell, counts, labels, bf16=True):
    """Mean (over rows) masked softmax cross-entropy and its gradient (flat)."""
    n = X.shape[0]
    A = forward(params, cells_n, X, bf16)
    Z = A["Z"]
    cnt = np.minimum(np.asarray(counts)[point_cell], CAND)
    lab = np.asarray(labels)
    valid = (lab >= 0) & (lab < cnt)
    mask = np.arange(CAND)[None, :] < cnt[:, None]
    zm = np.where(mask, Z, -np.inf)
    mx = np.max(np.where(mask, Z, -np.inf), 1, keepdims=True)
    mx = np.where(np.isfinite(mx), mx, 0.0)
    e = np.where(mask, np.exp(zm - mx), 0.0)
    s = e.sum(1, keepdims=True)
    p = np.where(s > 0, e / np.where(s > 0, s, 1.0), 0.0)
    onehot = np.zeros_like(p)
    onehot[np.arange(n)[valid], lab[valid]] = 1.0
    li = np.where(valid, np.log(np.where(s[:, 0] > 0, s[:, 0], 1.0)) -
                  (Z[np.arange(n), np.clip(lab, 0, CAND - 1)] - mx[:, 0]), 0.0)
    loss = li.sum() / n
    dZ = np.where(valid[:, None], (p - onehot) / n, 0.0)
    dZ = _r(dZ, bf16)
    g = {}
    g["W3"] = dZ.T @ A["H2"]
    g["b3"] = dZ.sum(0)
    dH2 = _r((dZ @ A["W3"]) * (A["H2"] > 0), bf16)
    g["W2"] = dH2.T @ A["H1"]
    g["b2"] = dH2.sum(0)
    dH1 = _r((dH2 @ A["W2"]) * (A["H1"] > 0), bf16)
    g["W1"] = dH1.T @ A["X"]
    g["b1"] = dH1.sum(0)
    dX = dH1 @ A["W1"]
    gE = np.zeros((cells_n, EMB))
    np.add.at(gE, np.asarray(point_cell), dX[:, :EMB])
    g["E"] = gE
    correct = np.sum(valid & (np.argmax(zm, 1) == lab))
    return loss, pack(g, cells_n), {"loss_sum": float(li.sum()), "correct": int(correct),
                                    "labelled": int(valid.sum()), "dX": dX}


def adam(p, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    p = np.asarray(p, np.float64)
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


# ---------------------------------------------------------------------------
# Candidate grid (the brute-force statement of what the lists must hold)
# ---------------------------------------------------------------------------
def polygon_boxes(part, ring, px, py):
    part = np.asarray(part, np.int64)
    ring = np.asarray(ring, np.int64)
    boxes = np.full((len(part) - 1, 4), np.nan)
    for i in range(len(part) - 1):
        a, b = ring[part[i]], ring[part[i + 1]]
        if b > a:
            boxes[i] = [px[a:b].min(), py[a:b].min(), px[a:b].max(), py[a:b].max()]
    return boxes


def candidate_lists(part, ring, px, py, x0, y0, cell, dim):
    """{cell: sorted polygon ids whose box reaches the cell} (brute force)."""
    out = {}
    for i, (mnx, mny, mxx, mxy) in enumerate(polygon_boxes(part, ring, px, py)):
        if np.isnan(mnx):
            continue
        for iy in range(int(grid_coord(mny, y0, cell, dim)), int(grid_coord(mxy, y0, cell, dim)) + 1):
            for ix in range(int(grid_coord(mnx, x0, cell, dim)),
                            int(grid_coord(mxx, x0, cell, dim)) + 1):
                out.setdefault(iy * dim + ix, []).append(i)
    return out

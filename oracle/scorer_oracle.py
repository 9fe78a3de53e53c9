"""CPU restatement of the candidate scorer, FP64 (test infrastructure only:
the checker of the GPU forward and backward passes, never a product path).

The scorer has no gaffa counterpart.  Network, per candidate:
  x0 = cube [32 subint][64 phase] read as 32 input channels over phase;
  conv_l (l = 1..4): y[p][co] = relu(b[co] + sum_{t,ci} W[co][t][ci] x[(p + t - 1) mod 64][ci]),
                      channels 32 -> 64 -> 64 -> 64 -> 64;
  f  = [mean_p y4 (64), max_p y4 (64, first phase on ties), patch (128)];
  h  = relu(W5 f + b5) (512);  z = W6 h + b6;
  loss = mean_b softplus(z) - y z (binary cross-entropy with logits).

`emulate=True` rounds to bfloat16 (round to nearest even, via float32) at
every point the GPU stores a bfloat16 value (weights, the input cube and
patch, each layer's output, the pooled features, and the backward signals
fed to the MFMA GEMMs), so that the GPU differs from it only by FP32
accumulation order.  `emulate=False` is the plain FP64 network, pinned by
finite differences and by torch autograd (tests/test_scorer.py).
"""

from __future__ import annotations

import numpy as np

SUBINTS, PHASE, PATCH, CHANNELS, HIDDEN = 32, 64, 128, 64, 512
CIN = (SUBINTS, CHANNELS, CHANNELS, CHANNELS)
FEATURES = 2 * CHANNELS + PATCH


def layout() -> dict:
    """Offsets of each layer's weight and bias in the parameter vector."""
    rows = [CHANNELS] * 4 + [HIDDEN, 1]
    cols = [3 * c for c in CIN] + [FEATURES, HIDDEN]
    weight, bias, o = [], [], 0
    for r, c in zip(rows, cols):
        weight.append(o)
        o += r * c
        bias.append(o)
        o += r
    return {"weight": weight, "bias": bias, "fan_in": cols, "nparams": (o + 7) // 8 * 8}


def bf16(x) -> np.ndarray:
    """float64 -> float32 -> bfloat16 (nearest even) -> float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def f32(x) -> np.ndarray:
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def _unpack(params, lay):
    p = np.asarray(params, dtype=np.float64)
    ws, bs = [], []
    for i in range(6):
        w = p[lay["weight"][i]:lay["bias"][i]]
        nb = CHANNELS if i < 4 else (HIDDEN if i == 4 else 1)
        ws.append(w)
        bs.append(p[lay["bias"][i]:lay["bias"][i] + nb])
    return ws, bs


def _im2col(x):
    """x [B][64][C] -> [B][64][3C], tap t reads phase p + t - 1."""
    return np.concatenate([np.roll(x, 1, axis=1), x, np.roll(x, -1, axis=1)], axis=2)


def forward(params, cubes, patches, *, emulate: bool = True):
    q = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    lay = layout()
    ws, bs = _unpack(params, lay)
    B = cubes.shape[0]
    x = q(np.asarray(cubes, np.float64).transpose(0, 2, 1))  # [B][64][32]
    cache = {"xcol": [], "y": [], "q": q, "ws": ws, "bs": bs}
    for l in range(4):
        xcol = _im2col(x)
        w = q(ws[l]).reshape(CHANNELS, 3 * CIN[l])
        pre = xcol @ w.T + bs[l]
        y = np.maximum(pre, 0.0)
        y = q(y) if emulate else y
        cache["xcol"].append(xcol)
        cache["y"].append(y)
        x = y
    mean = x.mean(axis=1)
    if emulate:
        mean = q(f32(mean))
    mx = x.max(axis=1)
    cache["argmax"] = x.argmax(axis=1)  # first maximum
    f = np.concatenate([mean, mx, q(np.asarray(patches, np.float64))], axis=1)
    w5 = q(ws[4]).reshape(HIDDEN, FEATURES)
    h = np.maximum(f @ w5.T + bs[4], 0.0)
    h = q(h) if emulate else h
    w6 = q(ws[5])
    z = h @ w6 + bs[5][0]
    if emulate:
        z = f32(z)
    cache.update(f=f, h=h, z=z, B=B)
    return z, cache


def loss_and_gradients(params, cubes, patches, labels, *, emulate: bool = True):
    z, c = forward(params, cubes, patches, emulate=emulate)
    q, ws, B = c["q"], c["ws"], c["B"]
    y = np.asarray(labels, np.float64)
    lay = layout()
    g = np.zeros(lay["nparams"], np.float64)
    loss = np.mean(np.maximum(z, 0.0) - z * y + np.log1p(np.exp(-np.abs(z))))
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / B
    dzq = q(dz)
    g[lay["bias"][5]] = dz.sum()
    h, f = c["h"], c["f"]
    g[lay["weight"][5]:lay["bias"][5]] = dzq @ h
    w6 = q(ws[5])
    dh = np.where(h > 0, dzq[:, None] * w6[None, :], 0.0)
    dhq = q(dh)
    g[lay["bias"][4]:lay["bias"][4] + HIDDEN] = dh.sum(axis=0)
    g[lay["weight"][4]:lay["bias"][4]] = (dhq.T @ f).ravel()
    w5 = q(ws[4]).reshape(HIDDEN, FEATURES)
    df = dhq @ w5[:, :2 * CHANNELS]
    if emulate:
        df = f32(df)
    y4 = c["y"][3]
    dy = np.repeat(df[:, None, :CHANNELS] / PHASE, PHASE, axis=1)
    onehot = np.arange(PHASE)[None, :, None] == c["argmax"][:, None, :]
    dy = dy + onehot * df[:, None, CHANNELS:]
    dy = np.where(y4 > 0, dy, 0.0)
    for l in range(3, -1, -1):
        dyq = q(dy)
        cin = CIN[l]
        xcol = c["xcol"][l]
        gw = dyq.reshape(-1, CHANNELS).T @ xcol.reshape(-1, 3 * cin)
        g[lay["weight"][l]:lay["bias"][l]] = gw.ravel()
        g[lay["bias"][l]:lay["bias"][l] + CHANNELS] = dy.sum(axis=(0, 1))
        if l == 0:
            break
        w = q(ws[l]).reshape(CHANNELS, 3 * cin)
        dxcol = dyq @ w  # [B][64][3 cin]
        if emulate:
            dxcol = f32(dxcol)
        dx = (np.roll(dxcol[:, :, :cin], -1, axis=1) + dxcol[:, :, cin:2 * cin]
              + np.roll(dxcol[:, :, 2 * cin:], 1, axis=1))
        dy = np.where(c["y"][l - 1] > 0, dx, 0.0)
    return loss, g

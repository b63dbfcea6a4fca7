"""CPU oracle of the learned SolidLiquid classifier (test infrastructure only:
used by tests/ and bench.py's CPU baseline, never by the product path).

* ``forward`` / ``gradients``: the 11 -> 64 -> 64 -> 1 ReLU network with the
  GPU kernels' numerics (inputs standardised and rounded to BF16, weights
  rounded to BF16 for the two matrix layers, hidden activations of layer 1
  rounded to BF16 before layer 2, FP32 accumulation), its BCE-with-logits loss
  and the analytic gradients of that loss.
* ``features``: q_l and normalised w_l for l = 4, 6, 8, 10, 12 and the bond
  count, from the oracle's Steinhardt restatement (Steinhardt.cc:120-222,
  291-359).
"""

import numpy as np

N_IN, N_USED, N_H = 16, 11, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_H * N_IN
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_H
N_PARAMS = OFF_B3 + 1


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(params, dtype=np.float32):
    p = np.asarray(params, dtype)
    return dict(
        W1=p[OFF_W1:OFF_B1].reshape(N_H, N_IN), b1=p[OFF_B1:OFF_W2], W2=p[OFF_W2:OFF_B2].reshape(N_H, N_H),
        b2=p[OFF_B2:OFF_W3], w3=p[OFF_W3:OFF_B3], b3=p[OFF_B3])


def standardise(feat, mu, sigma, exact=False):
    if exact:
        x = np.zeros((len(feat), N_IN), np.float64)
        x[:, :N_USED] = (np.asarray(feat, np.float64)[:, :N_USED] - np.asarray(mu, np.float64)) / np.asarray(
            sigma, np.float64)
        return x
    x = np.zeros((len(feat), N_IN), np.float32)
    isig = (np.float32(1.0) / np.asarray(sigma, np.float32)).astype(np.float32)
    x[:, :N_USED] = (np.asarray(feat, np.float32)[:, :N_USED] - np.asarray(mu, np.float32)) * isig
    return bf16(x)


def forward(params, feat, mu, sigma, exact=False):
    """Logits and the intermediate values the backward pass needs.  With
    exact=True: float64 without BF16 rounding (the mathematical network, for
    finite-difference checks of the gradient formulas)."""
    P = unpack(params, np.float64 if exact else np.float32)
    rnd = (lambda a: np.asarray(a, np.float64)) if exact else bf16
    dt = np.float64 if exact else np.float32
    xb = standardise(feat, mu, sigma, exact)
    z1 = (xb @ rnd(P["W1"]).T + P["b1"]).astype(dt)
    h1b = rnd(np.maximum(z1, 0))
    z2 = (h1b @ rnd(P["W2"]).T + P["b2"]).astype(dt)
    h2 = np.maximum(z2, 0)
    y = (h2 @ P["w3"] + P["b3"]).astype(dt)
    return y, dict(xb=xb, z1=z1, h1b=h1b, z2=z2, h2=h2)


def loss(y, t):
    y = np.asarray(y, np.float64)
    t = np.asarray(t, np.float64)
    return float(np.mean(np.maximum(y, 0) - t * y + np.log1p(np.exp(-np.abs(y)))))


def gradients(params, feat, labels, mu, sigma, exact=False):
    """Gradient of the mean BCE-with-logits loss in the GPU parameter layout
    (the backward pass uses the BF16-rounded W2 and layer-1 activations that
    the forward pass used)."""
    dt = np.float64 if exact else np.float32
    P = unpack(params, dt)
    y, c = forward(params, feat, mu, sigma, exact)
    t = np.asarray(labels, dt)
    n = len(y)
    dy = ((1.0 / (1.0 + np.exp(-y.astype(np.float64))) - t) / n).astype(dt)
    g = np.zeros(N_PARAMS, dt)
    dz2 = np.where(c["z2"] > 0, dy[:, None] * P["w3"][None, :], 0).astype(dt)
    g[OFF_W3:OFF_B3] = dy @ c["h2"]
    g[OFF_B3] = dy.sum()
    g[OFF_W2:OFF_B2] = (dz2.T @ c["h1b"]).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dh1 = dz2 @ (P["W2"] if exact else bf16(P["W2"]))
    dz1 = np.where(c["z1"] > 0, dh1, 0).astype(dt)
    g[OFF_W1:OFF_B1] = (dz1.T @ c["xb"]).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return g, loss(y, t)


def features(box, points, nlist, nthreads=1):
    """(N, 12): q4..q12, normalised w4..w12, bond count, 0."""
    import oracle

    ls = [4, 6, 8, 10, 12]
    q = oracle.steinhardt(box, points, nlist, ls, nthreads=nthreads)
    w = oracle.steinhardt(box, points, nlist, ls, wl=True, wl_normalize=True, nthreads=nthreads)
    n = len(points)
    out = np.zeros((n, 12), np.float32)
    out[:, 0:5] = q["ql"]
    out[:, 5:10] = w["wl"]
    ij = np.asarray(nlist["ij"])
    out[:, 10] = np.bincount(ij[:, 0], minlength=n) if len(ij) else 0
    return out

"""CPU oracle of the learned ORCA proposal network (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
forward, backward and loss kernels.  The product never imports it.

The network and loss restate include/ecm_mlp.h (no reference counterpart:
the reference has no learned component).  With `bf16=True` the values the
GPU kernels round to BF16 are rounded here at the same points (inputs,
weights, hidden activations, output and hidden-layer gradients that feed
the backward GEMMs); everything else is FP64, so GPU-vs-oracle differences
come only from FP32 accumulation order and the hardware exp / reciprocal.
Its backward pass is pinned separately by central finite differences in
FP64 (bf16=False).
"""
from __future__ import annotations

import numpy as np

N_IN_PAD, HID, N_OUT, N_SLOTS = 64, 128, 16, 14
VEL_SCALE = 0.5
# flat parameter layout (row-major W[out][in], then the bias)
SHAPES = [("W1", (HID, N_IN_PAD)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("W4", (N_OUT, HID)), ("b4", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unflatten(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shp)
        o += n
    return out


def flatten(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def bf16_round(x):
    """Round to the nearest BF16 value, ties to even (as v_cvt_pk_bf16_f32)."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    out = (((u + r) >> 16) << 16).astype(np.uint32).view(np.float32)
    return out.astype(np.float64)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def forward(params, x, bf16=True):
    """Raw outputs (n x 16) and the intermediates of the forward pass."""
    P = unflatten(params)
    q = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    h = q(x)
    cache = {"h0": h}
    for l in (1, 2, 3):
        z = h @ q(P[f"W{l}"]).T + P[f"b{l}"]
        h = q(silu(z))
        cache[f"z{l}"] = z
        cache[f"h{l}"] = h
    y = h @ q(P["W4"]).T + P["b4"]
    return y, cache


def loss_grad(params, x, label, v, pos_weight=2.0, vel_weight=1.0, bf16=True):
    """Batch-mean loss and flat gradient, the loss of include/ecm_mlp.h:
    BCE over the slots that hold a constraint (positive class weighted by
    pos_weight) + vel_weight * 0.5 |y_v - v * 0.5|^2."""
    P = unflatten(params)
    q = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    y, c = forward(params, x, bf16)
    B = x.shape[0]
    label = np.asarray(label, np.uint64)
    slots = np.arange(N_SLOTS, dtype=np.uint64)
    t = ((label[:, None] >> slots) & 1).astype(np.float64)
    valid = ((label[:, None] >> (slots + 16)) & 1).astype(np.float64)
    lg = y[:, :N_SLOTS]
    sp = np.maximum(lg, 0) + np.log1p(np.exp(-np.abs(lg)))          # softplus
    bce = t * pos_weight * (sp - lg) + (1 - t) * sp
    tv = np.asarray(v, np.float64) * VEL_SCALE
    ev = y[:, N_SLOTS:] - tv
    loss = (np.sum(valid * bce) + 0.5 * vel_weight * np.sum(ev * ev)) / B
    s = sigmoid(lg)
    dy = np.zeros_like(y)
    dy[:, :N_SLOTS] = valid * (s * (1 - t + pos_weight * t) - pos_weight * t) / B
    dy[:, N_SLOTS:] = vel_weight * ev / B
    G = {}
    G["b4"] = dy.sum(0)
    dyq = q(dy)
    G["W4"] = dyq.T @ c["h3"]
    dh = dyq @ q(P["W4"])
    for l in (3, 2, 1):
        z = c[f"z{l}"]
        sg = sigmoid(z)
        dz = dh * (sg * (1 + z * (1 - sg)))
        G[f"b{l}"] = dz.sum(0)
        dzq = q(dz)
        G[f"W{l}"] = dzq.T @ c[f"h{l - 1}"]
        if l > 1:
            dh = dzq @ q(P[f"W{l}"])
    return float(loss), flatten(G)


def hidden_perm(k):
    """Input index of a hidden layer held at B-operand position k of the
    32x32x16 MFMA chain (the inference packing of the library)."""
    kc, kl = k // 16, k % 16
    return 32 * (kc // 2) + 16 * (kc % 2) + 8 * ((kl % 8) // 4) + 4 * (kl // 8) + (kl % 4)

"""CPU checker for the learned-apodization MLP (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker of the HIP kernels in
ultrasound-engine_amd/csrc/apod.hip; nothing in the product calls it.

The network has no counterpart in the reference (parity there is unpinned by
definition); this module states it in float64 numpy:

    H_0 = x (256 features, (re, im) interleaved per channel)
    Z_l = H_l W_l + b_l,  H_{l+1} = gelu(Z_l)  for l = 0, 1, 2
    O   = H_3 W_3 + b_3;  channel weight c = 1 + O[2c] + j O[2c+1]
    loss = (1/n) sum_b sw_b sum_{c active} |1 + o_c - g_c|^2

with gelu the tanh form 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))).
``emulate=True`` rounds to bf16 exactly where the kernels do (weights, the
layer inputs H_l, and the output gradients dZ_l before they enter a GEMM);
``emulate=False`` is the exact float64 network. The gradients are pinned by
central finite differences of the float64 loss (tests/test_apod.py), not by
restating them.

The frame rule, out = sum_c (1 + o_c) y_c over active channels with the DAS
sum below 4 active channels, and the input normalisation (rotate by
conj(sum y)/|sum y|, scale to unit mean power over the active channels) are
the ones apod.hip documents.
"""

from __future__ import annotations

import numpy as np

F = 256
LAYERS = 4
N_PARAMS = LAYERS * F * F + LAYERS * F
GELU_K = np.sqrt(2.0 / np.pi)
GELU_C = 0.044715


def bf16(a):
    """Round to the nearest bf16 (ties to even), returned as float64."""
    u = np.ascontiguousarray(np.asarray(a, dtype=np.float32)).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, 0x7FC0, r).astype(np.uint32)
    return (r << 16).view(np.float32).astype(np.float64)


def bf16_bits(a):
    """bf16 bit patterns (uint16) of the rounded values."""
    v = bf16(a).astype(np.float32).view(np.uint32)
    return (v >> 16).astype(np.uint16)


def from_bits(bits):
    return (np.asarray(bits, np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(z):
    t = np.tanh(GELU_K * (z + GELU_C * z ** 3))
    return 0.5 * z * (1.0 + t)


def gelu_grad(z):
    t = np.tanh(GELU_K * (z + GELU_C * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_K * (1.0 + 3.0 * GELU_C * z * z)


def unpack(params):
    """(W list [in][out], b list) views of a flat parameter vector."""
    p = np.asarray(params, dtype=np.float64)
    Ws = [p[l * F * F:(l + 1) * F * F].reshape(F, F) for l in range(LAYERS)]
    bs = [p[LAYERS * F * F + l * F:LAYERS * F * F + (l + 1) * F] for l in range(LAYERS)]
    return Ws, bs


def pack(Ws, bs):
    return np.concatenate([w.ravel() for w in Ws] + [b.ravel() for b in bs])


def active_features(mask):
    """(n, 256) 0/1 array: feature 2c and 2c+1 are on when channel c is."""
    mask = np.asarray(mask, np.uint32).reshape(-1, 4)
    ch = np.arange(F // 2)
    on = (mask[:, ch // 32] >> (ch % 32).astype(np.uint32)) & 1
    return np.repeat(on, 2, axis=1).astype(np.float64)


def forward(params, x, emulate=True):
    """Layer inputs H_0..H_3, pre-activations Z_0..Z_2 and the output O."""
    Ws, bs = unpack(params)
    r = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    H = [r(x)]
    Z = []
    for l in range(LAYERS):
        z = H[l] @ r(Ws[l]) + (bs[l].astype(np.float32) if emulate else bs[l])
        if l + 1 < LAYERS:
            Z.append(z)
            H.append(r(gelu(z)))
        else:
            O = z
    return H, Z, O


def loss_and_grad(params, x, g, mask, sw, emulate=True):
    """Loss and flat gradient (C-ABI parameter layout) of one batch."""
    x = np.asarray(x, np.float64)
    n = x.shape[0]
    Ws, bs = unpack(params)
    r = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    H, Z, O = forward(params, x, emulate)
    act = active_features(mask)
    sw = np.asarray(sw, np.float64)[:, None]
    pred = O + np.tile([1.0, 0.0], F // 2)[None, :]
    e = (pred - np.asarray(g, np.float64)) * act
    loss = float(np.sum(sw * e * e) / n)
    dz = r(2.0 / n * sw * e)
    gW = [None] * LAYERS
    gb = [None] * LAYERS
    for l in range(LAYERS - 1, -1, -1):
        gW[l] = H[l].T @ dz
        gb[l] = dz.sum(axis=0)
        if l == 0:
            break
        dh = dz @ r(Ws[l]).T
        dz = r(dh * gelu_grad(Z[l - 1]))
    return loss, pack(gW, gb)


def features(y, mask):
    """Network input of pixels from their channel vectors y (n, 128)
    complex and 4-word masks: active channels rotated by conj(sum)/|sum|,
    scaled to unit mean power; float64 (round with bf16() for the kernels'
    bf16 input)."""
    y = np.asarray(y, np.complex128)
    act = active_features(mask)[:, ::2]
    ya = y * act
    s = ya.sum(axis=1)
    na = act.sum(axis=1)
    e = (np.abs(ya) ** 2).sum(axis=1)
    mag = np.abs(s)
    rot = np.where(mag > 0, np.conj(s) / np.where(mag > 0, mag, 1.0), 1.0)
    scale = np.where((e > 0) & (na > 0), 1.0 / np.sqrt(np.where(e > 0, e, 1.0) / np.maximum(na, 1)), 0.0)
    xc = ya * (rot * scale)[:, None]
    out = np.empty((y.shape[0], F))
    out[:, 0::2] = xc.real
    out[:, 1::2] = xc.imag
    return out


def apply_weights(params, y, mask, emulate=True):
    """Frame values of pixels: sum_c (1 + o_c) y_c over active channels
    (the DAS sum below 4 active channels)."""
    y = np.asarray(y, np.complex128)
    x = features(y, mask)
    _, _, O = forward(params, bf16(x) if emulate else x, emulate)
    w = (1.0 + O[:, 0::2]) + 1j * O[:, 1::2]
    act = active_features(mask)[:, ::2]
    na = act.sum(axis=1)
    learned = np.sum(w * y * act, axis=1)
    das = np.sum(y * act, axis=1)
    return np.where(na >= 4, learned, das)

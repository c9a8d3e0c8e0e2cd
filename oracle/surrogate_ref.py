"""CPU restatement of the learned input-layer surrogate (test infrastructure).

Only tests and bench.py's checks use this module; the product path is the HIP
library (include/cbm_surrogate.h).  The surrogate has no counterpart in the
reference: this restates the network and its gradient as the header defines
them, in binary64 numpy, with BF16 rounding at the same points as the device
kernels, so the device's forward pass, cross-entropy and every parameter
gradient can be compared with it:

    H1 = bf16(relu(conv(X, W1, dil0) + b1))       X: MF spikes as 0/1
    H2 = bf16(relu(conv(H1, W2, dil1) + b2))
    Z3 = conv(H2, W3, dil2) + b3,  p = sigmoid(Z3)
    loss = mean over valid outputs of softplus(Z3) - y*Z3
    dZ3 = bf16((p - y) / count)   (0 on rows without a full receptive field)
    dZ2 = bf16(convT(dZ3, W3) * [H2 > 0]),  dZ1 = bf16(convT(dZ2, W2) * [H1 > 0])
    conv(X, W, d)[t][o] = sum_tap sum_c X[t - tap*d][c] W[o][tap][c]   (zero before the segment)

Weights enter as their BF16 roundings (the device multiplies BF16 copies of
its FP32 master weights); biases stay FP32.  `exact=True` drops every BF16
rounding: the binary64 network that the finite-difference and torch autograd
checks pin.
"""
from __future__ import annotations

import numpy as np


def bf16(x) -> np.ndarray:
    """Round binary32 values to BF16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bits_to_f64(h) -> np.ndarray:
    return (np.asarray(h, np.uint32) << 16).view(np.float32).astype(np.float64)


class Dims:
    def __init__(self, num_in, hidden, num_out, taps=8, dil=(1, 2, 4)):
        self.num_in, self.hidden, self.num_out, self.taps, self.dil = num_in, hidden, num_out, taps, tuple(dil)
        self.cin = (num_in, hidden, hidden)
        self.cout = (hidden, hidden, num_out)
        self.lead = (taps - 1) * sum(dil)

    def shapes(self):
        out = []
        for l in range(3):
            out += [(self.cout[l], self.taps, self.cin[l]), (self.cout[l],)]
        return out

    def num_params(self):
        return sum(int(np.prod(s)) for s in self.shapes())

    def split(self, flat):
        flat = np.asarray(flat)
        out, o = [], 0
        for s in self.shapes():
            n = int(np.prod(s))
            out.append(flat[o:o + n].reshape(s))
            o += n
        return out

    def join(self, arrays):
        return np.concatenate([np.asarray(a, np.float64).ravel() for a in arrays])


def conv(X, W, d):
    """X: [S, P, cin] (float64), W: [cout, taps, cin] -> [S, P, cout], causal, zero-padded per segment."""
    S, P, _ = X.shape
    out = np.zeros((S, P, W.shape[0]))
    for tap in range(W.shape[1]):
        sh = tap * d
        if sh >= P:
            break
        out[:, sh:, :] += X[:, :P - sh, :] @ W[:, tap, :].T
    return out


def conv_t(G, W, d):
    """Data gradient of conv: G: [S, P, cout] -> [S, P, cin]."""
    S, P, _ = G.shape
    out = np.zeros((S, P, W.shape[2]))
    for tap in range(W.shape[1]):
        sh = tap * d
        if sh >= P:
            break
        out[:, :P - sh, :] += G[:, sh:, :] @ W[:, tap, :]
    return out


def conv_w(G, X, d, taps):
    """Weight gradient of conv: sum over rows of G[t][o] X[t - tap*d][c] -> [cout, taps, cin]."""
    S, P, cout = G.shape
    cin = X.shape[2]
    out = np.zeros((cout, taps, cin))
    for tap in range(taps):
        sh = tap * d
        if sh >= P:
            break
        out[:, tap, :] = np.einsum("spo,spc->oc", G[:, sh:, :], X[:, :P - sh, :])
    return out


def segments(stream, starts, P):
    """Rows starts[s] .. starts[s]+P-1 of a [T, C] stream -> [S, P, C]."""
    return np.stack([np.asarray(stream[int(s):int(s) + P]) for s in starts])


def forward(dims: Dims, flat, X, exact=False):
    """X: MF spikes [S, P, num_in] (nonzero = spike).  Returns H1, H2, Z3."""
    r = (lambda v: np.asarray(v, np.float64)) if exact else bf16
    W1, b1, W2, b2, W3, b3 = dims.split(flat)
    x = (np.asarray(X) != 0).astype(np.float64)
    H1 = r(np.maximum(conv(x, r(W1), dims.dil[0]) + b1, 0.0))
    H2 = r(np.maximum(conv(H1, r(W2), dims.dil[1]) + b2, 0.0))
    Z3 = conv(H2, r(W3), dims.dil[2]) + b3
    return H1, H2, Z3


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def loss_grad(dims: Dims, flat, X, Y, exact=False):
    """Mean cross-entropy over rows lead.. of each segment and its gradient
    (flat, cbm_sg_get_params order).  Y: targets [S, P, num_out] in [0, 1]."""
    r = (lambda v: np.asarray(v, np.float64)) if exact else bf16
    W1, b1, W2, b2, W3, b3 = dims.split(flat)
    x = (np.asarray(X) != 0).astype(np.float64)
    H1, H2, Z3 = forward(dims, flat, X, exact)
    S, P, O = Z3.shape
    valid = np.zeros((S, P, 1))
    valid[:, dims.lead:, :] = 1.0
    count = S * (P - dims.lead) * O
    y = np.asarray(Y, np.float64)
    ce = np.maximum(Z3, 0) + np.log1p(np.exp(-np.abs(Z3))) - y * Z3
    loss = float(np.sum(ce * valid) / count)
    dZ3 = r((sigmoid(Z3) - y) * valid / count)
    gW3 = conv_w(dZ3, H2, dims.dil[2], dims.taps)
    gb3 = dZ3.sum(axis=(0, 1))
    dZ2 = r(conv_t(dZ3, r(W3), dims.dil[2]) * (H2 > 0))
    gW2 = conv_w(dZ2, H1, dims.dil[1], dims.taps)
    gb2 = dZ2.sum(axis=(0, 1))
    dZ1 = r(conv_t(dZ2, r(W2), dims.dil[1]) * (H1 > 0))
    gW1 = conv_w(dZ1, x, dims.dil[0], dims.taps)
    gb1 = dZ1.sum(axis=(0, 1))
    return loss, dims.join([gW1, gb1, gW2, gb2, gW3, gb3])

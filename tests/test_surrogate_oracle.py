"""Pins of the surrogate's CPU restatement (oracle/surrogate_ref.py), on the CPU.

The oracle's network and gradient are checked against two independent
computations, so that a transposed tap, a wrong shift direction, a missing
segment boundary or a dropped term fails one of them:
* torch's autograd through conv1d (left zero padding, dilation) in binary64,
  with the oracle's BF16 roundings switched off;
* central finite differences of the oracle's own binary64 loss.
Its BF16 rounding is checked against torch's float32 -> bfloat16 conversion
(round to nearest even).
"""
import numpy as np
import pytest
import torch

import surrogate_ref as R

DIMS = R.Dims(num_in=16, hidden=8, num_out=12, taps=3, dil=(1, 2, 3))


def _case(seed=0, S=3, L=7, T=80):
    rng = np.random.default_rng(seed)
    P = DIMS.lead + L
    flat = rng.normal(0.0, 0.4, DIMS.num_params())
    stream = (rng.random((T, DIMS.num_in)) < 0.3).astype(np.uint8)
    starts = rng.integers(0, T - P, S)
    X = R.segments(stream, starts, P)
    Y = rng.random((S, P, DIMS.num_out)) * 0.3
    return flat, X, Y


def _torch_loss(flat, X, Y):
    W1, b1, W2, b2, W3, b3 = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in DIMS.split(flat)]
    x = torch.tensor((X != 0).astype(np.float64)).permute(0, 2, 1)  # [S, C, P]

    def conv(h, W, b, d):
        # W[o][tap][c] weighs h[t - tap*d]: conv1d's kernel index k weighs h[t + k*d - pad]
        k = torch.flip(W, dims=[1]).permute(0, 2, 1)  # [o, c, taps]
        pad = (W.shape[1] - 1) * d
        return torch.nn.functional.conv1d(torch.nn.functional.pad(h, (pad, 0)), k, b, dilation=d)

    h1 = torch.relu(conv(x, W1, b1, DIMS.dil[0]))
    h2 = torch.relu(conv(h1, W2, b2, DIMS.dil[1]))
    z = conv(h2, W3, b3, DIMS.dil[2])[:, :, DIMS.lead:]
    y = torch.tensor(Y, dtype=torch.float64).permute(0, 2, 1)[:, :, DIMS.lead:]
    loss = torch.nn.functional.binary_cross_entropy_with_logits(z, y, reduction="mean")
    loss.backward()
    return float(loss.detach()), np.concatenate([a.grad.numpy().ravel() for a in (W1, b1, W2, b2, W3, b3)])


def test_oracle_matches_torch_autograd():
    for seed in range(3):
        flat, X, Y = _case(seed)
        loss, grad = R.loss_grad(DIMS, flat, X, Y, exact=True)
        tl, tg = _torch_loss(flat, X, Y)
        assert abs(loss - tl) <= 1e-12 * max(1.0, abs(tl))
        for a, b, shape in zip(DIMS.split(grad), DIMS.split(tg), DIMS.shapes()):
            np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-13, err_msg=str(shape))


def test_oracle_finite_differences():
    flat, X, Y = _case(7)
    _, grad = R.loss_grad(DIMS, flat, X, Y, exact=True)
    rng = np.random.default_rng(1)
    idx = rng.choice(flat.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(flat)
        e[i] = h
        lp, _ = R.loss_grad(DIMS, flat + e, X, Y, exact=True)
        lm, _ = R.loss_grad(DIMS, flat - e, X, Y, exact=True)
        num = (lp - lm) / (2 * h)
        assert abs(num - grad[i]) <= 1e-6 * max(1e-3, abs(grad[i])) + 1e-9, (i, num, grad[i])


def test_receptive_field_is_causal_and_lead_long():
    # an output depends on exactly the `lead + 1` rows that end at it
    flat, X, _ = _case(3, S=1, L=5)
    _, _, Z = R.forward(DIMS, flat, X, exact=True)
    X2 = X.copy()
    X2[0, -1, :] ^= 1  # the last row: only the last output may change
    _, _, Z2 = R.forward(DIMS, flat, X2, exact=True)
    assert np.array_equal(Z[0, :-1], Z2[0, :-1]) and not np.array_equal(Z[0, -1], Z2[0, -1])
    X3 = X.copy()
    X3[0, 0, :] ^= 1  # row 0 reaches outputs up to row lead only
    _, _, Z3 = R.forward(DIMS, flat, X3, exact=True)
    assert np.array_equal(Z[0, DIMS.lead + 1:], Z3[0, DIMS.lead + 1:])
    assert not np.array_equal(Z[0, DIMS.lead], Z3[0, DIMS.lead])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 10, 10000), [1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, -2.5, 0.0, 1e-30]]).astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(R.bf16(x), ref)
    assert np.array_equal(R.bits_to_f64(R.bf16_bits(x)), ref)


@pytest.mark.parametrize("exact", [False, True])
def test_gradient_of_bias_is_sum_of_pre_activation_gradient(exact):
    # b3's gradient is the sum over valid outputs of (p - y)/count: a closed form
    flat, X, Y = _case(5)
    _, _, Z = R.forward(DIMS, flat, X, exact)
    _, g = R.loss_grad(DIMS, flat, X, Y, exact)
    S, P, O = Z.shape
    want = ((R.sigmoid(Z) - Y)[:, DIMS.lead:, :] / (S * (P - DIMS.lead) * O)).sum(axis=(0, 1))
    got = DIMS.split(g)[5]
    tol = 0 if exact else 2**-8
    np.testing.assert_allclose(got, want, rtol=1e-12 + tol, atol=1e-12 + tol * np.abs(want).max())

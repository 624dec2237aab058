"""Independent pins of the network oracle's forward pass (orc_mlp).

The surrogate's GPU kernels are compared with ``oracle.mlp``; these tests
check ``oracle.mlp`` itself against what does not share its code:

  * a numpy forward pass of the 3 -> 64 -> 64 -> 1 network written from the
    parameter layout of include/tlq.h (W1[64][3], b1[64], W2[64][64] with
    row = output unit, b2[64], W3[64], b3) with the textbook tanh GELU,
    0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))), not the oracle's
    sigmoid rewrite x / (1 + exp(-u)); bf16 rounding done by scaling to the
    value's binade and numpy's round-half-even, not by bit manipulation;
  * a single-path network (one live unit per layer) whose output reduces to
    nested scalar GELUs evaluated by hand with math.tanh;
  * the limits GELU(0) = 0, GELU(x) -> x for x >= 6 and -> 0 for x <= -6,
    through networks whose output those limits fix.

A slip in the oracle's GELU argument scale, a transposed W2 or a dropped
bias fails the first test; the input normalisation (training box onto
[-1, 1]) is restated from the box bounds.
"""
import math

import numpy as np
import pytest

import oracle

H = 64
OW1, OB1, OW2, OB2, OW3, OB3 = 0, 3 * H, 4 * H, 4 * H + H * H, 5 * H + H * H, 6 * H + H * H
NP = OB3 + 1
# training box of the inputs (log10 n, log10 T, log10(dt/t_cool)), mapped onto [-1, 1]
BOX = np.array([[-3.0, 4.0], [4.0, 9.0], [-2.0, 2.0]])


def gelu_tanh(x):
    return 0.5 * x * (1.0 + np.tanh(np.sqrt(2.0 / np.pi) * (x + 0.044715 * x ** 3)))


def bf16(x):
    """Round to float32, then to bf16 (8 significant bits), nearest-even,
    as value arithmetic: scale by the binade and use numpy's rint."""
    f = np.asarray(x, dtype=np.float32).astype(np.float64)
    out = np.zeros_like(f)
    nz = f != 0.0
    e = np.floor(np.log2(np.abs(f[nz])))
    ulp = np.exp2(e - 7.0)
    out[nz] = np.rint(f[nz] / ulp) * ulp
    return out


def forward(theta, x, quant):
    q = bf16 if quant else (lambda v: np.asarray(v, dtype=np.float32).astype(np.float64))
    th = np.asarray(theta, dtype=np.float32)
    W1 = q(th[OW1:OB1]).reshape(H, 3)
    b1 = q(th[OB1:OW2])
    W2 = q(th[OW2:OB2]).reshape(H, H)
    b2 = q(th[OB2:OW3])
    W3 = q(th[OW3:OB3])
    b3 = q(th[OB3:OB3 + 1])[0]
    xn = 2.0 * (x - BOX[:, 0]) / (BOX[:, 1] - BOX[:, 0]) - 1.0
    z1 = xn @ W1.T + b1
    h1 = gelu_tanh(z1)
    if quant:
        h1 = bf16(h1)
    z2 = h1 @ W2.T + b2
    return gelu_tanh(z2) @ W3 + b3


def _inputs(rng, n):
    return np.stack([rng.uniform(lo, hi, n) for lo, hi in BOX], 1)


@pytest.mark.parametrize("quant", [True, False])
def test_forward_matches_numpy_textbook_gelu(quant):
    rng = np.random.default_rng(11)
    th = (rng.standard_normal(NP) * 0.3).astype(np.float32)
    x = _inputs(rng, 512)
    got = oracle.mlp(th, x, quant=quant)[0]
    ref = forward(th, x, quant)
    # Same roundings at the same points; the sums differ in order only, and a
    # bf16 rounding of h1 could flip at a tie by that -- none does at this seed.
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    if quant:
        assert np.abs(got - forward(th, x, False)).max() > 1e-6   # the roundings matter


def test_forward_transposed_w2_differs():
    # the test above would catch a transposed W2: the layouts give different outputs
    rng = np.random.default_rng(12)
    th = (rng.standard_normal(NP) * 0.3).astype(np.float32)
    tt = th.copy()
    tt[OW2:OB2] = th[OW2:OB2].reshape(H, H).T.ravel()
    x = _inputs(rng, 64)
    assert np.abs(oracle.mlp(th, x)[0] - oracle.mlp(tt, x)[0]).max() > 1e-3


def _single_path(a, c, w, d, v, e):
    """Unit 0 of layer 1 sees input 0 (weight a, bias c); unit 0 of layer 2
    sees it (weight w, bias d); the output reads unit 0 (weight v, bias e).
    Every other parameter is zero, so every other unit is GELU(0) = 0."""
    th = np.zeros(NP, dtype=np.float32)
    th[OW1 + 0] = a
    th[OB1 + 0] = c
    th[OW2 + 0] = w
    th[OB2 + 0] = d
    th[OW3 + 0] = v
    th[OB3] = e
    return th


def _gelu_scalar(x):
    return 0.5 * x * (1.0 + math.tanh(math.sqrt(2.0 / math.pi) * (x + 0.044715 * x ** 3)))


@pytest.mark.parametrize("quant", [True, False])
def test_single_path_network_is_nested_scalar_gelus(quant):
    # parameters exactly representable in bf16, so the weight rounding is exact
    a, c, w, d, v, e = 1.5, 0.25, -1.25, 0.375, 2.0, -0.0625
    th = _single_path(a, c, w, d, v, e)
    for x0 in (-3.0, -1.1, 0.3, 2.2, 4.0):
        x = np.array([[x0, 6.0, 0.5]])
        xn0 = 2.0 * (x0 + 3.0) / 7.0 - 1.0
        h = _gelu_scalar(a * xn0 + c)
        if quant:
            h = float(bf16(np.array([h]))[0])
        want = e + v * _gelu_scalar(w * h + d)
        assert oracle.mlp(th, x, quant=quant)[0][0] == pytest.approx(want, rel=1e-14, abs=1e-15)


def test_zero_network_outputs_its_bias():
    th = np.zeros(NP, dtype=np.float32)
    th[OB3] = 0.71875                       # bf16-exact
    rng = np.random.default_rng(13)
    out = oracle.mlp(th, _inputs(rng, 32))[0]
    assert np.all(out == 0.71875)           # GELU(0) = 0 in both layers


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_saturated_gelus_are_linear_or_zero(sign):
    # b1 = +-8 drives every layer-1 unit to GELU(+-8) = +-8 (|error| < 1e-14)
    # or 0 (|GELU(-8)| < 1e-14); layer 2 then sees 64 h1 through W2 = 0.125 and
    # its bias 0.5, giving z2 = 64*8*0.125 + 0.5 = 64.5 (GELU = 64.5) or 0.5.
    th = np.zeros(NP, dtype=np.float32)
    th[OB1:OW2] = 8.0 * sign
    th[OW2:OB2] = 0.125
    th[OB2:OW3] = 0.5
    th[OW3:OB3] = 1.0 / 64
    th[OB3] = 0.0
    x = _inputs(np.random.default_rng(14), 16)
    out = oracle.mlp(th, x)[0]
    want = 64.5 if sign > 0 else _gelu_scalar(0.5)
    assert np.allclose(out, want, rtol=1e-12, atol=1e-13)
    assert abs(_gelu_scalar(6.0) - 6.0) < 1e-7 and abs(_gelu_scalar(-6.0)) < 1e-7

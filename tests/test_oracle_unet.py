"""Pins of the oracle's U-Net (the learned warm start's network): analytic
gradients against central finite differences for every parameter block, a
1x1-equivalent case against numpy, and the bfloat16 rounding helper against
exact bit patterns."""
import ctypes as C

import numpy as np
import pytest

import nusif_api as A
from conftest import lib

DP = A._DP


def cfg_arr(L, ch, cin):
    a = (C.c_int * 6)(L, *(list(ch) + [0] * (4 - len(ch))), cin)
    return a


def loss(cfg, n, params, x, t, grad=None, pred=None, bf16=0):
    g = grad.ctypes.data_as(DP) if grad is not None else None
    p = pred.ctypes.data_as(DP) if pred is not None else None
    return lib("orc_unet").orcUnetLoss(cfg, n, params.ctypes.data_as(DP), x.ctypes.data_as(DP),
                                      t.ctypes.data_as(DP), g, p, bf16)


def blocks(L, ch, cin):
    """(name, start, length) of every parameter block, in the flat order."""
    cins, couts = [], []
    for l in range(L):
        cins += [cin if l == 0 else ch[l - 1], ch[l]]
        couts += [ch[l], ch[l]]
    for l in range(L - 2, -1, -1):
        cins += [ch[l + 1] + ch[l], ch[l]]
        couts += [ch[l], ch[l]]
    out, off = [], 0
    for q, (ci, co) in enumerate(zip(cins, couts)):
        out.append(("W%d" % q, off, co * 27 * ci))
        off += co * 27 * ci
        out.append(("b%d" % q, off, co))
        off += co
    out.append(("headW", off, ch[0]))
    off += ch[0]
    out.append(("headB", off, 1))
    return out, off + 1


@pytest.mark.parametrize("L,ch", [(3, (4, 6, 8)), (2, (3, 5))])
def test_unet_gradient_matches_finite_differences(L, ch):
    cin, n = 2, 8
    cfg = cfg_arr(L, ch, cin)
    bl, total = blocks(L, ch, cin)
    assert lib("orc_unet").orcUnetParamCount(cfg) == total
    rng = np.random.default_rng(7)
    params = rng.normal(0, 0.3, total)
    x = rng.uniform(-1, 1, (n, n, n, cin))
    t = rng.uniform(-1, 1, (n, n, n))
    grad = np.zeros(total)
    loss(cfg, n, params, x, t, grad)
    eps = 1e-6
    for name, off, length in bl:
        for idx in rng.choice(length, size=min(4, length), replace=False):
            q = off + int(idx)
            pp = params.copy(); pp[q] += eps
            pm = params.copy(); pm[q] -= eps
            fd = (loss(cfg, n, pp, x, t) - loss(cfg, n, pm, x, t)) / (2 * eps)
            assert abs(fd - grad[q]) <= 1e-6 * max(1.0, abs(fd)), (name, q, fd, grad[q])


def test_single_level_reduces_to_two_convs_and_head():
    """L = 1: head(relu(conv(relu(conv(x))))) against a direct numpy version
    with 3x3x3 zero-padded correlation."""
    cin, co, n = 2, 3, 5
    cfg = cfg_arr(1, (co,), cin)
    total = lib("orc_unet").orcUnetParamCount(cfg)
    rng = np.random.default_rng(3)
    params = rng.normal(0, 0.5, total)
    x = rng.uniform(-1, 1, (n, n, n, cin))
    t = np.zeros((n, n, n))
    pred = np.zeros((n, n, n))
    loss(cfg, n, params, x, t, pred=pred)

    def conv(inp, W, b):
        ci = inp.shape[-1]
        W = W.reshape(-1, 27, ci)
        pad = np.pad(inp, ((1, 1), (1, 1), (1, 1), (0, 0)))
        out = np.zeros(inp.shape[:3] + (W.shape[0],)) + b
        for tap in range(27):
            dz, dy, dx = tap // 9, (tap // 3) % 3, tap % 3
            sl = pad[dz:dz + n, dy:dy + n, dx:dx + n, :]
            out += sl @ W[:, tap, :].T
        return np.maximum(out, 0)
    o = 0
    W0 = params[o:o + co * 27 * cin]; o += W0.size
    b0 = params[o:o + co]; o += co
    W1 = params[o:o + co * 27 * co]; o += W1.size
    b1 = params[o:o + co]; o += co
    hw = params[o:o + co]; hb = params[o + co]
    y = conv(conv(x, W0, b0), W1, b1) @ hw + hb
    np.testing.assert_allclose(pred, y, rtol=1e-12, atol=1e-12)


def test_bf16_rounding_is_round_to_nearest_even():
    f = lib("orc_unet").orcBf16
    assert f(1.0) == 1.0
    assert f(1.0 + 2 ** -8) == 1.0            # halfway, even mantissa stays
    assert f(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6  # halfway, odd rounds up
    assert f(1.0 + 2 ** -7) == 1.0 + 2 ** -7
    assert f(-3.0000001) == -3.0

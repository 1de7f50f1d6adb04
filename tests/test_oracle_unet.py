"""Pins of the oracle's U-Net (the learned warm start's network): analytic
gradients against central finite differences for every parameter block, the
whole multi-level forward against an independent numpy restatement on cubes
and boxes, the pooling/upsampling layers against reshape-mean and np.repeat
and their backward passes against the adjoint identities, and the bfloat16
rounding helper against exact bit patterns."""
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


# ---- independent numpy restatement of the forward --------------------------
# Tensors are [z][y][x][c] (x fastest, channels last), as the oracle stores
# them. Pooling is a reshape-mean over 2x2x2 blocks, upsampling np.repeat
# along each axis, the decoder input concat(upsampled, skip) in that order,
# and each conv a zero-padded 3x3x3 correlation whose tap t is
# (dz+1)*9 + (dy+1)*3 + (dx+1).

def np_conv(inp, W, b):
    nz, ny, nx, ci = inp.shape
    W = W.reshape(-1, 27, ci)
    pad = np.pad(inp, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((nz, ny, nx, W.shape[0])) + b
    for tap in range(27):
        dz, dy, dx = tap // 9, (tap // 3) % 3, tap % 3
        out += pad[dz:dz + nz, dy:dy + ny, dx:dx + nx, :] @ W[:, tap, :].T
    return np.maximum(out, 0.0)


def np_pool(a):
    nz, ny, nx, c = a.shape
    return a.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(1, 3, 5))


def np_up(a):
    return a.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def np_unet(params, x, L, ch, cin):
    """The network's prediction, [z][y][x], from the flat parameter vector."""
    bl, total = blocks(L, ch, cin)
    assert params.size == total
    it = iter(bl)

    def take():
        (_, ow, lw), (_, ob, lb) = next(it), next(it)
        return params[ow:ow + lw], params[ob:ob + lb]
    skips, cur = [], x
    for l in range(L):
        a = np_conv(cur, *take())
        s = np_conv(a, *take())
        skips.append(s)
        if l < L - 1:
            cur = np_pool(s)
    prev = skips[L - 1]
    for l in range(L - 2, -1, -1):
        cat = np.concatenate([np_up(prev), skips[l]], axis=-1)
        prev = np_conv(np_conv(cat, *take()), *take())
    (_, oh, lh), (_, ob, _) = bl[-2], bl[-1]
    return prev @ params[oh:oh + lh] + params[ob]


@pytest.mark.parametrize("L,ch,ext", [
    (2, (16, 32), (8, 8, 8)),
    (2, (4, 6), (10, 6, 4)),                 # box, x longest
    (4, (16, 32, 64, 128), (16, 16, 16)),    # the warm start's network
    (4, (16, 32, 64, 128), (8, 16, 24)),     # box, z longest: nx, ny, nz differ
])
def test_unet_forward_matches_numpy_restatement(L, ch, ext):
    nx, ny, nz = ext
    cin = 8
    cfg = cfg_arr(L, ch, cin)
    _, total = blocks(L, ch, cin)
    assert lib("orc_unet").orcUnetParamCount(cfg) == total
    rng = np.random.default_rng(11 + L)
    params = np.empty(total)
    bl = blocks(L, ch, cin)[0]
    for (_, ow, lw), (_, ob, lb) in zip(bl[0::2], bl[1::2]):
        # He-normal weights keep the ReLU activations O(1) through the levels
        fan_in = lw // lb  # 27 cin for a conv, ch0 for the head
        params[ow:ow + lw] = rng.normal(0, np.sqrt(2.0 / fan_in), lw)
        params[ob:ob + lb] = rng.normal(0, 0.1, lb)
    x = rng.uniform(-1, 1, (nz, ny, nx, cin))
    t = rng.uniform(-1, 1, (nz, ny, nx))
    pred = np.zeros((nz, ny, nx))
    ref = np_unet(params, x, L, ch, cin)
    lo = lib("orc_unet").orcUnetLossBox(cfg, nx, ny, nz, params.ctypes.data_as(DP), x.ctypes.data_as(DP),
                                        t.ctypes.data_as(DP), None, pred.ctypes.data_as(DP), 0)
    assert np.abs(ref).max() > 1e-3  # the deep levels reach the output
    np.testing.assert_allclose(pred, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    assert lo == pytest.approx(np.mean((ref - t) ** 2), rel=1e-12)
    if nx == ny == nz:  # the cube entry point is the box one
        pred2 = np.zeros_like(pred)
        loss(cfg, nx, params, x, t, pred=pred2)
        assert np.array_equal(pred, pred2)


@pytest.mark.parametrize("ext", [(4, 4, 4), (3, 2, 5)])
def test_pool_and_upsample_layers(ext):
    """pool = reshape-mean, up = np.repeat; their backward passes are the
    adjoints: <pool f, g> = <f, pool^T g> and <up g, f> = <g, up^T f>."""
    nx, ny, nz = ext
    c = 3
    tr = lib("orc_unet").orcUnetTransfer
    rng = np.random.default_rng(5)
    f = rng.normal(size=(2 * nz, 2 * ny, 2 * nx, c))
    g = rng.normal(size=(nz, ny, nx, c))
    out = np.zeros_like(g)
    tr(0, nx, ny, nz, c, f.ctypes.data_as(DP), out.ctypes.data_as(DP))
    np.testing.assert_allclose(out, np_pool(f), rtol=1e-14, atol=1e-14)
    fin = np.zeros_like(f)
    tr(2, nx, ny, nz, c, fin.ctypes.data_as(DP), g.ctypes.data_as(DP))
    assert np.array_equal(fin, np_up(g))
    # adjoints (the backward passes add into their output)
    ptg = np.zeros_like(f)
    tr(1, nx, ny, nz, c, ptg.ctypes.data_as(DP), g.copy().ctypes.data_as(DP))
    assert np.sum(np_pool(f) * g) == pytest.approx(np.sum(f * ptg), rel=1e-12)
    utf = np.zeros_like(g)
    tr(3, nx, ny, nz, c, f.copy().ctypes.data_as(DP), utf.ctypes.data_as(DP))
    assert np.sum(np_up(g) * f) == pytest.approx(np.sum(g * utf), rel=1e-12)
    # and they are the transposes, not merely adjoint in one random direction
    np.testing.assert_allclose(ptg, np_up(g) / 8.0, rtol=1e-14)
    np.testing.assert_allclose(utf, 8.0 * np_pool(f), rtol=1e-13)


def test_box_gradient_matches_finite_differences():
    """the backward on a box (nx, ny, nz all different), spot-checked"""
    L, ch, cin = 2, (3, 5), 2
    nx, ny, nz = 6, 4, 8
    cfg = cfg_arr(L, ch, cin)
    bl, total = blocks(L, ch, cin)
    rng = np.random.default_rng(9)
    params = rng.normal(0, 0.3, total)
    x = rng.uniform(-1, 1, (nz, ny, nx, cin))
    t = rng.uniform(-1, 1, (nz, ny, nx))
    f = lib("orc_unet").orcUnetLossBox

    def lossb(p, grad=None):
        return f(cfg, nx, ny, nz, p.ctypes.data_as(DP), x.ctypes.data_as(DP), t.ctypes.data_as(DP),
                 grad.ctypes.data_as(DP) if grad is not None else None, None, 0)
    grad = np.zeros(total)
    lossb(params, grad)
    eps = 1e-6
    for name, off, length in bl:
        for idx in rng.choice(length, size=min(3, length), replace=False):
            q = off + int(idx)
            pp = params.copy(); pp[q] += eps
            pm = params.copy(); pm[q] -= eps
            fd = (lossb(pp) - lossb(pm)) / (2 * eps)
            assert abs(fd - grad[q]) <= 1e-6 * max(1.0, abs(fd)), (name, q, fd, grad[q])

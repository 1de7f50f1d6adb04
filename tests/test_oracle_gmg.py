"""Pins of the oracle's geometric multigrid (the added `gmg` solver) from
the mathematics: transfer operators exact where they must be, convergence to
a manufactured discrete solution, V-cycle contraction, FP32 defect correction
reaching the FP64 answer."""
import ctypes as C
import math

import numpy as np
import pytest

import nusif_api as A
from conftest import lib

DP = A._DP


def ptr(a):
    return a.ctypes.data_as(DP)


def centre_grid(ni, nj, nk, h=(1.0, 1.0, 1.0)):
    x = (np.arange(ni + 2) - 0.5) * h[0]
    y = (np.arange(nj + 2) - 0.5) * h[1]
    z = (np.arange(nk + 2) - 0.5) * h[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def test_prolongation_is_exact_on_linear_fields():
    ni, nj, nk = 16, 12, 8
    Xc, Yc, Zc = centre_grid(ni // 2, nj // 2, nk // 2, (2.0, 2.0, 2.0))
    Xf, Yf, Zf = centre_grid(ni, nj, nk)
    lin = lambda X, Y, Z: 0.3 + 1.5 * X - 0.7 * Y + 0.25 * Z
    xc = np.ascontiguousarray(lin(Xc, Yc, Zc))
    xf = np.zeros_like(Xf)
    lib("orc_gmg").orcGmgProlong(ni, nj, nk, ptr(xc), ptr(xf))
    inner = (slice(2, nk), slice(2, nj), slice(2, ni))  # away from the clamped border
    np.testing.assert_allclose(xf[inner], lin(Xf, Yf, Zf)[inner], rtol=1e-13, atol=1e-13)
    # constants are reproduced everywhere (weights sum to 1, clamping included)
    xc[...] = 2.5
    xf[...] = 0.0
    lib("orc_gmg").orcGmgProlong(ni, nj, nk, ptr(xc), ptr(xf))
    np.testing.assert_allclose(xf[1:-1, 1:-1, 1:-1], 2.5, rtol=1e-15)


def test_restriction_averages_children():
    ni, nj, nk = 8, 6, 4
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, (nk + 2, nj + 2, ni + 2))
    x = np.zeros_like(b)                      # residual = b
    bc = np.zeros((nk // 2 + 2, nj // 2 + 2, ni // 2 + 2))
    lib("orc_gmg").orcGmgRestrict(ni, nj, nk, 0.1, 0.1, 0.1, ptr(x), ptr(b), ptr(bc))
    inner = b[1:-1, 1:-1, 1:-1].reshape(nk // 2, 2, nj // 2, 2, ni // 2, 2)
    np.testing.assert_allclose(bc[1:-1, 1:-1, 1:-1], inner.mean(axis=(1, 3, 5)), rtol=1e-14)


def solve(p, rhs, n, levels=5, tol=1e-10, maxc=60, mixed=0, project=1, h=None):
    ni, nj, nk = n
    h = h or (1.0 / ni, 1.0 / nj, 1.0 / nk)
    rel = C.c_double()
    cyc = lib("orc_gmg").orcGmgSolve(ni, nj, nk, h[0], h[1], h[2], ptr(p), ptr(rhs), levels, 2, 2,
                                     1.0, tol, maxc, project, mixed, C.byref(rel))
    return cyc, rel.value


@pytest.mark.parametrize("mixed", [0, 1])
def test_gmg_converges_to_manufactured_solution(mixed):
    n = (32, 16, 24)
    h = (1.0 / 32, 0.5 / 16, 0.75 / 24)
    X, Y, Z = centre_grid(*n, h=h)
    L = (1.0, 0.5, 0.75)
    m = (1, 2, 1)
    sol = np.cos(m[0] * math.pi * X / L[0]) * np.cos(m[1] * math.pi * Y / L[1]) * \
        np.cos(m[2] * math.pi * Z / L[2])
    lam = sum(-4.0 / hh ** 2 * math.sin(mm * math.pi * hh / (2 * ll)) ** 2
              for hh, ll, mm in zip(h, L, m))
    rhs = np.ascontiguousarray(lam * sol)
    p = np.zeros_like(rhs)
    cyc, rel = solve(p, rhs, n, levels=4, tol=1e-11, mixed=mixed, h=h)
    assert rel <= 1e-11 and cyc < 25
    got = p[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(got - got.mean(), sol[1:-1, 1:-1, 1:-1], atol=2e-9)


def test_vcycle_contraction():
    n = (64, 64, 64)
    rng = np.random.default_rng(1)
    rhs = rng.uniform(-1, 1, (66, 66, 66))
    rels = []
    for c in (1, 4):
        p = np.zeros_like(rhs)
        _, rel = solve(p, rhs.copy(), n, levels=6, tol=0.0, maxc=c)
        rels.append(rel)
    rate = (rels[1] / rels[0]) ** (1.0 / 3.0)
    assert rate < 0.2, rate


def test_mixed_matches_fp64():
    n = (32, 32, 32)
    rng = np.random.default_rng(2)
    rhs = rng.uniform(-1, 1, (34, 34, 34))
    p64 = np.zeros_like(rhs)
    p32 = np.zeros_like(rhs)
    c64, r64 = solve(p64, rhs.copy(), n, tol=1e-10)
    c32, r32 = solve(p32, rhs.copy(), n, tol=1e-10, mixed=1)
    assert r64 <= 1e-10 and r32 <= 1e-10
    assert abs(c64 - c32) <= 1
    a = p64[1:-1, 1:-1, 1:-1]
    b = p32[1:-1, 1:-1, 1:-1]
    assert np.max(np.abs((a - a.mean()) - (b - b.mean()))) <= 1e-9 * np.max(np.abs(a))


def solve_grouped(p, rhs, n, first, levels=5, tol=1e-10, maxc=60, mixed=0, project=1):
    ni, nj, nk = n
    rel = C.c_double()
    cyc = lib("orc_gmg").orcGmgSolveGrouped(ni, nj, nk, 1.0 / ni, 1.0 / nj, 1.0 / nk, ptr(p), ptr(rhs),
                                            levels, 2, 2, 1.0, tol, maxc, project, mixed, first,
                                            C.byref(rel))
    return cyc, rel.value


@pytest.mark.parametrize("mixed", [0, 1])
def test_group_of_one_is_the_plain_loop(mixed):
    n = (32, 16, 24)
    rhs = np.random.default_rng(7).uniform(-1, 1, (26, 18, 34))
    pa, pb = np.zeros_like(rhs), np.zeros_like(rhs)
    ca = solve(pa, rhs.copy(), n, levels=4, tol=1e-9, mixed=mixed)
    cb = solve_grouped(pb, rhs.copy(), n, 1, levels=4, tol=1e-9, mixed=mixed)
    assert ca == cb
    np.testing.assert_array_equal(pa, pb)


@pytest.mark.parametrize("m", [2, 3])
def test_fp64_group_gives_the_plain_iterate(m):
    """The stop test does not touch p: m FP64 cycles in one group are the
    plain loop's m-th iterate, bit for bit."""
    n = (32, 32, 16)
    rhs = np.random.default_rng(8).uniform(-1, 1, (18, 34, 34))
    pa, pb = np.zeros_like(rhs), np.zeros_like(rhs)
    ca, ra = solve(pa, rhs.copy(), n, levels=4, tol=0.0, maxc=m)
    cb, rb = solve_grouped(pb, rhs.copy(), n, m, levels=4, tol=0.0, maxc=m)
    assert ca == cb == m and ra == rb
    np.testing.assert_array_equal(pa, pb)


def test_mixed_group_continues_the_correction():
    """Three FP32 cycles on one defect reduce the residual like three plain
    mixed cycles (within 10x), and well below one cycle (10x; one cycle
    contracts by ~0.03 here): a group that restarted e from zero, or ran one
    cycle, fails this."""
    n = (32, 32, 32)
    rhs = np.random.default_rng(9).uniform(-1, 1, (34, 34, 34))
    rels = {}
    for name, first, maxc in (("one", 1, 1), ("plain3", 1, 3), ("group3", 3, 3)):
        p = np.zeros_like(rhs)
        cyc, rel = solve_grouped(p, rhs.copy(), n, first, levels=5, tol=0.0, maxc=maxc, mixed=1)
        assert cyc == maxc
        rels[name] = rel
    assert rels["group3"] <= 10.0 * rels["plain3"]
    assert rels["group3"] <= 0.1 * rels["one"]
    # and a grouped solve still stops at the tolerance, overshooting by at
    # most the group's extra cycles
    p0, p1 = np.zeros_like(rhs), np.zeros_like(rhs)
    c0, r0 = solve_grouped(p0, rhs.copy(), n, 1, tol=1e-9, mixed=1)
    c1, r1 = solve_grouped(p1, rhs.copy(), n, 4, tol=1e-9, mixed=1)
    assert r0 <= 1e-9 and r1 <= 1e-9 and c0 <= c1 <= max(c0, 4)


def test_coarse_dct_solve_is_the_min_norm_least_squares_solution():
    """The coarsest-level solve equals the pseudo-inverse of the dense
    Neumann Laplacian (numpy lstsq) on a small anisotropic box."""
    ni, nj, nk = 6, 4, 5
    h = (0.3, 0.5, 0.4)
    n = ni * nj * nk
    A = np.zeros((n, n))
    idx = lambda i, j, k: (k * nj + j) * ni + i
    for k in range(nk):
        for j in range(nj):
            for i in range(ni):
                r = idx(i, j, k)
                for d, (ext, hh) in enumerate(zip((ni, nj, nk), h)):
                    c = (i, j, k)
                    for s in (-1, 1):
                        q = list(c)
                        q[d] += s
                        if 0 <= q[d] < ext:         # interior neighbour
                            A[r, idx(*q)] += 1.0 / hh ** 2
                            A[r, r] -= 1.0 / hh ** 2
                        # ghost copy: neighbour equals the cell itself -> no term
    rng = np.random.default_rng(4)
    bvec = rng.uniform(-1, 1, n)
    bvec -= bvec.mean()
    x_ref = np.linalg.lstsq(A, bvec, rcond=None)[0]
    b = np.zeros((nk + 2, nj + 2, ni + 2))
    b[1:-1, 1:-1, 1:-1] = bvec.reshape(nk, nj, ni)
    x = np.zeros_like(b)
    lib("orc_gmg").orcGmgCoarseSolve(ni, nj, nk, h[0], h[1], h[2], ptr(b), ptr(x))
    np.testing.assert_allclose(x[1:-1, 1:-1, 1:-1].ravel(), x_ref, atol=1e-12)

"""Oracle pins from the mathematics, independent of the reference's code:
exact discrete identities a dropped term, a wrong sign, index or factor
would break."""
import math

import numpy as np
import pytest

import nusif_api as A
from conftest import lib

BASE = dict(name="none", imax=10, jmax=8, kmax=6, xlength=1.0, ylength=2.0, zlength=1.5,
            re=50.0, gamma=0.0, tau=0.5, dt=0.01, omg=1.0, eps=1e-12, itermax=1)


def centres(case):
    g = case.d.grid
    nk, nj, ni = case.shape
    x = (np.arange(ni) - 0.5) * g.dx
    y = (np.arange(nj) - 0.5) * g.dy
    z = (np.arange(nk) - 0.5) * g.dz
    return np.meshgrid(z, y, x, indexing="ij")[::-1]  # X, Y, Z in (k, j, i) layout


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def test_rhs_is_exact_divergence_of_linear_faces():
    c = A.Case(lib("orc_rb"), **BASE)
    X, Y, Z = centres(c)
    g = c.d.grid
    # F at x-faces i+1/2: x = X + dx/2, etc.; div = 2 - 3 + 0.5
    c.set_fields(f=2.0 * (X + g.dx / 2), g=-3.0 * (Y + g.dy / 2), h=0.5 * (Z + g.dz / 2))
    c.call("computeRHS")
    rhs = interior(c.fields()["rhs"])
    np.testing.assert_allclose(rhs, (2.0 - 3.0 + 0.5) / c.d.dt, rtol=1e-11)


def test_adapt_subtracts_exact_gradient():
    c = A.Case(lib("orc_rb"), **BASE)
    X, Y, Z = centres(c)
    c.set_fields(p=1.5 * X - 0.25 * Y + 4.0 * Z, f=np.ones(c.shape), g=np.zeros(c.shape),
                 h=np.full(c.shape, -1.0))
    c.call("adaptUV")
    f = c.fields()
    dt = c.d.dt
    np.testing.assert_allclose(interior(f["u"]), 1.0 - 1.5 * dt, rtol=1e-12)
    np.testing.assert_allclose(interior(f["v"]), 0.25 * dt, rtol=1e-10)
    np.testing.assert_allclose(interior(f["w"]), -1.0 - 4.0 * dt, rtol=1e-12)


def test_timestep_formula_counts_ghost_cells():
    c = A.Case(lib("orc_rb"), **BASE)
    u = np.zeros(c.shape); v = np.zeros(c.shape); w = np.zeros(c.shape)
    u[0, 0, 0] = -40.0          # a ghost corner sets umax
    v[3, 4, 5] = 7.0
    c.set_fields(u=u, v=v, w=w)
    c.call("computeTimestep")
    g = c.d.grid
    bound = 0.5 * c.d.re / (1 / g.dx ** 2 + 1 / g.dy ** 2 + 1 / g.dz ** 2)
    expect = 0.5 * min(bound, g.dx / 40.0, g.dy / 7.0)  # w == 0: DBL_MIN seed, no limit
    assert c.d.dt == pytest.approx(expect, rel=1e-15)


@pytest.mark.parametrize("gamma", [0.0, 1.0, 0.6])
def test_fg_donor_cell_on_linear_u(gamma):
    """u = a x + b (v = w = 0): Laplacian 0; d(u^2)/dx central = 2 a u,
    donor-cell correction -gamma a^2 dx / 2 where u > a dx (Griebel et al.)."""
    prm = dict(BASE, gamma=gamma, gx=0.3, gy=-0.1, gz=0.2)
    c = A.Case(lib("orc_rb"), **prm)
    X, _, _ = centres(c)
    g = c.d.grid
    a, b = 0.8, 3.0
    U = a * (X + g.dx / 2) + b
    c.set_fields(u=U, v=np.zeros(c.shape), w=np.zeros(c.shape))
    c.call("computeFG")
    f = c.fields()
    dt = c.d.dt
    Ui = interior(U)
    expect_f = Ui + dt * (-(2 * a * Ui - gamma * a * a * g.dx / 2) + 0.3)
    got = interior(f["f"])[:, :, :-1]  # F(imax) is the wall value
    np.testing.assert_allclose(got, expect_f[:, :, :-1], rtol=1e-12)
    np.testing.assert_array_equal(interior(f["f"])[:, :, -1], Ui[:, :, -1])
    np.testing.assert_allclose(interior(f["g"])[:, :-1, :], dt * -0.1, rtol=1e-10)
    np.testing.assert_allclose(interior(f["h"])[:-1, :, :], dt * 0.2, rtol=1e-10)


def test_neumann_eigenfunction_is_fixed_point_of_sor():
    """p = cos(pi x) cos(pi y / Ly) cos(pi z / Lz) at cell centres satisfies the
    ghost copy exactly; with rhs = Lap_h p (closed-form eigenvalue) every
    defect is zero and SOR leaves p unchanged."""
    c = A.Case(lib("orc_rb"), **dict(BASE, omg=1.7, itermax=3))
    X, Y, Z = centres(c)
    g = c.d.grid
    L = (g.xlength, g.ylength, g.zlength)
    h = (g.dx, g.dy, g.dz)
    p = np.cos(math.pi * X / L[0]) * np.cos(math.pi * Y / L[1]) * np.cos(math.pi * Z / L[2])
    lam = sum(-4.0 / hh ** 2 * math.sin(math.pi * hh / (2 * ll)) ** 2 for hh, ll in zip(h, L))
    c.set_fields(p=p, rhs=lam * p)
    res = c.lib.solve(A.C.byref(c.s), c.d.p, c.d.rhs)
    np.testing.assert_allclose(interior(c.fields()["p"]), interior(p), atol=1e-12)
    n = g.imax * g.jmax * g.kmax
    # carried residual: 1.0 / N after one iteration with zero defects, then 1/N^2 ...
    assert res == pytest.approx(((1.0 / n) / n) / n, rel=1e-6)


def test_sor_converges_to_manufactured_solution():
    c = A.Case(lib("orc_rb"), **dict(BASE, omg=1.6, itermax=4000, eps=1e-14))
    X, Y, Z = centres(c)
    g = c.d.grid
    L = (g.xlength, g.ylength, g.zlength)
    h = (g.dx, g.dy, g.dz)
    p = np.cos(2 * math.pi * X / L[0]) * np.cos(math.pi * Z / L[2])
    lam = sum(-4.0 / hh ** 2 * math.sin(math.pi * m * hh / (2 * ll)) ** 2
              for hh, ll, m in zip(h, L, (2, 0, 1)))
    c.set_fields(p=np.zeros(c.shape), rhs=lam * p)
    c.lib.solve(A.C.byref(c.s), c.d.p, c.d.rhs)
    got = interior(c.fields()["p"])
    got = got - got.mean()
    np.testing.assert_allclose(got, interior(p), atol=1e-7)


def test_zero_fields_stay_zero():
    c = A.Case(lib("orc_rb"), **dict(BASE, name="none", itermax=5)).run(3)
    f = c.fields()
    for k in ("u", "v", "w", "p"):
        assert not np.any(f[k]), k

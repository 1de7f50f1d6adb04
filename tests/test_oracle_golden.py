"""The oracle against the reference's own outputs (golden fixtures written by
tests/golden/make_golden.py from the compiled reference, strict IEEE build).
Both are strict IEEE C, so agreement is exact."""
import numpy as np
import pytest

import make_golden as G
import nusif_api as A
from conftest import lib


def _run_case(libname, name, kind, prm, seed, funcs):
    case = A.Case(lib("%s_%s" % (libname, kind)), **prm)
    inputs = G.random_fields(case.shape, seed)
    case.set_fields(**inputs)
    out = {}
    for fn in funcs:
        if fn == "solve":
            out["res"] = case.lib.solve(A.C.byref(case.s), case.d.p, case.d.rhs)
        else:
            case.call(fn)
    out.update(case.fields())
    out["dt"] = case.d.dt
    out["iterations"] = case.iterations()
    case.close()
    return inputs, out


@pytest.mark.parametrize("case", G.kat_cases(), ids=lambda c: c[0])
def test_oracle_kat(case, kat):
    name, kind, prm, seed, funcs = case
    inputs, out = _run_case("orc", *case)
    for k in A.FIELDS:
        key = "%s/%s" % (name, k)
        expect = kat[key] if key in kat.files else inputs[k]
        np.testing.assert_array_equal(out[k], expect, err_msg=key)
    assert out["dt"] == float(kat[name + "/dt"])
    if name + "/res" in kat.files:
        assert out["res"] == float(kat[name + "/res"])


@pytest.mark.parametrize("name", list(G.TRAJ))
def test_oracle_trajectory(name, traj):
    kind, prm, n = G.TRAJ[name]
    case = A.Case(lib("orc_" + kind), **prm).run(n)
    f = case.fields()
    np.testing.assert_array_equal(np.array(case.trace), traj[name + "/trace"])
    for k in ("p", "u", "v", "w"):
        np.testing.assert_array_equal(f[k], traj[name + "/" + k], err_msg=k)
    case.close()


def test_fixture_has_quirk_bearing_content(kat, traj):
    """The fixtures exercise what only a reference run can pin: the lid only on
    1 <= i < imax, 1 <= k < kmax (ghost row keeps -U elsewhere), carried RB
    residual with itermax cut-off, canal inflow 2.0."""
    lid = kat["lid/u"]
    inputs = G.random_fields(lid.shape, 2)
    nk, nj, ni = (s - 2 for s in lid.shape)
    row = lid[:, nj + 1, :]
    assert np.array_equal(row[1:nk, 1:ni], 2.0 - inputs["u"][1:nk, nj, 1:ni])
    assert np.array_equal(row[nk, :], inputs["u"][nk, nj + 1, :])  # k = kmax untouched
    assert np.all(kat["inflow/u"][1:-1, 1:-1, 0] == 2.0)
    assert traj["c1_rb/trace"].shape == (100, 3)

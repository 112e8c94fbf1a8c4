"""The oracle against the golden fixtures written from the reference build:
bit-for-bit equality (both are unfused IEEE double code in the same order)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.golden import all_fixtures

FIXTURES = all_fixtures()


def oracle_params(meta_params):
    p = dict(meta_params)
    for k in ("scheme", "diff", "width"):
        p[k] = tuple(p[k])
    return O.Params(**p)


@pytest.mark.parametrize("fx", FIXTURES, ids=[f.name for f in FIXTURES])
def test_oracle_matches_reference_fixture(fx):
    assert len(FIXTURES) >= 10
    blocks = fx.blocks()
    oc = O.OracleCase(oracle_params(fx.params()), blocks)
    if fx.meta["mode"] == "rhs":
        oc.rhs()
        for b in range(len(blocks)):
            np.testing.assert_array_equal(oc.get_rhs(b), fx.out(b, "rhs"))
            np.testing.assert_array_equal(oc.get_prim(b), fx.out(b, "prim"))
            assert oc.dt_local(b) == float(fx.out(b, "dt_local"))
            np.testing.assert_array_equal(oc.max_eigen(b), fx.out(b, "max_eigen"))
    else:
        dts = oc.steps(fx.meta["nsteps"], fx.meta["cfl"])
        np.testing.assert_array_equal(dts, fx.arrays["out_dts"])
        for b in range(len(blocks)):
            np.testing.assert_array_equal(oc.get_q(b), fx.out(b, "q"))

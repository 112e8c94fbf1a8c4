"""The HIP path (through the C ABI) against the golden fixtures written from
the reference build: <= 1e-12 relative max-norm per conserved component
(tests/parity.py), or within the reference's own -O0/-O2 build spread on the
components of a fixture where that spread exceeds 1e-12, for both kernel
variants (reference operation order, and the contracted fast form)."""
import numpy as np
import pytest

from nuc3d_amd import solver
from tests.golden import all_fixtures
from tests.parity import check, relerr_components

pytestmark = pytest.mark.gpu
FIXTURES = all_fixtures()


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("fx", FIXTURES, ids=[f.name for f in FIXTURES])
def test_gpu_matches_reference_fixture(fx, exact):
    print("case: %s exact=%d" % (fx.name, exact), flush=True)
    pk = fx.params()
    for k in ("scheme", "diff", "width"):
        pk[k] = tuple(pk[k])
    if solver.CRWENO5 in pk["scheme"] and not solver.crweno5_available():
        pytest.skip("CRWENO5 line solver not in this build")
    params = solver.Params(exact=exact, **pk)
    blocks = fx.blocks()
    g = solver.Case(params, blocks)
    try:
        if fx.meta["mode"] == "rhs":
            g.rhs()
            for b in range(len(blocks)):
                got, want = g.get_rhs(b), fx.out(b, "rhs")
                err = check(got, want, case=fx.name, exact=exact)
                print("  block %d rhs parity ratio %.3f  per component %s" % (
                    b, err, " ".join("%.1e" % e for e in relerr_components(got, want))), flush=True)
                assert err <= 1.0
                np.testing.assert_allclose(g.max_eigen(b), fx.out(b, "max_eigen"), rtol=1e-14)
        else:
            g.steps(fx.meta["nsteps"], fx.meta["cfl"])
            for b in range(len(blocks)):
                got, want = g.get_q(b), fx.out(b, "q")
                err = check(got, want, case=fx.name, exact=exact)
                print("  block %d q parity ratio %.3f  per component %s" % (
                    b, err, " ".join("%.1e" % e for e in relerr_components(got, want))), flush=True)
                assert err <= 1.0
            dt = fx.arrays["out_dts"][-1]
            assert abs(g.dt() - dt) <= 1e-13 * dt
    finally:
        g.close()

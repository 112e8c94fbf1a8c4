"""The oracle against the reference compiled from its own sources
(oracle/_ref/ref_harness), on seeded cases beyond the fixtures.  Runs where
that build exists (this container); skipped elsewhere."""
import tempfile

import numpy as np
import pytest

from nuc3d_amd import cases
from oracle import oracle as O

pytestmark = [pytest.mark.ref,
              pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref absent")]


def compare(params, blocks, mode="rhs", nsteps=0, cfl=0.5):
    d = tempfile.mkdtemp(prefix="nuc3d_ref_")
    O.write_ref_case(d, params, blocks)
    ref = O.run_ref(d, len(blocks), mode, nsteps, cfl)
    oc = O.OracleCase(params, blocks)
    if mode == "rhs":
        oc.rhs()
    else:
        np.testing.assert_array_equal(oc.steps(nsteps, cfl), ref[0]["dts"])
    for b in range(len(blocks)):
        np.testing.assert_array_equal(oc.get_rhs(b), ref[b]["rhs"])
        np.testing.assert_array_equal(oc.get_q(b), ref[b]["q"])
        np.testing.assert_array_equal(oc.max_eigen(b), ref[b]["max_eigen"])


@pytest.mark.parametrize("seed", [21, 22])
@pytest.mark.parametrize("scheme", [O.WENO5JS, O.WENO5Z, O.CRWENO5])
def test_live_rhs_random_tgv(seed, scheme):
    rng = np.random.default_rng(seed)
    dims = tuple(int(x) for x in rng.integers(9, 19, 3))
    blocks = cases.taylor_green(dims=dims, noise=1e-2, seed=seed)
    compare(O.Params(model=O.NS, scheme=(scheme,) * 3), blocks)


def test_live_multiblock_two_ranks_crweno5_steps():
    spec = cases.taylor_green(dims=(24, 10, 9), noise=1e-3, seed=23)[0]
    compare(O.Params(model=O.NS, scheme=(O.CRWENO5,) * 3), cases.split_block_x(spec, 2), "steps", 2)


def test_live_crweno5_long_line_within_stack_limit():
    # the reference's Thomas arrays hold 256 rows (crweno5.cpp:69-72): lines up to 261 cells
    blocks = cases.shu_osher(260, 8, 8, seed=3)
    compare(O.Params(model=O.EULER, mach=1.0, scheme=(O.CRWENO5,) * 3), blocks, "steps", 1, 0.4)


def test_live_energy_fixup_path():
    # a cell with negative internal energy takes the e<0 fix-up of physicsModel.cpp:324-332
    b = cases.taylor_green(dims=(12, 10, 9), noise=1e-3, seed=24)[0]
    b.q[4, 4, 5, 6] = 0.2 * b.q[4, 4, 5, 6]
    b.q[1, 4, 5, 6] = 40.0 * b.q[0, 4, 5, 6] * (2 * np.pi / 12) ** 0
    p = O.Params(model=O.NS)
    oc = O.OracleCase(p, [b])
    oc.rhs()
    assert oc.fixups() >= 1
    compare(p, [b])


@pytest.mark.skipif(not O.ref_available(optimised=True), reason="-O2 reference build absent")
@pytest.mark.parametrize("name", ["shuosher_euler_ausmp_steps2", "shuosher_euler_crweno5_steps3"])
def test_reference_builds_agree_under_parity_measure(name):
    """The reference's -O0 and -O2 builds of the same sources differ by more
    than 1e-12 in the near-zero v/w momentum of the quasi-1D Shu-Osher flow,
    but agree to < 1e-12 with momentum measured as a vector."""
    from tests.golden import all_fixtures
    from tests.parity import relerr, relerr_components

    fx = [f for f in all_fixtures() if f.name == name][0]
    pk = fx.params()
    for k in ("scheme", "diff", "width"):
        pk[k] = tuple(pk[k])
    p = O.Params(**pk)
    outs = []
    for opt in (False, True):
        d = tempfile.mkdtemp(prefix="nuc3d_ref_")
        O.write_ref_case(d, p, fx.blocks())
        outs.append(O.run_ref(d, 1, "steps", fx.meta["nsteps"], fx.meta["cfl"], optimised=opt)[0]["q"])
    assert max(relerr_components(outs[1], outs[0])) > 1e-12
    assert relerr(outs[1], outs[0]) < 1e-13


def test_live_curvilinear_multiblock_two_ranks():
    """All nine metric terms non-zero (wavy periodic mesh), cut into two
    blocks exchanged over MPI."""
    spec = cases.taylor_green_wavy((24, 12, 10), seed=25)[0]
    compare(O.Params(model=O.NS, scheme=(O.WENO5Z,) * 3), cases.split_block_x(spec, 2), "steps", 2)

"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same inputs.  The oracle agrees bit-for-bit with the reference compiled from
its own sources (tests/test_oracle_ref_live.py, tests/golden/).

Bar: <= 1e-12 relative max-norm per conserved component, as defined in
tests/parity.py.
"""
import numpy as np
import pytest

from nuc3d_amd import cases, solver
from oracle import oracle as O
from tests.parity import check, relerr_components

pytestmark = pytest.mark.gpu


def both_params(**kw):
    """GPU parameters (exact variant unless `exact` is given) and oracle parameters."""
    return solver.Params(**kw), O.Params(**{k: v for k, v in kw.items() if k != "exact"})


def run_rhs(sp, op, blocks):
    g = solver.Case(sp, blocks)
    g.rhs()
    o = O.OracleCase(op, blocks)
    o.rhs()
    return g, o


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("model", [solver.EULER3D, solver.NAVIERSTOKES3D])
@pytest.mark.parametrize("scheme", [solver.WENO5JS, solver.WENO5Z])
def test_rhs_tgv(exact, model, scheme):
    print(f"case: rhs tgv 20x18x22 model={model} scheme={scheme} exact={exact}", flush=True)
    blocks = cases.taylor_green(dims=(20, 18, 22), noise=1e-3, seed=3)
    sp, op = both_params(model=model, scheme=(scheme,) * 3, exact=exact)
    g, o = run_rhs(sp, op, blocks)
    err = check(g.get_rhs(0), o.get_rhs(0))
    print("  parity ratio %.3f" % err, flush=True)
    assert err <= 1.0
    np.testing.assert_allclose(g.max_eigen(0), o.max_eigen(0), rtol=1e-14)
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
def test_steps_tgv_ns(exact):
    print(f"case: 3 steps NS weno5js tgv 16^3 exact={exact}", flush=True)
    blocks = cases.taylor_green(16, noise=1e-3, seed=4)
    sp, op = both_params(exact=exact)
    g = solver.Case(sp, blocks)
    o = O.OracleCase(op, blocks)
    g.steps(3, 0.5)
    dts = o.steps(3, 0.5)
    err = check(g.get_q(0), o.get_q(0))
    print("  parity ratio %.3f dt %.17g %.17g" % (err, g.dt(), dts[-1]), flush=True)
    assert err <= 1.0
    assert abs(g.dt() - dts[-1]) <= 1e-14 * dts[-1]
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
def test_multiblock_weno5z(exact):
    print("case: 3 blocks split in x, NS weno5z, 2 steps, exact=%d" % exact, flush=True)
    spec = cases.taylor_green(dims=(36, 12, 14), noise=1e-3, seed=5)[0]
    blocks = cases.split_block_x(spec, 3)
    sp, op = both_params(scheme=(solver.WENO5Z,) * 3, exact=exact)
    g = solver.Case(sp, blocks)
    o = O.OracleCase(op, blocks)
    g.steps(2, 0.5)
    o.steps(2, 0.5)
    for b in range(3):
        err = check(g.get_q(b), o.get_q(b))
        print("  block %d parity ratio %.3f" % (b, err), flush=True)
        assert err <= 1.0
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("scheme", [solver.WENO5JS, solver.WENO5Z])
@pytest.mark.parametrize("riemann", [solver.LF, solver.AUSM, solver.AUSMP])
def test_shu_osher_inflow_outflow(scheme, riemann, exact):
    print(f"case: shu-osher 48x10x8 euler scheme={scheme} riemann={riemann} exact={exact}", flush=True)
    blocks = cases.shu_osher(48, 10, 8, seed=2)
    sp, op = both_params(model=solver.EULER3D, mach=1.0, scheme=(scheme,) * 3, riemann=riemann, exact=exact)
    g = solver.Case(sp, blocks)
    o = O.OracleCase(op, blocks)
    g.steps(3, 0.4)
    o.steps(3, 0.4)
    err = check(g.get_q(0), o.get_q(0))
    print("  parity ratio %.3f" % err, flush=True)
    assert err <= 1.0
    g.close()


@pytest.mark.parametrize("faces", [
    [(-1, 2, None), (-1, 1, None), (-1, 2, None), (-1, 3, None), (0, 5, None), (0, 4, None)],
    [(-1, 3, None), (-1, 3, None), (-1, 0, [1.0, 0.1, 0.0, 0.0, 71.0]), (-1, 1, None), (-1, 2, None), (-1, 3, None)],
])
@pytest.mark.parametrize("exact", [1, 0])
def test_physical_faces_ns(faces, exact):
    print(f"case: NS walls/symmetry/inlet faces={faces} exact={exact}", flush=True)
    b = cases.taylor_green(dims=(16, 12, 10), noise=1e-3, seed=6)[0]
    b.faces = faces
    sp, op = both_params(t_wall=1.2, exact=exact)
    g, o = run_rhs(sp, op, [b])
    err = check(g.get_rhs(0), o.get_rhs(0))
    print("  rhs parity ratio %.3f" % err, flush=True)
    assert err <= 1.0
    g.close()
    g = solver.Case(sp, [b])
    o = O.OracleCase(op, [b])
    g.steps(2, 0.5)
    o.steps(2, 0.5)
    err = check(g.get_q(0), o.get_q(0))
    print("  steps parity ratio %.3f" % err, flush=True)
    assert err <= 1.0
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
def test_conservation_periodic_64(exact):
    """Periodic block: the cell sum of every RHS component telescopes to zero,
    so mass, momentum and energy are conserved to round-off."""
    print("case: conservation NS tgv 64^3 exact=%d" % exact, flush=True)
    blocks = cases.taylor_green(64, noise=1e-3, seed=7)
    g = solver.Case(solver.Params(exact=exact), blocks)
    q0 = g.get_q(0)
    g.steps(2, 0.5)
    q1 = g.get_q(0)
    for c in range(5):
        s0, s1 = q0[c].sum(), q1[c].sum()
        # momentum components share the scale of the momentum magnitude
        scale = np.abs(q0[c]).sum() if c in (0, 4) else np.abs(q0[1:4]).sum()
        print("  comp %d drift %.3e" % (c, abs(s1 - s0) / scale), flush=True)
        assert abs(s1 - s0) <= 1e-13 * scale
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("model", [solver.EULER3D, solver.NAVIERSTOKES3D])
def test_crweno5_rhs_tgv(exact, model):
    print(f"case: crweno5 rhs tgv 20x18x22 model={model} exact={exact}", flush=True)
    blocks = cases.taylor_green(dims=(20, 18, 22), noise=1e-3, seed=31)
    sp, op = both_params(model=model, scheme=(solver.CRWENO5,) * 3, exact=exact)
    g, o = run_rhs(sp, op, blocks)
    err = check(g.get_rhs(0), o.get_rhs(0))
    print("  parity ratio %.3f" % err, flush=True)
    assert err <= 1.0
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
def test_crweno5_long_lines_shu_osher(exact):
    """x-lines of 300 cells: past the 261-cell limit of the reference's stack
    arrays (crweno5.cpp:69-72); pinned to the oracle, which has no cap."""
    print("case: crweno5 shu-osher 300x8x8 euler 2 steps exact=%d" % exact, flush=True)
    blocks = cases.shu_osher(300, 8, 8, seed=4)
    sp, op = both_params(model=solver.EULER3D, mach=1.0, scheme=(solver.CRWENO5,) * 3, exact=exact)
    g = solver.Case(sp, blocks)
    o = O.OracleCase(op, blocks)
    g.steps(2, 0.5)
    o.steps(2, 0.5)
    # the reference is undefined on these lines: the components' spread is the
    # restatement's between libm squares (the -O0 build) and rounded squares
    with O.square_mode(O.SQ_EXACT):
        o2 = O.OracleCase(op, blocks)
        o2.steps(2, 0.5)
    spread = relerr_components(o2.get_q(0), o.get_q(0))
    err = check(g.get_q(0), o.get_q(0), exact=exact, spread=spread)
    print("  parity ratio %.3f (square spread %s)" % (err, " ".join("%.1e" % x for x in spread)), flush=True)
    assert err <= 1.0
    if exact:
        # against the restatement that squares as the GPU does, every value is
        # the same (the path calls no libm function but sqrt on Euler cases)
        np.testing.assert_array_equal(g.get_q(0), o2.get_q(0))
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
def test_mixed_schemes_per_direction(exact):
    print("case: x crweno5, y weno5z, z weno5js, NS, 2 steps exact=%d" % exact, flush=True)
    blocks = cases.taylor_green(dims=(18, 16, 14), noise=1e-3, seed=32)
    sp, op = both_params(scheme=(solver.CRWENO5, solver.WENO5Z, solver.WENO5JS), exact=exact)
    g = solver.Case(sp, blocks)
    o = O.OracleCase(op, blocks)
    g.steps(2, 0.5)
    o.steps(2, 0.5)
    err = check(g.get_q(0), o.get_q(0))
    print("  parity ratio %.3f" % err, flush=True)
    assert err <= 1.0
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("scheme", [solver.WENO5JS, solver.WENO5Z, solver.CRWENO5])
def test_curvilinear_multiblock(scheme, exact):
    """Wavy periodic mesh (every metric term non-zero) cut into 2 blocks, NS, 2 steps."""
    print("case: wavy mesh 2 blocks NS scheme=%d exact=%d" % (scheme, exact), flush=True)
    spec = cases.taylor_green_wavy((28, 16, 14), seed=33)[0]
    blocks = cases.split_block_x(spec, 2)
    sp, op = both_params(scheme=(scheme,) * 3, exact=exact)
    g = solver.Case(sp, blocks)
    o = O.OracleCase(op, blocks)
    g.steps(2, 0.5)
    o.steps(2, 0.5)
    for b in range(2):
        err = check(g.get_q(b), o.get_q(b))
        print("  block %d parity ratio %.3f" % (b, err), flush=True)
        assert err <= 1.0
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("walls", [False, True])
def test_central8_two_blocks(walls, exact):
    """The eighth-order central scheme (the smooth-region scheme of the learned
    hybrid, four ghost layers): interior and halo faces central, physical faces
    the reference's closures; NS, 2 blocks, 2 steps."""
    print("case: central8 NS 2 blocks walls=%s exact=%d" % (walls, exact), flush=True)
    spec = cases.taylor_green(dims=(24, 14, 12), noise=1e-3, seed=34)[0]
    if walls:
        spec.faces = [(-1, 2, None), (-1, 2, None), (-1, 3, None), (-1, 3, None), (0, 5, None), (0, 4, None)]
    blocks = cases.split_block_x(spec, 2)
    sp, op = both_params(scheme=(solver.CENTRAL8,) * 3, width=(4, 4, 4), t_wall=1.2, exact=exact)
    g = solver.Case(sp, blocks)
    o = O.OracleCase(op, blocks)
    g.steps(2, 0.5)
    o.steps(2, 0.5)
    for b in range(2):
        err = check(g.get_q(b), o.get_q(b))
        print("  block %d parity ratio %.3f" % (b, err), flush=True)
        assert err <= 1.0
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
def test_central8_x_weno5js_yz(exact):
    """Central in x only (four ghost layers everywhere, WENO5-JS closures of
    width 3 in y and z): one RHS and 2 steps against the oracle."""
    print("case: central8 x, weno5js y z, Euler exact=%d" % exact, flush=True)
    blocks = cases.taylor_green(dims=(20, 16, 12), noise=1e-3, seed=35)
    sp, op = both_params(model=solver.EULER3D, scheme=(solver.CENTRAL8, solver.WENO5JS, solver.WENO5JS),
                         width=(4, 3, 3), exact=exact)
    g, o = run_rhs(sp, op, blocks)
    err = check(g.get_rhs(0), o.get_rhs(0))
    print("  rhs parity ratio %.3f" % err, flush=True)
    assert err <= 1.0
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("walls", [False, True])
def test_cd8_viscous_terms_two_blocks(walls, exact):
    """centraldifference8th at the stage level (centraldifference8th.cpp:22-192,
    routed by fieldOperator.cpp:508-520): the eighth-order viscous derivatives
    in the gradients and the divergence, four ghost layers, the reference's
    exterior closures at walls; with the central scheme (width 4), NS, 2
    blocks, 2 steps against the oracle."""
    print("case: central8 + cd8 NS 2 blocks walls=%s exact=%d" % (walls, exact), flush=True)
    spec = cases.taylor_green(dims=(24, 14, 12), noise=1e-3, seed=36)[0]
    if walls:
        spec.faces = [(-1, 2, None), (-1, 2, None), (-1, 3, None), (-1, 3, None), (0, 5, None), (0, 4, None)]
    blocks = cases.split_block_x(spec, 2)
    sp, op = both_params(scheme=(solver.CENTRAL8,) * 3, diff=(solver.CD8,) * 3, width=(4, 4, 4), t_wall=1.2,
                         exact=exact)
    g = solver.Case(sp, blocks)
    o = O.OracleCase(op, blocks)
    g.steps(2, 0.5)
    o.steps(2, 0.5)
    for b in range(2):
        err = check(g.get_q(b), o.get_q(b))
        print("  block %d parity ratio %.3f" % (b, err), flush=True)
        assert err <= 1.0
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
def test_cd8_in_x_only(exact):
    """cd8 in x (central8, width 4), cd6 in y and z (WENO5-JS, width 3): one
    RHS evaluation and 2 steps against the oracle."""
    print("case: x central8 + cd8, y z weno5js + cd6, NS exact=%d" % exact, flush=True)
    blocks = cases.taylor_green(dims=(22, 16, 12), noise=1e-3, seed=37)
    sp, op = both_params(scheme=(solver.CENTRAL8, solver.WENO5JS, solver.WENO5JS),
                         diff=(solver.CD8, solver.CD6, solver.CD6), width=(4, 3, 3), exact=exact)
    g, o = run_rhs(sp, op, blocks)
    err = check(g.get_rhs(0), o.get_rhs(0))
    print("  rhs parity ratio %.3f" % err, flush=True)
    assert err <= 1.0
    g.close()
    g = solver.Case(sp, blocks)
    o = O.OracleCase(op, blocks)
    g.steps(2, 0.5)
    o.steps(2, 0.5)
    err = check(g.get_q(0), o.get_q(0))
    print("  steps parity ratio %.3f" % err, flush=True)
    assert err <= 1.0
    g.close()


def test_cd8_needs_width_4():
    blocks = cases.taylor_green(dims=(16, 12, 10), noise=1e-3, seed=38)
    with pytest.raises(solver.Nuc3dError, match=r"\(-4\).*cd8 needs ghost width 4"):
        solver.Case(solver.Params(diff=(solver.CD8,) * 3), blocks)

"""The configurations over their whole horizons (200 steps for the
Taylor-Green vortex, 400 for Shu-Osher, 300 for the batch; 300 for the flat
plate and the turbulence, whose configurations name none), on the GPU at
their full sizes: every value finite, no positivity failure, and where every
block is periodic the cell sums of the conserved variables constant to
round-off.

Over such horizons the discrete flows amplify a last-bit difference by many
orders of magnitude (the WENO5 weights respond steeply to their smoothness
indicators): the reference's own -O0 and -O2 builds of the same sources
differ by 4e-5 after 20 steps and by 5e-2 after 40 on a 16^3 Navier-Stokes
Taylor-Green vortex.  So the comparison over the whole horizon is made on
small Euler blocks against the restatement evaluated with the GPU's squares
(correctly rounded x*x; the reference's -O0 build squares through libm pow),
which the exact variant equals bit for bit at every step."""
import numpy as np
import pytest

from nuc3d_amd import cases, solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _drifts(q0, q1):
    """Change of each conserved variable's cell sum (extended precision) over
    its scale; momentum on the scale of its magnitude before or after."""
    out = []
    for c in range(5):
        s0, s1 = q0[c].astype(np.longdouble).sum(), q1[c].astype(np.longdouble).sum()
        scale = np.abs(q0[c]).sum() if c in (0, 4) else max(np.abs(q0[1:4]).sum(), np.abs(q1[1:4]).sum())
        out.append(float(abs(s1 - s0)) / scale)
    return out


def _drift(q0, q1):
    return max(_drifts(q0, q1))


def test_tgv_200_steps():
    print("case: Taylor-Green 64^3 NS, 200 graph-launched steps; 16^3 against the oracle", flush=True)
    blocks = cases.taylor_green(64)
    g = solver.Case(solver.Params(), blocks)
    q0 = g.get_q(0)
    g.steps_graph(200, 0.5)
    g.sync()
    assert g.positivity() == (0, 0) and g.fixups() == 0
    q1 = g.get_q(0)
    g.close()
    assert np.isfinite(q1).all()
    d = _drift(q0, q1)
    print("  conservation drift %.3e" % d, flush=True)
    assert d <= 1e-13
    small = cases.taylor_green(16)
    g = solver.Case(solver.Params(model=solver.EULER3D, exact=1), small)
    g.steps_graph(200, 0.5)
    with O.square_mode(O.SQ_EXACT):
        o = O.OracleCase(O.Params(model=O.EULER), small)
        o.steps(200, 0.5)
    a, r = g.get_q(0), o.get_q(0)
    print("  16^3 Euler after 200 steps: %d values differ" % np.count_nonzero(a.view(np.int64) != r.view(np.int64)),
          flush=True)
    np.testing.assert_array_equal(a, r)
    g.close()


def test_shu_osher_400_steps():
    """CFL 0.4: at 0.5 the reference's own CRWENO5 loses positivity on this
    problem within 30 steps and ends in NaN; so does this implementation,
    which reports it (NUC3D_ERR_NEG_DENSITY) instead of running on."""
    print("case: Shu-Osher 512x64x64 Euler, WENO5-Z and CRWENO5, 400 graph-launched steps each, CFL 0.4", flush=True)
    blocks = cases.shu_osher(512, 64, 64, seed=0)
    jac = blocks[0].jac
    rho = {}
    for name, sch in (("weno5z", solver.WENO5Z), ("crweno5", solver.CRWENO5)):
        c = solver.Case(solver.Params(model=solver.EULER3D, mach=1.0, scheme=(sch,) * 3), blocks)
        c.steps_graph(400, 0.4)
        c.sync()
        assert c.positivity() == (0, 0)
        q = c.get_q(0)
        print("  %s: fix-ups %d" % (name, c.fixups()), flush=True)
        c.close()
        assert np.isfinite(q).all()
        ke = 0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / q[0] * jac
        p = 0.4 * (q[4] * jac - ke)
        assert (q[0] * jac).min() > 0.0 and p.min() > 0.0
        rho[name] = q[0] * jac
    # the two schemes resolve the entropy waves behind the shock differently;
    # the shock itself (the right-most cell of the y-z mean density above 2.5,
    # between the post-shock 3.86 and the pre-shock 1 +- 0.2) is at the same
    # place within a few cells (5 apart after 400 steps at 512 cells)
    front = {k: int(np.nonzero(v.mean(axis=(0, 1)) > 2.5)[0].max()) for k, v in rho.items()}
    diff = np.abs(rho["crweno5"] - rho["weno5z"])
    print("  shock front at cell %d (WENO5-Z), %d (CRWENO5); |rho_crweno5 - rho_weno5z| mean %.3e of mean rho" % (
        front["weno5z"], front["crweno5"], diff.mean() / rho["weno5z"].mean()), flush=True)
    assert 100 < front["weno5z"] < 500 and abs(front["crweno5"] - front["weno5z"]) <= 8
    small = cases.shu_osher(48, 10, 8, seed=2)
    for sch, osch in ((solver.WENO5Z, O.WENO5Z), (solver.CRWENO5, O.CRWENO5)):
        g = solver.Case(solver.Params(model=solver.EULER3D, mach=1.0, scheme=(sch,) * 3, exact=1), small)
        g.steps_graph(400, 0.4)
        with O.square_mode(O.SQ_EXACT):
            o = O.OracleCase(O.Params(model=O.EULER, mach=1.0, scheme=(osch,) * 3), small)
            o.steps(400, 0.4)
        a, r = g.get_q(0), o.get_q(0)
        print("  48x10x8 scheme %d after 400 steps: %d values differ" % (
            sch, np.count_nonzero(a.view(np.int64) != r.view(np.int64))), flush=True)
        np.testing.assert_array_equal(a, r)
        g.close()


def test_turbulence_256_300_steps():
    print("case: 256^3 turbulence NS, 300 graph-launched steps", flush=True)
    blocks, phys = cases.isotropic_turbulence(n=256, seed=0)
    g = solver.Case(solver.Params(mach=phys.mach, reynolds=phys.reynolds), blocks)
    q0 = g.get_q(0)
    g.steps_graph(300, 0.5)
    g.sync()
    assert g.positivity() == (0, 0)
    print("  fix-ups %d" % g.fixups(), flush=True)
    q1 = g.get_q(0)
    g.close()
    assert np.isfinite(q1).all()
    d = _drift(q0, q1)
    print("  conservation drift %.3e" % d, flush=True)
    assert d <= 1e-13


def test_flatplate_300_steps():
    blocks = cases.flat_plate_blocks()
    print("case: flat plate, 16 blocks, %d cells, 300 graph-launched steps" % sum(int(np.prod(b.dims)) for b in blocks),
          flush=True)
    g = solver.Case(solver.Params(mach=2.9, reynolds=2.0e4, t_wall=2.5), blocks)
    cases.setup_flat_plate(g, blocks)
    g.steps_graph(300, 0.5)
    g.sync()
    assert g.positivity() == (0, 0)
    print("  fix-ups %d" % g.fixups(), flush=True)
    for b in range(len(blocks)):
        q = g.get_q(b)
        assert np.isfinite(q).all()
        assert (q[0] * g.get_metrics(b)[0]).min() > 0.0
    g.close()


def test_batch48_hybrid_300_steps():
    from nuc3d_amd import train_sensor
    members = cases.batch_cases(n_cases=48, lo=32, hi=160, seed=0)
    blocks = [m.block for m in members]
    print("case: batch of 48, %d cells, learned hybrid with WENO5-JS fallback, 300 graph-launched steps at CFL 0.2"
          % sum(int(np.prod(b.dims)) for b in blocks), flush=True)
    g = solver.Case(solver.Params(scheme=(solver.HYBRID,) * 3, width=(4, 4, 4)), blocks)
    for i, m in enumerate(members):
        g.set_group(i, m.group, m.mach, m.reynolds)
    params, thr = train_sensor.load()
    g.set_sensor(params, thr)
    q0 = [g.get_q(b) for b in range(len(blocks))]
    g.steps_graph(300, 0.2)
    g.sync()
    assert g.positivity() == (0, 0)
    cen, net, fb = g.hybrid_counts()
    print("  fix-ups %d; central %d, network WENO %d, fallback %d" % (g.fixups(), cen, net, fb), flush=True)
    # the e<0 fix-up (physicsModel.cpp:322-332) replaces the energy of a cell,
    # so the energy of a case is conserved only where none was needed; mass
    # and momentum always are
    fixed = g.fixups()
    worst = [0.0] * 5
    for b in range(len(blocks)):
        q1 = g.get_q(b)
        assert np.isfinite(q1).all()
        worst = [max(x, y) for x, y in zip(worst, _drifts(q0[b], q1))]
    print("  worst per-case conservation drift per component %s" % " ".join("%.1e" % x for x in worst), flush=True)
    assert max(worst[:4]) <= 1e-13
    assert worst[4] <= 1e-13 or fixed > 0
    g.close()

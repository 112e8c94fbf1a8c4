"""Properties that do not depend on the size, checked at the benchmark's full
sizes (the oracle is too slow there): 256^3 decaying isotropic turbulence
(Navier-Stokes) and the 512x64x64 Shu-Osher problem (Euler, CRWENO5 and
WENO5-Z).

- Periodic blocks: the cell sums of the conserved variables stay constant to
  round-off (the flux-difference form telescopes), plain and hybrid.
- Graph-launched steps equal eagerly launched ones, and the uniform-mesh path
  equals the field loads, bit for bit.
- Shu-Osher: the state stays finite with positive density and pressure, and
  CRWENO5 and WENO5-Z stay close to each other."""
import numpy as np
import pytest

from nuc3d_amd import cases, solver

pytestmark = pytest.mark.gpu


def _drift_ok(q0, q1, tol=1e-14):
    # sums in extended precision: a float64 sum of 16.7 M terms rounds at the
    # 1e-14 level itself
    for c in range(5):
        s0, s1 = q0[c].astype(np.longdouble).sum(), q1[c].astype(np.longdouble).sum()
        scale = np.abs(q0[c]).sum() if c in (0, 4) else np.abs(q0[1:4]).sum()
        drift = float(abs(s1 - s0)) / scale
        print("  comp %d drift %.3e" % (c, drift), flush=True)
        assert drift <= tol


@pytest.fixture(scope="module")
def turbulence():
    return cases.isotropic_turbulence(n=256, seed=0)


def test_turbulence_256_conservation_graph_and_uniform(turbulence):
    blocks, phys = turbulence
    p = solver.Params(mach=phys.mach, reynolds=phys.reynolds)
    print("case: 256^3 turbulence NS, 2 graph steps; eager and field-load copies", flush=True)
    g = solver.Case(p, blocks)
    assert g.uniform_mesh(0)
    q0 = g.get_q(0)
    g.steps_graph(2, 0.5)
    qg = g.get_q(0)
    g.close()
    _drift_ok(q0, qg)
    e = solver.Case(p, blocks)
    e.set_uniform_mesh(False)
    assert not e.uniform_mesh(0)
    e.steps(2, 0.5)
    np.testing.assert_array_equal(e.get_q(0), qg)
    e.close()


def test_turbulence_256_hybrid_conservation(turbulence):
    from nuc3d_amd import train_sensor
    blocks, phys = turbulence
    params, thr = train_sensor.load()
    p = solver.Params(mach=phys.mach, reynolds=phys.reynolds, scheme=(solver.HYBRID,) * 3, width=(4, 4, 4))
    print("case: 256^3 turbulence NS, learned hybrid, 2 graph steps", flush=True)
    h = solver.Case(p, blocks)
    h.set_sensor(params, thr)
    h.set_sensor_interval(2)
    q0 = h.get_q(0)
    h.steps_graph(2, 0.5)
    q1 = h.get_q(0)
    cen, net, fb = h.hybrid_counts()
    h.close()
    print("  central %d, network WENO %d, fallback %d" % (cen, net, fb), flush=True)
    assert cen > 0 and (cen + net + fb) % (3 * 257 * 256 * 256) == 0
    _drift_ok(q0, q1)


@pytest.mark.parametrize("exact", [1, 0])
def test_shu_osher_512_crweno5_and_weno5z(exact):
    blocks = cases.shu_osher(512, 64, 64, seed=0)
    print("case: Shu-Osher 512x64x64 Euler, CRWENO5 and WENO5-Z, 3 steps, exact=%d" % exact, flush=True)
    out = {}
    for name, sch in (("crweno5", solver.CRWENO5), ("weno5z", solver.WENO5Z)):
        c = solver.Case(solver.Params(model=solver.EULER3D, mach=1.0, scheme=(sch,) * 3, exact=exact), blocks)
        c.steps_graph(2, 0.5)
        c.steps(1, 0.5)
        q = c.get_q(0)
        c.close()
        jac = blocks[0].jac
        rho = q[0] * jac
        ke = 0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / q[0] * jac
        p = 0.4 * (q[4] * jac - ke)
        assert np.isfinite(q).all()
        assert rho.min() > 0.0 and p.min() > 0.0
        out[name] = rho
    diff = np.abs(out["crweno5"] - out["weno5z"])
    d, m = diff.max() / out["weno5z"].max(), diff.mean() / out["weno5z"].mean()
    print("  |rho_crweno5 - rho_weno5z|: max %.3e of max rho, mean %.3e of mean rho" % (d, m), flush=True)
    assert d < 0.25 and m < 5e-3


@pytest.mark.parametrize("method", ["plain", "hybrid"])
def test_batch48_fullsize(method):
    """The 48 independent cases of seeded sizes 32^3..160^3 (57.0 M cells; 24
    Taylor-Green vortices, 24 spherical blasts, Navier-Stokes, one block group
    each), 20 graph-launched steps at CFL 0.2: every case stays finite, passes
    the positivity check and conserves mass, momentum and energy per case to
    round-off (every case is a periodic block)."""
    from nuc3d_amd import train_sensor
    members = cases.batch_cases(n_cases=48, lo=32, hi=160, seed=0)
    blocks = [m.block for m in members]
    print("case: batch of 48, %d cells, %s, 20 steps" % (sum(int(np.prod(b.dims)) for b in blocks), method),
          flush=True)
    kw = dict(scheme=(solver.HYBRID,) * 3, width=(4, 4, 4)) if method == "hybrid" else {}
    g = solver.Case(solver.Params(**kw), blocks)
    for i, m in enumerate(members):
        g.set_group(i, m.group, m.mach, m.reynolds)
    if method == "hybrid":
        params, thr = train_sensor.load()
        g.set_sensor(params, thr)
    q0 = [g.get_q(b) for b in range(len(blocks))]
    g.steps_graph(20, 0.2)
    g.sync()                                   # raises on a positivity failure
    assert g.positivity() == (0, 0)
    print("  fix-ups %d" % g.fixups(), flush=True)
    worst = 0.0
    for b in range(len(blocks)):
        q1 = g.get_q(b)
        assert np.isfinite(q1).all()
        for c in range(5):
            s0, s1 = q0[b][c].astype(np.longdouble).sum(), q1[c].astype(np.longdouble).sum()
            # momentum on the scale of its magnitude (the blasts start at rest)
            scale = np.abs(q0[b][c]).sum() if c in (0, 4) else max(np.abs(q0[b][1:4]).sum(), np.abs(q1[1:4]).sum())
            worst = max(worst, float(abs(s1 - s0)) / scale)
    print("  worst per-case conservation drift %.3e" % worst, flush=True)
    assert worst <= 1e-13
    g.close()


def test_flatplate_fullsize():
    """The 16-block supersonic flat plate over a bump (60.5 M cells,
    curvilinear tanh-stretched mesh with its metrics computed on the GPU,
    isothermal wall, inflow/outflow, z periodic), 20 graph-launched steps at
    CFL 0.5: finite, with positive density and pressure everywhere."""
    blocks = cases.flat_plate_blocks()
    print("case: flat plate, 16 blocks, %d cells, 20 steps" % sum(int(np.prod(b.dims)) for b in blocks), flush=True)
    g = solver.Case(solver.Params(mach=2.9, reynolds=2.0e4, t_wall=2.5), blocks)
    cases.setup_flat_plate(g, blocks)
    g.steps_graph(20, 0.5)
    g.sync()
    assert g.positivity() == (0, 0)
    print("  fix-ups %d" % g.fixups(), flush=True)
    for b in range(len(blocks)):
        q = g.get_q(b)
        assert np.isfinite(q).all()
        jac = g.get_metrics(b)[0]
        rho = q[0] * jac
        assert rho.min() > 0.0
    g.close()

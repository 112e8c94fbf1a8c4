"""Persistent direction sweeps (nuc3d_set_persistent): the x and y sweeps
(and the Euler z sweep) of WENO5-JS / WENO5-Z directions are each one grid of
as many workgroups as the GPU holds, taking (block, tile) work items from
per-XCD device queues.  Which workgroup computes which tile does
not enter the arithmetic, so every case must equal the ordinary launches
(one workgroup per tile) bit for bit: state, dt, residual, fix-up and
hybrid decision counts, eagerly and inside captured step graphs."""
import numpy as np
import pytest

from nuc3d_amd import cases, solver

pytestmark = pytest.mark.gpu


def _run(make, steps):
    out = []
    for persistent in (False, True):
        c = make()
        c.set_persistent(persistent)
        steps(c)
        c.sync()
        res = c.residual()
        out.append(([c.get_q(b) for b in range(len(c.blocks))], c.dt(), res, c.fixups(), c.positivity(),
                    c.hybrid_counts() if solver.HYBRID in c.params.scheme else None))
        c.close()
    (qa, dta, ra, fa, pa, ha), (qb, dtb, rb, fb, pb, hb) = out
    for a, b in zip(qa, qb):
        np.testing.assert_array_equal(a, b)
    assert dta == dtb and fa == fb and pa == pb and ha == hb
    np.testing.assert_array_equal(np.asarray(ra[1]), np.asarray(rb[1]))
    assert ra[0] == rb[0]


def test_persistent_turbulence_128_many_tiles_per_workgroup():
    """128^3 Navier-Stokes: tens of thousands of tiles per sweep, several per
    resident workgroup; 32-face tiles for x and y, 16 for z."""
    print("case: persistent vs per-tile launches, turbulence 128^3 NS, graph + eager steps", flush=True)
    blocks, phys = cases.isotropic_turbulence(n=128, seed=5)

    def make():
        c = solver.Case(solver.Params(model=solver.NAVIERSTOKES3D, mach=phys.mach, reynolds=phys.reynolds), blocks)
        c.set_sweep_tiles(0, (32, 32, 16))
        return c

    _run(make, lambda c: (c.steps_graph(2, 0.5), c.steps(1, 0.5)))


def test_persistent_flat_plate_uneven_curvilinear_blocks():
    print("case: persistent vs per-tile launches, flat plate 6 uneven curvilinear blocks, walls", flush=True)
    blocks = cases.flat_plate([12, 16, 20], [8, 12], 16)
    phys = dict(mach=2.9, reynolds=2.0e4, t_wall=2.5)

    def make():
        c = solver.Case(solver.Params(**phys), blocks)
        cases.setup_flat_plate(c, blocks)
        return c

    _run(make, lambda c: (c.steps_graph(2, 0.5), c.steps(2, 0.5)))


def test_persistent_batch_hybrid():
    """A batch of uneven independent cases with the learned hybrid: the
    deciding and reusing sweeps, per-group dt."""
    from nuc3d_amd import train_sensor
    params, thr = train_sensor.load()
    members = cases.batch_cases(n_cases=4, lo=16, hi=28, seed=7)
    print("case: persistent vs per-tile launches, hybrid batch of %d" % len(members), flush=True)

    def make():
        c = solver.Case(solver.Params(scheme=(solver.HYBRID,) * 3, width=(4, 4, 4)), [m.block for m in members])
        for b, m in enumerate(members):
            c.set_group(b, m.group, m.mach, m.reynolds)
        c.set_sensor(params, thr - 1.0)
        c.set_sensor_interval(2)
        return c

    _run(make, lambda c: (c.steps(1, 0.5), c.steps_graph(2, 0.5), c.steps(1, 0.5)))


def test_persistent_euler_weno5z_crweno5_lines():
    """Euler Shu-Osher with inflow/outflow faces: WENO5-Z x and y sweeps on
    the queue beside CRWENO5 z lines (whose kernel is not a sweep)."""
    print("case: persistent vs per-tile launches, Shu-Osher 48x10x8 Euler, weno5z x/y, crweno5 z", flush=True)
    blocks = cases.shu_osher(48, 10, 8, seed=9)

    def make():
        return solver.Case(solver.Params(model=solver.EULER3D, mach=1.0,
                                         scheme=(solver.WENO5Z, solver.WENO5Z, solver.CRWENO5)), blocks)

    _run(make, lambda c: (c.steps(2, 0.4), c.steps_graph(2, 0.4)))


def test_persistent_euler_all_weno5z_ausm():
    """Euler with AUSM splitting and WENO5-Z in every direction: the z sweep,
    which applies the RK stage, is persistent too."""
    print("case: persistent vs per-tile launches, Shu-Osher 48x10x8 Euler AUSM weno5z x/y/z", flush=True)
    blocks = cases.shu_osher(48, 10, 8, seed=11)

    def make():
        return solver.Case(solver.Params(model=solver.EULER3D, mach=1.0, riemann=solver.AUSM,
                                         scheme=(solver.WENO5Z,) * 3), blocks)

    _run(make, lambda c: (c.steps(2, 0.4), c.steps_graph(2, 0.4)))

"""The shortened FP64 division of the exact kernels (nuc3d_device.h:
recip_nr / div_nr) gives the compiler's IEEE quotient bit for bit on its
operand range, and the kernels that use it give the same results as with the
IEEE division everywhere."""
import numpy as np
import pytest

from nuc3d_amd import cases, solver

pytestmark = pytest.mark.gpu


def test_selftest_division():
    n = 1 << 28
    print("case: %d hashed operand pairs x 7 quotients and a square root" % n, flush=True)
    for seed in (1, 2):
        assert solver.selftest_division(n, seed) == 0


def _run(params, blocks, ieee, nsteps, cfl):
    g = solver.Case(params, blocks)
    g.set_division(ieee)
    if solver.HYBRID in params.scheme:
        from nuc3d_amd import train_sensor
        sp, thr = train_sensor.load()
        g.set_sensor(sp, thr - 2.0)   # more WENO5-JS faces than at the trained threshold
    if nsteps:
        g.steps(nsteps, cfl)
        out = [g.get_q(b) for b in range(len(blocks))]
    else:
        g.rhs()
        out = [g.get_rhs(b) for b in range(len(blocks))]
    g.close()
    return out


CASES = {
    "tgv_ns_weno5js_rhs": (dict(), lambda: cases.taylor_green(dims=(20, 18, 22), noise=1e-3, seed=3), 0, 0.5),
    "tgv_ns_weno5z_steps3": (dict(scheme=(solver.WENO5Z,) * 3),
                             lambda: cases.taylor_green(16, noise=1e-3, seed=4), 3, 0.5),
    "shuosher_euler_weno5js_steps3": (dict(model=solver.EULER3D, mach=1.0),
                                      lambda: cases.shu_osher(48, 10, 8, seed=2), 3, 0.4),
    "shuosher_euler_weno5z_ausmp_steps3": (dict(model=solver.EULER3D, mach=1.0, scheme=(solver.WENO5Z,) * 3,
                                                riemann=solver.AUSMP),
                                           lambda: cases.shu_osher(48, 10, 8, seed=2), 3, 0.4),
    "walls_multiblock_ns_steps2": (dict(t_wall=1.2), lambda: _walls(), 2, 0.5),
    # CRWENO5 lines read the state; their Euler z lines, or with Navier-Stokes
    # the separate divergence pass, write it (and keep its range flag)
    "tgv_ns_crweno5_steps3": (dict(scheme=(solver.CRWENO5,) * 3),
                              lambda: cases.taylor_green(16, noise=1e-3, seed=6), 3, 0.5),
    "shuosher_euler_crweno5_steps3": (dict(model=solver.EULER3D, mach=1.0, scheme=(solver.CRWENO5,) * 3),
                                      lambda: cases.shu_osher(48, 10, 8, seed=3), 3, 0.4),
    "hybrid_ns_steps3": (dict(scheme=(solver.HYBRID,) * 3, width=(4, 4, 4)),
                         lambda: cases.taylor_green(dims=(20, 18, 16), noise=5e-2, seed=9), 3, 0.5),
    "hybrid_walls_ns_rhs": (dict(scheme=(solver.HYBRID,) * 3, width=(4, 4, 4), t_wall=1.2), lambda: _walls(), 0, 0.5),
    # split fluxes of ~1e-45 (and ~1e-160, whose smoothness indicators are
    # subnormal): below the range the WENO5-Z and fast-variant weights need,
    # so their workgroups take the IEEE division; the exact WENO5-JS weights
    # need only the upper bound and take the reduced one.  Either way the
    # results must be the same
    "tiny_state_ns_rhs": (dict(), lambda: _tiny(1e-45), 0, 0.5),
    "tinier_state_ns_rhs": (dict(), lambda: _tiny(1e-160), 0, 0.5),
    "tiny_state_ns_weno5z_rhs": (dict(scheme=(solver.WENO5Z,) * 3), lambda: _tiny(1e-45), 0, 0.5),
    # momenta below the range of the shared reciprocal of rho (subnormal, and
    # normal below 2^-400) at a quarter of the cells of a resting field, which
    # keeps them over the steps: every RK stage writes such cells, so every
    # reader of the state divides by rho with the IEEE division
    "resting_tiny_momenta_ns_steps3": (dict(), lambda: _resting_tiny_momenta(), 3, 0.5),
}


def _walls():
    spec = cases.taylor_green(dims=(24, 14, 12), noise=1e-3, seed=34)[0]
    spec.faces = [(-1, 2, None), (-1, 2, None), (-1, 3, None), (-1, 3, None), (0, 5, None), (0, 4, None)]
    return cases.split_block_x(spec, 2)


def _tiny(scale):
    b = cases.taylor_green(dims=(16, 12, 10), noise=1e-3, seed=8)[0]
    b.q = b.q * scale
    return [b]


def _resting_tiny_momenta():
    """Uniform pressure, density with 1e-3 noise, no flow; the x and z momenta
    of a quarter of the cells set to +-[1e-320, 2e-308] (subnormal) or
    +-[1e-300, 1e-125] (normal, below 2^-400).  The inviscid momentum fluxes
    stay uniform to the last bit, so the momenta change only through the
    viscous terms of their own size and stay out of range."""
    dims = (16, 12, 10)
    h = tuple(2.0 * np.pi / d for d in dims)
    phys = cases.Physics()
    rng = np.random.default_rng(23)
    shape = dims[::-1]
    rho = 1.0 + 1e-3 * rng.uniform(-1.0, 1.0, shape)
    p = np.full(shape, 1.0 / (phys.gamma * phys.mach * phys.mach))
    zero = np.zeros(shape)
    jac, met, dxyz = cases.cartesian_metrics(dims, h)
    q = cases.conserved_over_j(rho, zero, zero, zero, p, phys.gamma, jac)
    for c in (1, 3):
        pick = rng.uniform(size=shape) < 0.25
        sub = rng.uniform(size=shape) < 0.5
        mag = np.where(sub, 10.0 ** rng.uniform(-320, np.log10(2e-308), shape), 10.0 ** rng.uniform(-300, -125, shape))
        q[c] = np.where(pick, np.where(rng.uniform(size=shape) < 0.5, -mag, mag), 0.0)
    assert (np.abs(q[1]) < 2.0 ** -400).all() and (np.abs(q[1][q[1] != 0.0]) < 2.3e-308).any()
    return [cases.BlockSpec(dims, cases.periodic_faces(), jac, met, dxyz, q)]


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("name", sorted(CASES))
def test_reduced_division_bit_identical(name, exact):
    kw, make, nsteps, cfl = CASES[name]
    print("case: %s exact=%d" % (name, exact), flush=True)
    params = solver.Params(exact=exact, **kw)
    blocks = make()
    a = _run(params, blocks, False, nsteps, cfl)
    b = _run(params, blocks, True, nsteps, cfl)
    for x, y in zip(a, b):
        assert np.isfinite(x).all()
        np.testing.assert_array_equal(x.view(np.int64), y.view(np.int64))


def test_range_flags_across_graphs_and_uploads():
    """The range flags of the state buffers (set by one RK stage for the
    buffer the next one writes, cleared by a cell out of range, zeroed by a
    refresh) through captured and eager steps and through new states uploaded
    between steps: in range, then momenta out of range, then in range again.
    The same bits as the IEEE division throughout."""
    print("case: TGV 16x12x10 NS, graph and eager steps across uploads of in- and out-of-range states", flush=True)
    blocks = cases.taylor_green(dims=(16, 12, 10), noise=1e-3, seed=31)
    tiny = _resting_tiny_momenta()[0]
    outs = []
    for ieee in (False, True):
        g = solver.Case(solver.Params(), blocks)
        g.set_division(ieee)
        g.steps_graph(2, 0.5)
        g.steps(1, 0.5)
        a = g.get_q(0)
        g.set_q(0, tiny.q)
        g.steps_graph(2, 0.5)
        b = g.get_q(0)
        g.set_q(0, a)
        g.steps(1, 0.5)
        g.steps_graph(4, 0.5)
        c = g.get_q(0)
        g.sync()
        outs.append((a, b, c))
        g.close()
    for x, y in zip(*outs):
        assert np.isfinite(x).all()
        np.testing.assert_array_equal(x.view(np.int64), y.view(np.int64))

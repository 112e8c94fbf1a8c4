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
    print("case: %d hashed operand pairs x 7 quotients" % n, flush=True)
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
}


def _walls():
    spec = cases.taylor_green(dims=(24, 14, 12), noise=1e-3, seed=34)[0]
    spec.faces = [(-1, 2, None), (-1, 2, None), (-1, 3, None), (-1, 3, None), (0, 5, None), (0, 4, None)]
    return cases.split_block_x(spec, 2)


def _tiny(scale):
    b = cases.taylor_green(dims=(16, 12, 10), noise=1e-3, seed=8)[0]
    b.q = b.q * scale
    return [b]


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

"""Diagnostics (postproc::solveQ: vorticity, enstrophy density, Q-criterion,
enstrophy and kinetic-energy sums) on the GPU against the oracle, whose sums
are pinned to the reference's enstrophy.dat."""
import numpy as np
import pytest

from nuc3d_amd import cases, solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def check(g, o, nb):
    for b in range(nb):
        e, k, f = g.diagnostics(b)
        eo, ko, fo = o.diagnostics(b)
        print("  block %d enstrophy %.15e vs %.15e, kinetic %.15e vs %.15e" % (b, e, eo, k, ko), flush=True)
        assert abs(e - eo) <= 1e-12 * abs(eo)
        assert abs(k - ko) <= 1e-12 * abs(ko)
        for c in range(5):
            err = np.abs(f[c] - fo[c]).max() / np.abs(fo[c]).max()
            assert err <= 1e-12, (c, err)


@pytest.mark.parametrize("steps", [0, 2])
def test_diagnostics_curvilinear_two_blocks(steps):
    print("case: diagnostics wavy 2 blocks NS, %d steps" % steps, flush=True)
    spec = cases.taylor_green_wavy((24, 16, 12), seed=61)[0]
    blocks = cases.split_block_x(spec, 2)
    g = solver.Case(solver.Params(), blocks)
    o = O.OracleCase(O.Params(model=O.NS), blocks)
    if steps:
        g.steps(steps, 0.5)
        o.steps(steps, 0.5)
    check(g, o, 2)
    g.close()


def test_diagnostics_euler_physical_faces():
    print("case: diagnostics Shu-Osher 40x8x8 Euler", flush=True)
    blocks = cases.shu_osher(40, 8, 8, seed=62)
    p = dict(model=solver.EULER3D, mach=1.0)
    g = solver.Case(solver.Params(**p), blocks)
    o = O.OracleCase(O.Params(**p), blocks)
    g.steps(1, 0.4)
    o.steps(1, 0.4)
    e, k, _ = g.diagnostics(0)
    eo, ko, _ = o.diagnostics(0)
    assert abs(k - ko) <= 1e-12 * abs(ko)
    assert abs(e - eo) <= 1e-10 * max(abs(eo), 1e-300) + 1e-300
    g.close()

"""Positivity failures reported as a status (the reference exits on a negative
pressure, physicsModel.cpp:1009-1016).  The failing states are injected from
the host into an ordinary case: the kernels run as always, every cell is
computed, and the synchronising calls return NUC3D_ERR_NEG_DENSITY /
NUC3D_ERR_NEG_PRESSURE.  The oracle, run on the same state, meets the same
cells at its con2prim."""
import numpy as np
import pytest

from nuc3d_amd import cases, solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NEG_DENSITY, NEG_PRESSURE = -8, -9


def injected(kind):
    b = cases.taylor_green(dims=(16, 12, 10), noise=1e-3, seed=41)[0]
    if kind == "density":
        b.q[0, 4, 5, 6] = -b.q[0, 4, 5, 6]
    else:
        b.q[4, 4, 5, 6] = np.nan
    return b


@pytest.mark.parametrize("kind,code", [("density", NEG_DENSITY), ("pressure", NEG_PRESSURE)])
@pytest.mark.parametrize("exact", [1, 0])
def test_positivity_status(kind, code, exact):
    print("case: NS TGV 16x12x10, injected %s failure, exact=%d" % (kind, exact), flush=True)
    b = injected(kind)
    g = solver.Case(solver.Params(exact=exact), [b])
    assert g.positivity() == (0, 0)
    # the uploaded state alone (its fix-up and check run when the case is
    # refreshed, which the uniform-mesh query does): one cell
    g.uniform_mesh(0)
    with pytest.raises(solver.Nuc3dError, match=r"\(%d\)" % code):
        g.sync()
    assert g.positivity() == ((1, 0) if kind == "density" else (0, 1))
    g.stage(0, 0.5)
    # a stage spreads a NaN into its neighbours' density as well: density
    # failures are reported first
    with pytest.raises(solver.Nuc3dError, match=r"\((%d|%d)\)" % (code, NEG_DENSITY)):
        g.sync()
    counts = g.positivity()
    print("  GPU counts (density, pressure) after one stage:", counts, flush=True)
    assert counts[0 if kind == "density" else 1] >= 1
    o = O.OracleCase(O.Params(), [b])
    o.rhs()
    ocounts = o.positivity()
    print("  oracle counts at con2prim:", ocounts, flush=True)
    assert ocounts[0 if kind == "density" else 1] >= 1
    # a healthy state clears the status (between full steps)
    g.stage(1, 0.5)
    g.stage(2, 0.5)
    with pytest.raises(solver.Nuc3dError, match=r"\((%d|%d)\)" % (code, NEG_DENSITY)):
        g.dt()
    g.set_q(0, cases.taylor_green(dims=(16, 12, 10), noise=1e-3, seed=41)[0].q)
    g.steps(1, 0.5)
    g.sync()
    assert g.positivity() == (0, 0)
    g.close()


def test_healthy_runs_report_nothing():
    blocks = cases.shu_osher(48, 10, 8, seed=2)
    g = solver.Case(solver.Params(model=solver.EULER3D, mach=1.0), blocks)
    g.steps(3, 0.4)
    g.sync()
    assert g.positivity() == (0, 0)
    g.close()

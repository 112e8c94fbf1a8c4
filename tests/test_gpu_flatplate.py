"""The multiblock supersonic flat plate over a Gaussian bump (curvilinear
mesh, tanh wall-normal stretching, isothermal wall, inflow/outflow, z
periodic, blocks of uneven sizes joined face to face): metrics computed on
the GPU from the mesh nodes, then two Navier-Stokes steps against the oracle
run on the same metrics and state."""
import numpy as np
import pytest

from nuc3d_amd import cases, solver
from oracle import oracle as O
from tests.parity import check

pytestmark = pytest.mark.gpu

PHYS = dict(mach=2.9, reynolds=2.0e4, t_wall=2.5)


@pytest.mark.parametrize("exact", [1, 0])
def test_flat_plate_small_against_oracle(exact):
    blocks = cases.flat_plate([12, 16, 20], [8, 12], 16)
    print("case: flat plate, %d blocks %s, exact=%d" % (len(blocks), [b.dims for b in blocks], exact), flush=True)
    g = solver.Case(solver.Params(exact=exact, **PHYS), blocks)
    cases.setup_flat_plate(g, blocks)
    oblocks = []
    for b, spec in enumerate(blocks):
        jac, met, dxyz = g.get_metrics(b)
        assert (jac > 0).all()
        oblocks.append(O.Block(spec.dims, spec.faces, jac, met, dxyz, g.get_q(b)))
    o = O.OracleCase(O.Params(**PHYS), oblocks)
    g.steps(2, 0.5)
    o.steps(2, 0.5)
    for b in range(len(blocks)):
        err = check(g.get_q(b), o.get_q(b), case="flatplate_small_steps2", exact=exact)
        print("  block %d parity ratio %.3f" % (b, err), flush=True)
        assert err <= 1.0
    assert np.isfinite(g.get_q(0)).all()
    g.close()


def test_flat_plate_hybrid_runs():
    """The learned hybrid on the flat plate: walls and inflow keep the
    reference's closures, the run stays finite and near plain WENO5-JS."""
    from nuc3d_amd import train_sensor
    params, thr = train_sensor.load()
    blocks = cases.flat_plate([16, 20], [12, 12], 20)
    print("case: flat plate hybrid, %d blocks" % len(blocks), flush=True)
    p = solver.Case(solver.Params(exact=0, **PHYS), blocks)
    h = solver.Case(solver.Params(exact=0, scheme=(solver.HYBRID,) * 3, width=(4, 4, 4), **PHYS), blocks)
    cases.setup_flat_plate(p, blocks)
    cases.setup_flat_plate(h, blocks)
    h.set_sensor(params, thr)
    p.steps(10, 0.5)
    h.steps(10, 0.5)
    num = den = 0.0
    for b, spec in enumerate(blocks):
        jac = p.get_metrics(b)[0]
        rp, rh = p.get_q(b)[0] * jac, h.get_q(b)[0] * jac
        assert np.isfinite(rh).all()
        num += float(np.sum((rh - rp) ** 2))
        den += float(np.sum(rp ** 2))
    dev = np.sqrt(num / den)
    print("  L2 density deviation from plain WENO5-JS %.2e of rms; decisions %s" % (dev, h.hybrid_counts()), flush=True)
    assert dev <= 1e-2
    p.close()
    h.close()

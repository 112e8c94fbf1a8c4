"""Metrics from mesh nodes on the GPU (nuc3d_upload_mesh) against the oracle
and against the reference's own block::initial (golden fixture
mesh_wavy_tgv_ns_steps2): cell centres, coordinate derivatives, J and metric
terms bit for bit, then the run from those metrics to the parity bar."""
import numpy as np
import pytest

from nuc3d_amd import cases, solver
from oracle import oracle as O
from tests.golden import all_fixtures
from tests.parity import check

pytestmark = pytest.mark.gpu


def case_for(dims, faces=None):
    jac, met, dxyz = cases.cartesian_metrics(dims, (1.0, 1.0, 1.0))
    nx, ny, nz = dims
    q = cases.conserved_over_j(*(np.ones((nz, ny, nx)),) * 4, 100.0 * np.ones((nz, ny, nx)), 1.4, jac)
    q[1:4] = 0.0
    return cases.BlockSpec(dims, faces or cases.periodic_faces(), jac, met, dxyz, q)


@pytest.mark.parametrize("dims,amp", [((16, 12, 8), 0.15), ((8, 24, 16), 0.3)])
def test_mesh_metrics_match_oracle(dims, amp):
    print("case: mesh metrics dims=%s amp=%.2f" % (dims, amp), flush=True)
    nodes = cases.wavy_nodes(dims, amp)
    g = solver.Case(solver.Params(), [case_for(dims)])
    centre = g.upload_mesh(0, *nodes)
    jac, met, dxyz = g.get_metrics(0)
    oc, od, oj, om = O.mesh_metrics(*nodes)
    np.testing.assert_array_equal(centre, oc)
    np.testing.assert_array_equal(dxyz, od)
    np.testing.assert_array_equal(jac, oj)
    np.testing.assert_array_equal(met, om)
    g.close()


def test_mesh_fixture_from_reference_block_initial():
    fx = [f for f in all_fixtures() if f.name == "mesh_wavy_tgv_ns_steps2"][0]
    print("case: %s" % fx.name, flush=True)
    nodes = fx.arrays["b0_nodes"]
    pk = fx.params()
    for k in ("scheme", "diff", "width"):
        pk[k] = tuple(pk[k])
    blk = fx.blocks()[0]
    g = solver.Case(solver.Params(**pk), [case_for(blk.dims)])
    centre = g.upload_mesh(0, nodes[0], nodes[1], nodes[2])
    jac, met, dxyz = g.get_metrics(0)
    np.testing.assert_array_equal(centre, fx.arrays["b0_centre"])
    np.testing.assert_array_equal(jac, blk.jac)
    np.testing.assert_array_equal(met, blk.met)
    np.testing.assert_array_equal(dxyz, blk.dxyz)
    g.set_q(0, blk.q)   # the reference's Taylor-Green initial state (block::initialQ)
    g.steps(fx.meta["nsteps"], fx.meta["cfl"])
    err = check(g.get_q(0), fx.out(0, "q"))
    print("  q parity ratio %.3f" % err, flush=True)
    assert err <= 1.0
    g.close()


def test_mesh_errors():
    dims = (8, 8, 12)
    g = solver.Case(solver.Params(), [case_for(dims)])
    xn, yn, zn = cases.wavy_nodes(dims, 0.1)
    with pytest.raises(solver.Nuc3dError, match=r"\(-1\).*Jacobian"):
        g.upload_mesh(0, -xn, yn, zn)
    g.close()
    g = solver.Case(solver.Params(), [case_for((10, 8, 8))])
    xn, yn, zn = cases.wavy_nodes((10, 8, 8), 0.1)
    with pytest.raises(solver.Nuc3dError, match=r"\(-4\).*divisible by 4"):
        g.upload_mesh(0, xn, yn, zn)
    g.close()


def test_restart_file_read_by_the_reference(tmp_path):
    """A restart file written from the GPU state is read back by the
    reference's own block::initial -> inputQ_binary (block.cpp:92-95, 780-810)
    bit for bit; nuc3d_load_q restores the state and time."""
    import os
    dims = (16, 12, 8)
    nodes = cases.wavy_nodes(dims, 0.2)
    g = solver.Case(solver.Params(), [case_for(dims)])
    centre = g.upload_mesh(0, *nodes)
    jac, _, _ = g.get_metrics(0)
    x, y, z = centre
    rho = 1.0 + 0.1 * np.sin(x) * np.cos(y)
    q = cases.conserved_over_j(rho, np.sin(y), np.cos(z), 0.2 * np.sin(x), 71.0 + np.cos(x + z), 1.4, jac)
    g.set_q(0, q)
    g.steps(2, 0.5)
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "flowData"), exist_ok=True)
    path = os.path.join(d, "flowData", "NUC3d_step_2_ID_0.bin")
    g.save_q(0, path, 0.125)
    want = g.get_q(0)
    raw = np.fromfile(path, dtype=np.float64)
    assert raw.size == 1 + want.size and raw[0] == 0.125
    np.testing.assert_array_equal(raw[1:].reshape(want.shape), want)
    g2 = solver.Case(solver.Params(), [case_for(dims)])
    g2.upload_mesh(0, *nodes)
    assert g2.load_q(0, path) == 0.125
    np.testing.assert_array_equal(g2.get_q(0), want)
    g.close()
    g2.close()
    # the reference's reader (built here by __graft_entry__.build(); it travels
    # with the tree)
    assert os.path.exists(O.REF_BIN_METRICS), "oracle/_ref/ref_metrics missing: run __graft_entry__.build()"
    O.write_mesh_case(d, O.Params(model=O.NS), [nodes], [cases.periodic_faces()], start_step=2, tecplot=True)
    ref = O.run_ref_metrics(d, 1, output=True)[0]
    np.testing.assert_array_equal(ref["q"], want)
    # Tecplot files: the reference's block::initial wrote the geometry file and
    # block::Output the flow file of this state; the GPU case writes the same
    # files byte for byte (block.cpp:692-752, 857-922)
    mine = os.path.join(d, "mine")
    os.makedirs(mine, exist_ok=True)
    g2 = solver.Case(solver.Params(), [case_for(dims)])
    g2.upload_mesh(0, *nodes)
    g2.load_q(0, path)
    g2.write_tecplot(0, os.path.join(mine, "flow.dat"), *nodes)
    g2.write_tecplot(0, os.path.join(mine, "geo.dat"), *nodes, geo=True)
    g2.close()
    for ours, theirs in (("flow.dat", "NUC3d_step_2_id_0.dat"), ("geo.dat", "GEO__id_0.dat")):
        a = open(os.path.join(mine, ours)).read()
        b = open(os.path.join(d, "flowData", theirs)).read()
        print("  %s: %d lines, reference %s: %d lines" % (ours, a.count("\n"), theirs, b.count("\n")), flush=True)
        assert a == b

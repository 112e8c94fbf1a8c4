"""The uniform-mesh path: blocks whose J and metric terms are one value each
take them from their descriptor.  Results equal the path that loads the
fields, bit for bit; curvilinear blocks are not taken for uniform ones."""
import numpy as np
import pytest

from nuc3d_amd import cases, solver

pytestmark = pytest.mark.gpu


def _run(params, blocks, uniform, steps=2, graph=False):
    c = solver.Case(params, blocks)
    c.set_uniform_mesh(uniform)
    flags = [c.uniform_mesh(b) for b in range(len(blocks))]
    if graph:
        c.steps_graph(steps, 0.5)
    else:
        c.steps(steps, 0.5)
    q = [c.get_q(b) for b in range(len(blocks))]
    c.close()
    return flags, q


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("scheme", [solver.WENO5JS, solver.WENO5Z, solver.CRWENO5])
def test_uniform_blocks_equal_field_loads(scheme, exact):
    spec = cases.taylor_green(dims=(36, 20, 18), noise=1e-3, seed=41)[0]
    blocks = cases.split_block_x(spec, 2)
    p = solver.Params(scheme=(scheme,) * 3, exact=exact)
    print("case: uniform path, TGV NS 2 blocks, scheme %d exact %d" % (scheme, exact), flush=True)
    fu, qu = _run(p, blocks, True)
    fl, ql = _run(p, blocks, False)
    assert fu == [True, True] and fl == [False, False]
    for a, b in zip(qu, ql):
        np.testing.assert_array_equal(a, b)


def test_uniform_euler_inflow_outflow():
    blocks = cases.shu_osher(64, 12, 10, seed=5)
    p = solver.Params(model=solver.EULER3D, mach=1.0, scheme=(solver.WENO5Z,) * 3)
    print("case: uniform path, Shu-Osher 64x12x10 Euler", flush=True)
    fu, qu = _run(p, blocks, True, steps=3, graph=True)
    fl, ql = _run(p, blocks, False, steps=3, graph=True)
    assert fu == [True] and fl == [False]
    np.testing.assert_array_equal(qu[0], ql[0])


def test_curvilinear_blocks_are_not_uniform():
    spec = cases.taylor_green_wavy((28, 16, 14), seed=33)[0]
    blocks = cases.split_block_x(spec, 2)
    c = solver.Case(solver.Params(), blocks)
    assert [c.uniform_mesh(b) for b in range(2)] == [False, False]
    c.close()


def test_uniform_hybrid():
    from nuc3d_amd import train_sensor
    params, thr = train_sensor.load()
    blocks = cases.taylor_green(dims=(32, 24, 20), noise=1e-3, seed=42)
    p = solver.Params(scheme=(solver.HYBRID,) * 3, width=(4, 4, 4), exact=0)
    print("case: uniform path, hybrid TGV", flush=True)
    out = []
    for uniform in (True, False):
        c = solver.Case(p, blocks)
        c.set_sensor(params, thr)
        c.set_uniform_mesh(uniform)
        c.steps_graph(2, 0.5)
        out.append(c.get_q(0))
        c.close()
    np.testing.assert_array_equal(out[0], out[1])

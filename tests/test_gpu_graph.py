"""The graph-launched step (nuc3d_step_graph) against ordinary launches
(nuc3d_steps) on the same case: bit-identical state and dt after every
replay, with eager work (downloads, odd steps, a cfl change) between replays."""
import numpy as np
import pytest

from nuc3d_amd import cases, solver

pytestmark = pytest.mark.gpu


def pair(params, blocks):
    return solver.Case(params, blocks), solver.Case(params, blocks)


def same(a, b, nb):
    for blk in range(nb):
        np.testing.assert_array_equal(a.get_q(blk), b.get_q(blk))
    assert a.dt() == b.dt()


@pytest.mark.parametrize("exact", [1, 0])
def test_graph_steps_match_eager_tgv_ns(exact):
    print("case: graph vs eager, NS tgv 24x20x18, exact=%d" % exact, flush=True)
    blocks = cases.taylor_green(dims=(24, 20, 18), noise=1e-3, seed=51)
    p = solver.Params(exact=exact)
    eager, graph = pair(p, blocks)
    for n, cfl in ((2, 0.5), (1, 0.5), (4, 0.5), (3, 0.5), (2, 0.3), (2, 0.3), (2, 0.5)):
        eager.steps(n, cfl)
        graph.steps_graph(n, cfl)
        same(eager, graph, 1)
        print("  %d steps at cfl %.1f: identical, dt %.6e" % (n, cfl, graph.dt()), flush=True)
    eager.close()
    graph.close()


def test_graph_steps_match_eager_multiblock_mixed():
    print("case: graph vs eager, 2 blocks, crweno5 x / weno5z y / weno5js z, NS", flush=True)
    spec = cases.taylor_green(dims=(24, 12, 14), noise=1e-3, seed=52)[0]
    blocks = cases.split_block_x(spec, 2)
    p = solver.Params(scheme=(solver.CRWENO5, solver.WENO5Z, solver.WENO5JS))
    eager, graph = pair(p, blocks)
    for n in (2, 2, 1, 2):
        eager.steps(n, 0.5)
        graph.steps_graph(n, 0.5)
        same(eager, graph, 2)
    assert graph.fixups() == eager.fixups()
    eager.close()
    graph.close()


def test_graph_steps_shu_osher_boundaries():
    print("case: graph vs eager, Shu-Osher 48x10x8 Euler AUSM weno5z", flush=True)
    blocks = cases.shu_osher(48, 10, 8, seed=53)
    p = solver.Params(model=solver.EULER3D, mach=1.0, riemann=solver.AUSM, scheme=(solver.WENO5Z,) * 3)
    eager, graph = pair(p, blocks)
    for _ in range(3):
        eager.steps(2, 0.4)
        graph.steps_graph(2, 0.4)
        same(eager, graph, 1)
    eager.close()
    graph.close()


@pytest.mark.parametrize("graph", [True, False])
def test_overlapped_viscous_flux_matches_serial(graph):
    """The viscous-flux pass on its own stream beside the x and y sweeps (the
    default) equals the passes run in order, bit for bit, eagerly and inside
    captured graphs; three uneven blocks with walls (curvilinear flat plate)."""
    print("case: overlap on/off, flat plate 3 blocks, graph=%s" % graph, flush=True)
    blocks = cases.flat_plate([12, 16, 20], [8], 16)
    phys = dict(mach=2.9, reynolds=2.0e4, t_wall=2.5)
    out = []
    for overlap in (True, False):
        c = solver.Case(solver.Params(**phys), blocks)
        cases.setup_flat_plate(c, blocks)
        c.set_overlap(overlap)
        if graph:
            c.steps_graph(4, 0.5)
        else:
            c.steps(3, 0.5)
        out.append([c.get_q(b) for b in range(len(blocks))] + [c.dt()])
        c.close()
    for a, b in zip(out[0][:-1], out[1][:-1]):
        np.testing.assert_array_equal(a, b)
    assert out[0][-1] == out[1][-1]

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

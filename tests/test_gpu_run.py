"""The step residual (PDEData3d::setRES) and the host time loop
(singleBlock::solvePDE with the IOController cadence) on the GPU, against the
oracle, whose residual is pinned bit for bit to the reference's."""
import os

import numpy as np
import pytest

from nuc3d_amd import cases, solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def tgv_two_blocks(seed):
    spec = cases.taylor_green(dims=(20, 12, 10), noise=1e-3, seed=seed)[0]
    return cases.split_block_x(spec, 2)


@pytest.mark.parametrize("model", [solver.NAVIERSTOKES3D, solver.EULER3D])
def test_residual_matches_oracle(model):
    print("case: residual per step, 2 blocks, model=%d" % model, flush=True)
    blocks = tgv_two_blocks(71)
    g = solver.Case(solver.Params(model=model), blocks)
    o = O.OracleCase(O.Params(model=model), blocks)
    assert g.residual()[0] == 0.0
    for s in range(3):
        g.steps(1, 0.5)
        o.steps(1, 0.5)
        res, per = g.residual()
        print("  step %d residual %.15e vs %.15e" % (s + 1, res, o.residual()), flush=True)
        assert abs(res - o.residual()) <= 1e-9 * abs(o.residual())
        for b in range(2):
            assert abs(per[b] - o.residual(b)) <= 1e-9 * abs(o.residual(b))
    g.close()


def test_residual_graph_equals_eager():
    blocks = tgv_two_blocks(72)
    a, b = solver.Case(solver.Params(), blocks), solver.Case(solver.Params(), blocks)
    a.steps(4, 0.5)
    b.steps_graph(4, 0.5)
    assert a.residual()[0] == b.residual()[0]
    a.steps(1, 0.5)
    b.steps_graph(1, 0.5)
    assert a.residual()[0] == b.residual()[0]
    a.close()
    b.close()


def test_run_loop_cadence_and_outputs(tmp_path):
    print("case: run loop 4 steps, save every 2, post every step", flush=True)
    blocks = tgv_two_blocks(73)
    d = str(tmp_path)
    os.makedirs(os.path.join(d, "flowData"))
    g = solver.Case(solver.Params(), blocks)
    st = g.run(end_step=4, cfl=0.5, save_step=2, post_step=1, dir=d)
    o = O.OracleCase(O.Params(model=O.NS), blocks)
    lines = open(os.path.join(d, "enstrophy.dat")).read().split("\n")[:-1]
    assert len(lines) == 4
    t = 0.0
    for s in range(4):
        t += o.steps(1, 0.5)[0]
        step, time, ens, kin = lines[s].split()
        eo, ko, _ = o.diagnostics()
        assert int(step) == s + 1
        assert abs(float(time) - t) <= 1e-11 * t
        assert abs(float(ens) - eo) <= 1e-10 * eo and abs(float(kin) - ko) <= 1e-10 * ko
    assert st["step"] == 4 and abs(st["time"] - t) <= 1e-13 * t
    assert abs(st["residual"] - o.residual()) <= 1e-9 * abs(o.residual())
    for step in (2, 4):
        for b in range(2):
            assert os.path.exists(os.path.join(d, "flowData", "NUC3d_step_%d_ID_%d.bin" % (step, b)))
    raw = np.fromfile(os.path.join(d, "flowData", "NUC3d_step_4_ID_1.bin"))
    assert raw[0] == st["time"]
    np.testing.assert_array_equal(raw[1:], g.get_q(1).ravel())
    assert not os.path.exists(os.path.join(d, "flowData", "NUC3d_step_3_ID_0.bin"))
    # time limit: `time <= tl` runs exactly one more step from time 0 with tl = 0
    g2 = solver.Case(solver.Params(), blocks)
    st2 = g2.run(end_step=100, tl=0.0)
    assert st2["step"] == 1
    g.close()
    g2.close()

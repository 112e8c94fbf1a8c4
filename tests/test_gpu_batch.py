"""Batches of independent cases in one case (block groups): each group takes
its own dt and its own Mach and Reynolds numbers, and every group's result
equals the same case run alone, bit for bit, and the oracle run on that case
alone to the parity bar."""
import numpy as np
import pytest

from nuc3d_amd import cases, solver
from oracle import oracle as O
from tests.parity import check

pytestmark = pytest.mark.gpu


def _batch(seed=3):
    return cases.batch_cases(n_cases=4, lo=16, hi=28, seed=seed)


@pytest.mark.parametrize("exact", [1, 0])
def test_batch_equals_cases_run_alone(exact):
    members = _batch()
    print("case: batch of %d: %s" % (len(members), ", ".join("%s %d^3 Re %.0f" % (m.kind, m.block.dims[0], m.reynolds)
                                                            for m in members)), flush=True)
    p = solver.Params(exact=exact)
    g = solver.Case(p, [m.block for m in members])
    for b, m in enumerate(members):
        g.set_group(b, m.group, m.mach, m.reynolds)
    g.steps_graph(2, 0.5)
    g.steps(1, 0.5)
    for b, m in enumerate(members):
        blk = cases.BlockSpec(m.block.dims, cases.periodic_faces(0), m.block.jac, m.block.met, m.block.dxyz, m.block.q)
        alone = solver.Case(solver.Params(exact=exact, mach=m.mach, reynolds=m.reynolds), [blk])
        alone.steps(3, 0.5)
        np.testing.assert_array_equal(g.get_q(b), alone.get_q(0))
        assert g.dt_group(b) == alone.dt()
        alone.close()
    # the groups really take different dt (equal-sized blasts share theirs)
    assert len({g.dt_group(b) for b in range(len(members))}) > 1
    g.close()


def test_batch_against_oracle():
    members = _batch(seed=4)
    print("case: batch of %d against the oracle per case" % len(members), flush=True)
    g = solver.Case(solver.Params(), [m.block for m in members])
    for b, m in enumerate(members):
        g.set_group(b, m.group, m.mach, m.reynolds)
    g.steps(2, 0.5)
    for b, m in enumerate(members):
        blk = O.Block(m.block.dims, cases.periodic_faces(0), m.block.jac, m.block.met, m.block.dxyz, m.block.q)
        o = O.OracleCase(O.Params(mach=m.mach, reynolds=m.reynolds), [blk])
        o.steps(2, 0.5)
        err = check(g.get_q(b), o.get_q(0))
        print("  %s %d^3: parity ratio %.3f" % (m.kind, m.block.dims[0], err), flush=True)
        assert err <= 1.0
    g.close()


def test_batch_hybrid():
    """The learned hybrid over a batch: each group equals its case run alone."""
    from nuc3d_amd import train_sensor
    params, thr = train_sensor.load()
    members = _batch(seed=5)
    print("case: hybrid batch of %d" % len(members), flush=True)
    p = solver.Params(scheme=(solver.HYBRID,) * 3, width=(4, 4, 4), exact=0)
    g = solver.Case(p, [m.block for m in members])
    g.set_sensor(params, thr)
    for b, m in enumerate(members):
        g.set_group(b, m.group, m.mach, m.reynolds)
    g.steps_graph(2, 0.5)
    for b, m in enumerate(members):
        blk = cases.BlockSpec(m.block.dims, cases.periodic_faces(0), m.block.jac, m.block.met, m.block.dxyz, m.block.q)
        alone = solver.Case(solver.Params(scheme=(solver.HYBRID,) * 3, width=(4, 4, 4), exact=0, mach=m.mach,
                                          reynolds=m.reynolds), [blk])
        alone.set_sensor(params, thr)
        alone.steps(2, 0.5)
        np.testing.assert_array_equal(g.get_q(b), alone.get_q(0))
        alone.close()
    g.close()


def test_groups_cannot_be_connected():
    """Connected blocks of different groups are refused when work is launched;
    moving both blocks of a connected case to another group, one after the
    other, is allowed and changes nothing."""
    spec = cases.taylor_green(dims=(24, 12, 12), noise=1e-3, seed=1)[0]
    blocks = cases.split_block_x(spec, 2)
    g = solver.Case(solver.Params(), blocks)
    g.set_group(1, 1)
    with pytest.raises(solver.Nuc3dError, match="group"):
        g.steps(1, 0.5)
    g.close()
    ref = solver.Case(solver.Params(), blocks)
    ref.steps(2, 0.5)
    g = solver.Case(solver.Params(), blocks)
    g.set_group(0, 1)
    g.set_group(1, 1)
    g.steps(2, 0.5)
    for b in range(2):
        np.testing.assert_array_equal(g.get_q(b), ref.get_q(b))
    g.close()
    ref.close()


@pytest.mark.parametrize("exact", [1, 0])
def test_autotune_keeps_results(exact):
    """The autotuner only changes the k segments of the column tiles: steps
    after it equal steps of an untuned case bit for bit (uneven batch and a
    two-block connected case)."""
    members = cases.batch_cases(n_cases=3, lo=16, hi=40, seed=7)
    spec = cases.taylor_green(dims=(24, 20, 36), noise=1e-3, seed=8)[0]
    two = cases.split_block_x(spec, 2)
    for label, blocks, groups in (("batch", [m.block for m in members], members), ("two blocks", two, None)):
        print("case: autotune, %s, exact=%d" % (label, exact), flush=True)
        a = solver.Case(solver.Params(exact=exact), blocks)
        b = solver.Case(solver.Params(exact=exact), blocks)
        if groups:
            for i, m in enumerate(groups):
                a.set_group(i, m.group, m.mach, m.reynolds)
                b.set_group(i, m.group, m.mach, m.reynolds)
        ks = a.autotune()
        print("  k segments %s, sweep tiles %s" % (ks, [a.sweep_tiles(i) for i in range(len(blocks))]), flush=True)
        assert all(k in (4, 8, 16, 32, 64, 128) for k in ks)
        assert all(nf in (16, 32) for i in range(len(blocks)) for nf in a.sweep_tiles(i))
        a.steps_graph(2, 0.5)
        b.steps_graph(2, 0.5)
        for i in range(len(blocks)):
            np.testing.assert_array_equal(a.get_q(i), b.get_q(i))
        ea, ka, _ = a.diagnostics(0)
        eb, kb, _ = b.diagnostics(0)
        assert abs(ea - eb) <= 1e-12 * abs(eb) and abs(ka - kb) <= 1e-12 * abs(kb)
        a.close()
        b.close()


@pytest.mark.parametrize("exact", [1, 0])
def test_batch_crweno5_line_groups(exact):
    """CRWENO5 lines run in groups of six lines of one block (the last group
    of a block padded): uneven Euler cases in one batch equal the same cases
    run alone, bit for bit, and the oracle on one of them."""
    members = cases.batch_cases(n_cases=4, lo=16, hi=28, seed=9)
    print("case: CRWENO5 batch of %d: %s, exact=%d" % (
        len(members), ", ".join("%d^3" % m.block.dims[0] for m in members), exact), flush=True)
    p = solver.Params(model=solver.EULER3D, scheme=(solver.CRWENO5,) * 3, exact=exact)
    g = solver.Case(p, [m.block for m in members])
    for b, m in enumerate(members):
        g.set_group(b, m.group, m.mach, m.reynolds)
    g.steps(2, 0.2)
    for b, m in enumerate(members):
        blk = cases.BlockSpec(m.block.dims, cases.periodic_faces(0), m.block.jac, m.block.met, m.block.dxyz, m.block.q)
        alone = solver.Case(solver.Params(model=solver.EULER3D, scheme=(solver.CRWENO5,) * 3, exact=exact,
                                          mach=m.mach, reynolds=m.reynolds), [blk])
        alone.steps(2, 0.2)
        np.testing.assert_array_equal(g.get_q(b), alone.get_q(0))
        alone.close()
    m = members[0]
    blk = O.Block(m.block.dims, cases.periodic_faces(0), m.block.jac, m.block.met, m.block.dxyz, m.block.q)
    o = O.OracleCase(O.Params(model=O.EULER, mach=m.mach, scheme=(solver.CRWENO5,) * 3), [blk])
    o.steps(2, 0.2)
    err = check(g.get_q(0), o.get_q(0))
    print("  block 0 (%d^3) against the oracle: parity ratio %.3f" % (m.block.dims[0], err), flush=True)
    assert err <= 1.0
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("scheme", [solver.WENO5JS, solver.WENO5Z])
def test_sweep_tile_shapes_keep_results(scheme, exact):
    """32-face x 8-line sweep tiles give the same steps as 16 x 16 bit for bit,
    in every direction or mixed per block (uneven blocks of a batch, Navier-
    Stokes with the viscous z sweep, and an Euler Shu-Osher block with
    inflow/outflow faces)."""
    members = cases.batch_cases(n_cases=3, lo=16, hi=40, seed=9)
    blocks = [m.block for m in members]
    for label, bl, params, cfl, groups in (
            ("batch NS", blocks, solver.Params(exact=exact, scheme=(scheme,) * 3), 0.2, members),
            ("shu-osher Euler", cases.shu_osher(70, 10, 12, seed=3),
             solver.Params(model=solver.EULER3D, mach=1.0, exact=exact, scheme=(scheme,) * 3), 0.4, None)):
        print("case: sweep tiles, %s, scheme %d, exact=%d" % (label, scheme, exact), flush=True)
        runs = []
        for shapes in ([(16, 16, 16)] * len(bl), [(32, 32, 32)] * len(bl),
                       [((32, 16, 32), (16, 32, 16), (32, 32, 16))[i % 3] for i in range(len(bl))]):
            g = solver.Case(params, bl)
            if groups:
                for i, m in enumerate(groups):
                    g.set_group(i, m.group, m.mach, m.reynolds)
            for i, sh in enumerate(shapes):
                g.set_sweep_tiles(i, sh)
                assert g.sweep_tiles(i) == tuple(sh)
            g.steps_graph(2, cfl)
            g.steps(1, cfl)
            runs.append([g.get_q(i) for i in range(len(bl))])
            g.close()
        for other in runs[1:]:
            for x, y in zip(runs[0], other):
                np.testing.assert_array_equal(x, y)

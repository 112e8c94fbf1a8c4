"""GPU tests of the learned sensor and the hybrid reconstruction against the
CPU restatement (oracle/sensor.py, the oracle's ORC_HYBRID).

  - features and teacher labels sampled on the GPU: bit-exact;
  - network logits (BF16 MFMA, FP32 accumulation) and training gradients:
    within 1e-2 of the float64 restatement on the same bf16 operands;
  - hybrid RHS with the GPU's recorded decisions: the parity bar (1e-12);
  - the GPU's decisions equal the restatement's wherever its logit is not
    within 0.05 of the threshold;
  - all-WENO5 and all-central thresholds reproduce WENO5-JS and the central
    scheme bit for bit; periodic hybrid steps conserve to round-off.
"""
import numpy as np
import pytest

from nuc3d_amd import cases, solver
from oracle import oracle as O
from oracle import sensor as S
from tests.parity import check

pytestmark = pytest.mark.gpu


def hybrid_params(model=solver.NAVIERSTOKES3D, exact=1, **kw):
    sp = solver.Params(model=model, scheme=(solver.HYBRID,) * 3, width=(4, 4, 4), exact=exact, **kw)
    op = O.Params(model=model, scheme=(O.HYBRID,) * 3, width=(4, 4, 4), **kw)
    return sp, op


def mixed_params(seed):
    """Random parameters whose logits straddle zero on smooth flow."""
    p = S.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    p[S.OB1:S.OW2] = rng.normal(0, 0.2, 32)
    p[S.OB2:S.OW3] = rng.normal(0, 0.2, 32)
    return p


def test_sample_features_labels_bit_exact():
    print("case: sensor samples TGV 16^3 NS (width 3), all faces, 3 directions", flush=True)
    blocks = cases.taylor_green(16, noise=3e-2, seed=21)
    g = solver.Case(solver.Params(), blocks)
    s = solver.Sensor(100000)
    n = s.sample(g, tau=1e-3)
    x, y = s.download()
    assert n == 3 * 17 * 16 * 16
    off = 0
    b = blocks[0]
    for d in range(3):
        xo, yo = S.sample_block(b.q, b.jac, 1.4, 1e-3, d, periodic=True, ghosts=3)
        m = yo.size
        np.testing.assert_array_equal(x[off:off + m], xo)
        np.testing.assert_array_equal(y[off:off + m], yo)
        off += m
    npos = s.count()[1]
    assert npos == int((y == 1).sum()) and npos > 0
    # strided sampling selects the same rows
    s.clear()
    s.sample(g, tau=1e-3, stride=7, offset=3)
    xs, ys = s.download()
    xo, yo = S.sample_block(b.q, b.jac, 1.4, 1e-3, 0, periodic=True, stride=7, offset=3, ghosts=3)
    np.testing.assert_array_equal(xs[:yo.size], xo)
    np.testing.assert_array_equal(ys[:yo.size], yo)
    s.close()
    g.close()


def test_sample_physical_faces():
    print("case: sensor samples, walls in x and y, NS", flush=True)
    spec = cases.taylor_green(dims=(20, 14, 12), noise=3e-2, seed=22)[0]
    spec.faces = [(-1, 2, None), (-1, 2, None), (-1, 3, None), (-1, 3, None), (0, 5, None), (0, 4, None)]
    g = solver.Case(solver.Params(t_wall=1.2), [spec])
    s = solver.Sensor(100000)
    s.sample(g, tau=1e-3)
    x, y = s.download()
    off = 0
    for d in range(3):
        xo, yo = S.sample_block(spec.q, spec.jac, 1.4, 1e-3, d, periodic=(d == 2), ghosts=3)
        np.testing.assert_array_equal(x[off:off + yo.size], xo)
        np.testing.assert_array_equal(y[off:off + yo.size], yo)
        off += yo.size
    s.close()
    g.close()


def _dataset(seed, n=5000):
    blocks = cases.taylor_green(16, noise=5e-2, seed=seed)
    g = solver.Case(solver.Params(), blocks)
    s = solver.Sensor(200000)
    s.sample(g, tau=1e-3)
    g.close()
    x, y = s.download()
    keep = np.nonzero(y >= 0)[0][:n]
    s.clear()
    s.upload(x[keep], y[keep])
    return s, x[keep], y[keep]


def test_forward_matches_oracle():
    print("case: sensor logits, 5000 TGV stencils, random parameters", flush=True)
    s, x, y = _dataset(23)
    params = mixed_params(1)
    s.set_params(params)
    z = s.infer()
    zo, _ = S.forward(x, params)
    err = np.abs(z - zo).max() / max(1.0, np.abs(zo).max())
    print("  max |dz| / max|z| = %.3e" % err, flush=True)
    # the same bf16 operands with FP32 accumulation in another order: measured
    # 2e-7 of max|z|
    assert err <= 2e-6
    s.close()


def test_gradient_matches_oracle():
    print("case: sensor gradient over 3000 stencils (23 workgroups + a partial one)", flush=True)
    s, x, y = _dataset(24)
    params = mixed_params(2)
    s.set_params(params)
    first, n = 100, 2950
    for pw in (1.0, 4.0):
        loss, g = s.grad(first, n, pos_weight=pw)
        lo, go = S.loss_grad(x[first:first + n], y[first:first + n], params, pos_weight=pw)
        rel = np.linalg.norm(g - go) / np.linalg.norm(go)
        print("  pos_weight %.1f loss %.6f / %.6f  |dg|/|g| = %.3e" % (pw, loss, lo, rel), flush=True)
        assert abs(loss - lo) <= 1e-4 * abs(lo)
        assert rel <= 1e-2
        # every parameter group agrees, not only the large ones
        for a, b in ((S.OW1, S.OB1), (S.OB1, S.OW2), (S.OW2, S.OB2), (S.OB2, S.OW3), (S.OW3, S.OB3)):
            assert np.linalg.norm(g[a:b] - go[a:b]) <= 2e-2 * np.linalg.norm(go[a:b]) + 1e-9
        assert abs(g[S.OB3] - go[S.OB3]) <= 1e-3 * abs(go[S.OB3]) + 1e-7
    s.close()


def test_training_step_matches_oracle_adam():
    """One GPU Adam step equals Adam of the restatement on the same hashed
    minibatch (rows drawn as k_sensor_fwdbwd draws them), wherever the
    gradient is not negligibly small (Adam's first step is lr sign(g))."""
    print("case: one Adam step on a hashed minibatch of 1000", flush=True)
    s, x, y = _dataset(25, n=4096)
    params = mixed_params(3)
    s.set_params(params)
    lr, pw, seed, batch = 1e-3, 2.0, 5, 1000
    losses = s.train(1, batch=batch, lr=lr, pos_weight=pw, seed=seed)
    got = s.params()
    idx = S.batch_indices(seed, 0, batch, x.shape[0])
    lo, go = S.loss_grad(x[idx], y[idx], params, pos_weight=pw, batch=batch)
    assert abs(losses[0] - lo) <= 1e-4 * abs(lo)
    want, _, _ = S.adam(params, go, np.zeros(S.NP), np.zeros(S.NP), 1, lr)
    big = np.abs(go) > 1e-2 * np.abs(go).max()
    np.testing.assert_allclose(got[big], want[big], rtol=1e-6, atol=1e-7)
    step = np.abs(got.astype(np.float64) - params)
    assert step.max() <= lr * (1 + 1e-5)
    assert (step[S.OW1:S.OB1][(np.arange(1024) % 32) >= 30] == 0).all()
    s.close()


def test_training_learns():
    """A few hundred Adam steps on TGV and blast stencils: the loss falls and
    the trained network classifies held-out stencils better than the same
    network before training."""
    print("case: training 400 steps", flush=True)
    s = solver.Sensor(2000000)
    g = solver.Case(solver.Params(model=solver.EULER3D, mach=1.0), cases.spherical_blast(32))
    g.steps(20, 0.5)
    s.sample(g, tau=1e-3)
    g.close()
    g = solver.Case(solver.Params(), cases.taylor_green(32, noise=1e-2, seed=30))
    s.sample(g, tau=1e-3)
    g.close()
    n, npos = s.count()
    x, y = s.download()
    valid = y >= 0
    print("  %d samples, %d positive" % (n, npos), flush=True)
    p0 = S.init_params(11)
    s.set_params(p0)
    acc = lambda z: float(((z >= 0) == (y[valid] > 0.5)).mean())
    z0 = s.infer()[valid]
    losses = s.train(400, batch=4096, lr=3e-3, pos_weight=1.0, seed=12)
    z1 = s.infer()[valid]
    print("  loss %.4f -> %.4f; accuracy %.4f -> %.4f" % (losses[:10].mean(), losses[-10:].mean(), acc(z0), acc(z1)),
          flush=True)
    assert losses[-10:].mean() < 0.5 * losses[:10].mean()
    assert acc(z1) > max(acc(z0), 0.9)
    s.close()


def _decisions_check(g, blocks, params, thr, periodic_dirs):
    b = blocks[0]
    for d in range(3):
        flags = g.hybrid_flags(0, d)
        fo, zo = S.decisions(b.q, b.jac, 1.4, d, d in periodic_dirs, params, thr)
        far = np.abs(zo - thr) > 0.05
        assert flags.shape == fo.shape
        np.testing.assert_array_equal(flags[far], fo[far])
        print("  dir %d: central %d weno %d fallback %d closure %d; near-threshold faces %d" % (
            d, (flags == 0).sum(), (flags == 1).sum(), (flags == 2).sum(), (flags == 3).sum(), (~far).sum()),
            flush=True)


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("model", [solver.EULER3D, solver.NAVIERSTOKES3D])
def test_hybrid_rhs_parity_periodic(model, exact):
    print("case: hybrid RHS TGV 18x16x14 model=%d exact=%d" % (model, exact), flush=True)
    blocks = cases.taylor_green(dims=(18, 16, 14), noise=3e-2, seed=26)
    sp, op = hybrid_params(model, exact)
    params = mixed_params(4)
    g = solver.Case(sp, blocks)
    g.set_sensor(params, 0.0)
    g.rhs()
    flags = [g.hybrid_flags(0, d) for d in range(3)]
    assert all((f == 0).any() and (f == 1).any() for f in flags)
    _decisions_check(g, blocks, params, 0.0, (0, 1, 2))
    o = O.OracleCase(op, blocks)
    for d in range(3):
        o.set_hybrid_mask(0, d, flags[d])
    o.rhs()
    err = check(g.get_rhs(0), o.get_rhs(0))
    print("  rhs parity ratio %.3f" % err, flush=True)
    assert err <= 1.0
    cen, net, fb = g.hybrid_counts()
    assert cen == sum(int((f == 0).sum()) for f in flags) and net == sum(int((f == 1).sum()) for f in flags)
    g.close()


@pytest.mark.parametrize("interval", [1, 2])
@pytest.mark.parametrize("exact", [1, 0])
def test_hybrid_steps_walls_two_blocks(exact, interval):
    """Two blocks joined in x, walls in x and y, periodic z; NS, 2 steps: the
    oracle replays every stage with the GPU's decisions of that stage."""
    print("case: hybrid NS 2 blocks walls, stage by stage, exact=%d, decision interval %d" % (exact, interval),
          flush=True)
    spec = cases.taylor_green(dims=(24, 14, 12), noise=3e-2, seed=27)[0]
    spec.faces = [(-1, 2, None), (-1, 2, None), (-1, 3, None), (-1, 3, None), (0, 5, None), (0, 4, None)]
    blocks = cases.split_block_x(spec, 2)
    sp, op = hybrid_params(solver.NAVIERSTOKES3D, exact, t_wall=1.2)
    params = mixed_params(5)
    g = solver.Case(sp, blocks)
    g.set_sensor(params, 0.0)
    g.set_sensor_interval(interval)
    o = O.OracleCase(op, blocks)
    for step in range(2):
        for st in range(3):
            counts = g.hybrid_counts()
            g.stage(st, 0.5)
            flags = {(b, d): g.hybrid_flags(b, d) for b in range(2) for d in range(3)}
            if st == 0 and step % interval == 0:
                step_flags = flags
                assert sum(g.hybrid_counts()) > sum(counts)
            else:
                # the other stages reuse the network's decisions (and count
                # none); only the positivity fallback may turn a central face
                # into a WENO5-JS one
                for k in flags:
                    same = flags[k] == step_flags[k]
                    assert (same | ((step_flags[k] == 0) & (flags[k] == 2))).all()
                assert g.hybrid_counts() == counts
            for (b, d), f in flags.items():
                o.set_hybrid_mask(b, d, f)
            o.stage(st, 0.5)
            for b in range(2):
                err = check(g.get_q(b), o.get_q(b))
                assert err <= 1.0, (step, st, b, err)
    f0 = g.hybrid_flags(0, 0)
    assert (f0[..., :4] == 3).all() and (f0[..., 4:] != 3).all()   # left wall closure of block 0
    print("  2 steps within the parity bar", flush=True)
    g.close()


def test_hybrid_thresholds_reproduce_pure_schemes():
    print("case: hybrid with all-WENO and all-central thresholds, NS TGV 16^3", flush=True)
    blocks = cases.taylor_green(16, noise=1e-3, seed=28)
    params = mixed_params(6)
    sp, _ = hybrid_params()
    h = solver.Case(sp, blocks)
    h.set_sensor(params, -1e30)
    h.rhs()
    w = solver.Case(solver.Params(), blocks)
    w.rhs()
    np.testing.assert_array_equal(h.get_rhs(0), w.get_rhs(0))
    h.set_sensor(params, 1e30)
    h.rhs()
    c = solver.Case(solver.Params(scheme=(solver.CENTRAL8,) * 3, width=(4, 4, 4)), blocks)
    c.rhs()
    np.testing.assert_array_equal(h.get_rhs(0), c.get_rhs(0))
    for x in (h, w, c):
        x.close()


def test_hybrid_conserves_periodic():
    print("case: hybrid Euler 32^3 periodic, 5 steps, conservation", flush=True)
    blocks = cases.taylor_green(32, noise=1e-2, seed=29)
    sp, _ = hybrid_params(solver.EULER3D, exact=0)
    g = solver.Case(sp, blocks)
    g.set_sensor(mixed_params(7), 0.0)
    q0 = g.get_q(0)
    g.steps(5, 0.5)
    q1 = g.get_q(0)
    for c in range(5):
        drift = abs(q1[c].sum() - q0[c].sum()) / max(np.abs(q0[c]).sum(), np.abs(q1[c]).sum())
        print("  var %d relative drift %.2e" % (c, drift), flush=True)
        assert drift <= 1e-13
    cen, net, fb = g.hybrid_counts()
    assert cen > 0 and net > 0
    g.close()


def test_hybrid_needs_sensor():
    sp, _ = hybrid_params()
    g = solver.Case(sp, cases.taylor_green(16))
    with pytest.raises(solver.Nuc3dError, match="sensor"):
        g.rhs()
    g.close()


@pytest.mark.parametrize("interval", [1, 2])
def test_hybrid_rhs_between_stages_keeps_step_decisions(interval):
    """A single RHS evaluation in the middle of a step decides for itself and
    leaves the step's decisions alone: stage 0, nuc3d_rhs, stages 1 and 2 give
    the same state as the step without the RHS call, bit for bit.  A new state
    makes the next stage 0 decide even where the interval would reuse."""
    print("case: hybrid NS TGV 18x16x14, rhs between stages, interval %d" % interval, flush=True)
    blocks = cases.taylor_green(dims=(18, 16, 14), noise=3e-2, seed=29)
    sp, _ = hybrid_params(solver.NAVIERSTOKES3D, 1)
    params = mixed_params(7)
    runs = []
    for with_rhs in (False, True):
        g = solver.Case(sp, blocks)
        g.set_sensor(params, 0.0)
        g.set_sensor_interval(interval)
        g.steps(1, 0.5)
        g.stage(0, 0.5)
        if with_rhs:
            g.rhs()
            assert all(g.hybrid_flags(0, d).size for d in range(3))
        g.stage(1, 0.5)
        g.stage(2, 0.5)
        runs.append(g.get_q(0))
        g.close()
    np.testing.assert_array_equal(runs[0], runs[1])
    # interval 2: step 1 reuses step 0's decisions unless the state was replaced
    g = solver.Case(sp, blocks)
    g.set_sensor(params, 0.0)
    g.set_sensor_interval(interval)
    g.steps(1, 0.5)
    counts = g.hybrid_counts()
    g.set_q(0, blocks[0].q)
    g.stage(0, 0.5)
    assert sum(g.hybrid_counts()) > sum(counts)
    g.close()


@pytest.mark.parametrize("exact", [1, 0])
def test_hybrid_wide_tiles_bit_identical(exact):
    """The hybrid's x and y sweeps on 32-face x 8-line tiles give the same
    states, decisions and counts as on 16 x 16 tiles: two blocks with walls
    (closure faces at tile edges), deciding and reusing stages, a single RHS
    between stages, eager and graph-launched steps."""
    print("case: hybrid NS 2 blocks walls, sweep tiles (16,16,16) vs (32,32,16), exact=%d" % exact, flush=True)
    spec = cases.taylor_green(dims=(40, 22, 12), noise=3e-2, seed=29)[0]
    spec.faces = [(-1, 2, None), (-1, 2, None), (-1, 3, None), (-1, 3, None), (0, 5, None), (0, 4, None)]
    blocks = cases.split_block_x(spec, 2)
    sp, _ = hybrid_params(solver.NAVIERSTOKES3D, exact, t_wall=1.2)
    params = mixed_params(6)
    out = []
    for nf in ((16, 16, 16), (32, 32, 16)):
        g = solver.Case(sp, blocks)
        for b in range(2):
            g.set_sweep_tiles(b, nf)
        g.set_sensor(params, 0.0)
        g.set_sensor_interval(2)
        g.steps(1, 0.5)
        g.stage(0, 0.5)
        g.rhs()
        r = [g.get_rhs(b) for b in range(2)]
        g.stage(1, 0.5)
        g.stage(2, 0.5)
        g.steps_graph(2, 0.5)
        g.sync()
        out.append(([g.get_q(b) for b in range(2)], r, [g.hybrid_flags(b, d) for b in range(2) for d in range(3)],
                    g.hybrid_counts(), g.fixups()))
        g.close()
    (qa, ra, fa, ca, xa), (qb, rb, fb, cb, xb) = out
    for a, b in zip(qa + ra, qb + rb):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(fa, fb):
        np.testing.assert_array_equal(a, b)
    assert ca == cb and xa == xb
    print("  counts %s" % (ca,), flush=True)


def test_hybrid_autotune_tiles():
    """The autotuner times both tile shapes of the hybrid's x and y sweeps
    (the z sweep keeps 16 x 16) and the steps after it equal an untuned
    case's."""
    print("case: hybrid NS TGV 32^3, autotuned sweep tiles", flush=True)
    blocks = cases.taylor_green(32, noise=1e-2, seed=31)
    sp, _ = hybrid_params(solver.NAVIERSTOKES3D, 1)
    params = mixed_params(7)
    qs = []
    for tune in (True, False):
        g = solver.Case(sp, blocks)
        g.set_sensor(params, 0.0)
        if tune:
            g.autotune()
            nf = g.sweep_tiles(0)
            print("  sweep tiles %s" % (nf,), flush=True)
            assert nf[0] in (16, 32) and nf[1] in (16, 32) and nf[2] == 16
        g.steps_graph(2, 0.5)
        qs.append(g.get_q(0))
        g.close()
    np.testing.assert_array_equal(qs[0], qs[1])


def test_hybrid_z_tiles_stay_16():
    """32-face tiles exist for the hybrid's x and y sweeps only: asking for
    them in z is refused with NUC3D_ERR_UNSUPPORTED and changes nothing."""
    blocks = cases.taylor_green(16, noise=1e-3, seed=2)
    sp, _ = hybrid_params(solver.NAVIERSTOKES3D, 1)
    g = solver.Case(sp, blocks)
    with pytest.raises(solver.Nuc3dError, match="32-face tiles"):
        g.set_sweep_tiles(0, (32, 32, 32))
    assert tuple(g.sweep_tiles(0)) == (16, 16, 16)
    g.set_sweep_tiles(0, (32, 32, 16))
    assert tuple(g.sweep_tiles(0)) == (32, 32, 16)
    g.close()

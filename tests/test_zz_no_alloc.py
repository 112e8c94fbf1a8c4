"""The entry point that allocates nothing (nuc3d_diagnostics_noalloc, fixed
device scratch of the library) against nuc3d_diagnostics: the same enstrophy
and kinetic-energy sums bit for bit, on two calls in a row, the second on a
smaller case (the scratch keeps the first call's contents)."""
import numpy as np
import pytest

from nuc3d_amd import cases, solver

pytestmark = pytest.mark.gpu


def test_noalloc_diagnostics_matches_allocating_entry_point():
    specs = [("NS TGV 48x40x36, 2 blocks", cases.split_block_x(cases.taylor_green(dims=(48, 40, 36), noise=1e-3,
                                                                                   seed=51)[0], 2)),
             ("NS TGV 20x16x12", cases.taylor_green(dims=(20, 16, 12), noise=1e-3, seed=52))]
    for name, blocks in specs:
        print("case: %s" % name, flush=True)
        g = solver.Case(solver.Params(), blocks)
        g.steps(1, 0.5)
        ens, kin, _ = g.diagnostics(0)
        e2, k2 = g.diagnostics_noalloc()
        print("  enstrophy %.17g / %.17g, kinetic %.17g / %.17g" % (ens, e2, kin, k2), flush=True)
        assert np.float64(e2).view(np.int64) == np.float64(ens).view(np.int64)
        assert np.float64(k2).view(np.int64) == np.float64(kin).view(np.int64)
        assert ens > 0 and kin > 0
        g.close()

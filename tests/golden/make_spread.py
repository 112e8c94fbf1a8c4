"""Writes tests/golden/ref_build_spread.json: per conserved component, how far
the reference's own two builds of the same sources -- the makefile's -O0 build
(oracle/_ref/ref_harness) and -O2 -march=x86-64-v3 (oracle/_ref/ref_harness_O2,
which folds pow(x, 2.0) into x*x and contracts into FMA) -- lie apart on each
named parity case.  The parity tests use it where a component's own scale is
too small for the reference to define it to 1e-12 (tests/parity.py).
Run in this container only, after `make -C oracle ref`:
    python tests/golden/make_spread.py
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from nuc3d_amd import cases  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.golden import all_fixtures  # noqa: E402
from tests.parity import relerr_components  # noqa: E402


def flat_plate_small():
    """The 6-block flat plate of the GPU flat-plate parity test, with metrics
    and state from the oracle's restatement of block::initial."""
    out = []
    for spec in cases.flat_plate([12, 16, 20], [8, 12], 16):
        centre, dxyz, jac, met = O.mesh_metrics(*spec.nodes)
        yn = spec.nodes[1]
        wall = 0.25 * (yn[:-1, :1, :-1] + yn[1:, :1, :-1] + yn[:-1, :1, 1:] + yn[1:, :1, 1:])
        q = cases.flat_plate_state(centre, jac, y_wall=wall)
        out.append(O.Block(spec.dims, spec.faces, jac, met, dxyz, q))
    return out


def spread(params, blocks, mode, nsteps, cfl):
    runs = []
    for opt in (False, True):
        d = tempfile.mkdtemp(prefix="nuc3d_spread_")
        O.write_ref_case(d, params, blocks)
        runs.append(O.run_ref(d, len(blocks), mode, nsteps, cfl, optimised=opt))
    key = "rhs" if mode == "rhs" else "q"
    per = np.zeros(5)
    for b in range(len(blocks)):
        per = np.maximum(per, relerr_components(runs[1][b][key], runs[0][b][key]))
    return [float(x) for x in per]


def main():
    out = {}
    for fx in all_fixtures():
        pk = fx.params()
        for k in ("scheme", "diff", "width"):
            pk[k] = tuple(pk[k])
        m = fx.meta["mode"]
        out[fx.name] = spread(O.Params(**pk), fx.blocks(), m, fx.meta.get("nsteps", 0), fx.meta.get("cfl", 0.5))
    out["flatplate_small_steps2"] = spread(O.Params(mach=2.9, reynolds=2.0e4, t_wall=2.5), flat_plate_small(),
                                           "steps", 2, 0.5)
    for k, v in out.items():
        print("%-34s %s" % (k, " ".join("%.2e" % x for x in v)))
    with open(os.path.join(HERE, "ref_build_spread.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

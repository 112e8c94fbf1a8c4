"""The parity measure shared by the GPU tests.

Bar: <= 1e-12 relative max-norm per conserved variable, each of the five
conserved components measured on its own scale:

    err_c = max|got_c - ref_c| / max|ref_c|,   c in (rho, rho u, rho v, rho w, E)

and err_c <= 1e-12 for every c.

Named exceptions.  Where a component is tiny next to the fluxes that make it
(v and w momentum of the quasi-1D Shu-Osher flow are ~1e-4 of rho u; the
spanwise momentum of the flat plate is ~1e-8 of it), the reference itself
defines it to less than 1e-12: its own -O0 (makefile) and -O2 builds of the
same sources differ by up to 7e-10 there.  For such a case the bound of each
component is max(1e-12, K * s_c), with s_c that component's spread between
the reference's two builds on the same case:
  * measured by tests/golden/make_spread.py (ref_build_spread.json) for every
    case the reference can run;
  * for CRWENO5 lines beyond the reference's 261-cell limit, where it is
    undefined, the spread of the restatement between libm squares (the -O0
    build) and correctly rounded squares (the GPU's), computed by the test.
K is 1 for the exact variant (the reference's operation order) and 8 for the
fast variant, whose WENO arithmetic is contracted into FMA as the -O2 build's
is, in other places.

When NUC3D_PARITY_LOG names a file, every measurement appends one JSON line
(the running test, the case, the five err_c and their bounds) to it, so a run
of the GPU suite leaves the per-component record of every parity case."""
import json
import os

import numpy as np

TOL = 1e-12
COMPONENTS = ("rho", "rho_u", "rho_v", "rho_w", "E")
K_EXACT, K_FAST = 1.0, 8.0

_HERE = os.path.dirname(os.path.abspath(__file__))
_SPREAD = None


def relerr_components(got, ref):
    """err_c of each conserved component, in the order of COMPONENTS."""
    return [float(np.abs(got[c] - ref[c]).max()) / max(float(np.abs(ref[c]).max()), 1e-300)
            for c in range(got.shape[0])]


def build_spread(case):
    """Per-component spread between the reference's -O0 and -O2 builds on a
    named case (None where it was not measured)."""
    global _SPREAD
    if _SPREAD is None:
        with open(os.path.join(_HERE, "golden", "ref_build_spread.json")) as f:
            _SPREAD = json.load(f)
    return _SPREAD.get(case)


def bounds(case=None, exact=1, spread=None):
    """Per-component bounds: TOL, or max(TOL, K s_c) where the case has a
    spread (given, or measured for the named case)."""
    s = spread if spread is not None else (build_spread(case) if case else None)
    if s is None:
        return [TOL] * 5
    k = K_EXACT if exact else K_FAST
    return [max(TOL, k * float(x)) for x in s]


def _record(errs, bnds, case):
    path = os.environ.get("NUC3D_PARITY_LOG")
    if not path:
        return
    test = os.environ.get("PYTEST_CURRENT_TEST", "").split(" ")[0]
    with open(path, "a") as f:
        f.write(json.dumps({"test": test, "case": case, "err": errs, "bound": bnds}) + "\n")


def check(got, ref, case=None, exact=1, spread=None):
    """The parity measure: max over components of err_c / bound_c (the bar is
    <= 1; without a named case every bound is TOL)."""
    errs = relerr_components(got, ref)
    bnds = bounds(case, exact, spread)
    _record(errs, bnds, case)
    return max(e / b for e, b in zip(errs, bnds))


def relerr(got, ref):
    """The largest per-component relative error (no exceptions)."""
    return max(relerr_components(got, ref))

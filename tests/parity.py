"""The parity measure shared by the GPU tests.

Bar: <= 1e-12 relative max-norm per conserved variable, the conserved
variables being density, the momentum vector and total energy:

    err = max over V in {rho, (rho u, rho v, rho w), E} of
          max_{c in V} max|got_c - ref_c|  /  max_{c in V} max|ref_c|

Momentum is measured as one vector.  Per Cartesian component, a component
that is a tiny fraction of the momentum (v and w in the quasi-1D Shu-Osher
flow are ~1e-4 of rho u) turns one-ulp differences of the x-momentum
arithmetic into ~1e-12 relative errors: the reference built from its own
sources at -O0 and at -O2 differs by up to 1.5e-12 there (and by <= 3e-13
on every well-conditioned case), and glibc's pow(x, 2.0) -- which the
reference calls for every square -- is not the correctly rounded x*x for
about 0.1% of arguments."""
import numpy as np

TOL = 1e-12
VARIABLES = ((0,), (1, 2, 3), (4,))


def relerr(got, ref):
    errs = []
    for comps in VARIABLES:
        scale = max(float(np.abs(ref[c]).max()) for c in comps)
        diff = max(float(np.abs(got[c] - ref[c]).max()) for c in comps)
        errs.append(diff / max(scale, 1e-300))
    return max(errs)


def relerr_components(got, ref):
    """Per Cartesian component, for the record (printed, not asserted)."""
    return [float(np.abs(got[c] - ref[c]).max()) / max(float(np.abs(ref[c]).max()), 1e-300)
            for c in range(got.shape[0])]

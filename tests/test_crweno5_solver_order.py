"""How much the order of the CRWENO5 tridiagonal solve matters.

The reference solves each line's system with the Thomas recurrence
(crweno5.cpp:207-218), and so do the GPU kernels, in the same order.  A
parallel solve (cyclic reduction) rounds differently.  This measures that
difference on the CPU: the system rows are restated here (crweno5.cpp:130-160,
checked bit for bit against the oracle's Thomas solution with correctly
rounded squares), then solved by Thomas and by parallel cyclic reduction.
On smooth lines and on a Mach-3 shock with its entropy wave the face values
agree to about 1e-15 of their magnitude, and their differences (the flux
difference of the right-hand side) to about 1e-13 of theirs, within the 1e-12
parity bar but not bit for bit."""
import numpy as np
import pytest

from oracle import oracle as O

EPS = 1e-6
CRW = (0.2, 0.5, 0.3)  # crweno5.cpp linear weights


def _smooth(f0, f1, f2, f3, f4):
    t, u = f0 - 2.0 * f1 + f2, f0 - 4.0 * f1 + 3.0 * f2
    b0 = 13.0 / 12.0 * (t * t) + 0.25 * (u * u)
    t, u = f1 - 2.0 * f2 + f3, f1 - f3
    b1 = 13.0 / 12.0 * (t * t) + 0.25 * (u * u)
    t, u = f2 - 2.0 * f3 + f4, 3.0 * f2 - 4.0 * f3 + f4
    b2 = 13.0 / 12.0 * (t * t) + 0.25 * (u * u)
    return b0, b1, b2


def system(ext, w, ends):
    """Rows of the + system for faces w..n-w (crweno5.cpp:130-160); rows 0 and
    M-1 are the explicit WENO5-JS faces `ends`."""
    n = ext.size - 2 * w
    p = ext[w:]
    M = n - 2 * w + 1
    a, b, c, d = np.zeros(M), np.ones(M), np.zeros(M), np.zeros(M)
    d[0], d[-1] = ends
    for r in range(1, M - 1):
        i = w - 1 + r
        b0, b1, b2 = _smooth(*[ext[w + i - 2 + k] for k in range(5)])
        a0 = CRW[0] / ((EPS + b0) * (EPS + b0))
        a1 = CRW[1] / ((EPS + b1) * (EPS + b1))
        a2 = CRW[2] / ((EPS + b2) * (EPS + b2))
        s = a0 + a1 + a2
        o0, o1, o2 = a0 / s, a1 / s, a2 / s
        a[r] = 2.0 / 3.0 * o0 + 1.0 / 3.0 * o1
        b[r] = 1.0 / 3.0 * o0 + 2.0 / 3.0 * (o1 + o2)
        c[r] = 1.0 / 3.0 * o2
        d[r] = 1.0 / 6.0 * o0 * p[i - 1] + (5.0 * (o0 + o1) + o2) * p[i] / 6.0 + (o1 + 5.0 * o2) * p[i + 1] / 6.0
    return a, b, c, d


def thomas(a, b, c, d):
    """crweno5.cpp:207-218 order."""
    c, d = c.copy(), d.copy()
    M = b.size
    for r in range(1, M):
        c[r] = c[r] / (b[r] - c[r - 1] * a[r])
        d[r] = (d[r] - d[r - 1] * a[r]) / (b[r] - c[r - 1] * a[r])
    x = np.zeros(M)
    x[-1] = d[-1]
    for r in range(M - 2, -1, -1):
        x[r] = d[r] - c[r] * x[r + 1]
    return x


def pcr(a, b, c, d):
    """Parallel cyclic reduction (Hockney): log2(M) levels, every row at once."""
    M = b.size
    s = 1
    while s < M:
        lo = np.arange(M) - s
        hi = np.arange(M) + s
        has_lo, has_hi = lo >= 0, hi < M
        lo, hi = np.clip(lo, 0, M - 1), np.clip(hi, 0, M - 1)
        al = np.where(has_lo, -a / b[lo], 0.0)
        ga = np.where(has_hi, -c / b[hi], 0.0)
        b, d, a, c = (b + al * c[lo] + ga * a[hi], d + al * d[lo] + ga * d[hi], al * a[lo], ga * c[hi])
        s *= 2
    return d / b


def _lines():
    n, w = 256, 3
    x = (np.arange(-w, n + w) + 0.5) / n
    shock = np.where(x < 0.3, 3.857143, 1.0 + 0.2 * np.sin(5.0 * (10.0 * x - 5.0)))
    return {
        "smooth, one wave": np.sin(2 * np.pi * x) + 2.0,
        "smooth, four waves": np.sin(8 * np.pi * x) * np.cos(2 * np.pi * x) + 1.5,
        "Mach-3 shock and entropy wave": shock,
        "momentum across the shock": shock * np.where(x < 0.3, 2.629369, 0.0) + 1e-3,
    }


@pytest.mark.parametrize("name", sorted(_lines()))
def test_cyclic_reduction_against_thomas(name):
    ext, w = _lines()[name], 3
    n = ext.size - 2 * w
    with O.square_mode(O.SQ_EXACT):
        ref = O.recon_line(O.CRWENO5, 1, ext, w)
    a, b, c, d = system(ext, w, (ref[w], ref[n - w]))
    xt = thomas(a, b, c, d)
    np.testing.assert_array_equal(xt, ref[w:n - w + 1])  # the restatement is the oracle's system
    xp = pcr(a, b, c, d)
    face = np.abs(xp - xt).max() / np.abs(xt).max()
    dft, dfp = np.diff(xt), np.diff(xp)
    diff = np.abs(dfp - dft).max() / np.abs(dft).max()
    print("%s: cyclic reduction against Thomas: faces %.1e, differences %.1e, equal faces %d of %d" % (
        name, face, diff, int(np.count_nonzero(xp == xt)), xt.size))
    assert face <= 1e-14 and diff <= 1e-12

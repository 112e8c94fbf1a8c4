"""Pins of the oracle's building blocks by what the mathematics fixes:
closed forms, exactness on polynomials, convergence orders, symmetries and
conservation.  Each check is chosen so that a plausible slip (a dropped term,
a wrong sign or index, swapped operands) makes it fail."""
import numpy as np
import pytest

from nuc3d_amd import cases
from oracle import oracle as O


def ext_line(f, w):
    """cell values f(i) for i in [-w, n+w) -> ext array"""
    return np.asarray(f, dtype=np.float64)


def line_of(fun, n, w, h=1.0):
    x = (np.arange(-w, n + w) + 0.5) * h
    return fun(x), x


@pytest.mark.parametrize("scheme", [O.WENO5JS, O.WENO5Z, O.CRWENO5])
@pytest.mark.parametrize("upwind", [1, -1])
def test_constant_and_linear_data_exact(scheme, upwind):
    n, w = 24, 3
    c, _ = line_of(lambda x: 0 * x + 3.25, n, w)
    np.testing.assert_allclose(O.recon_line(scheme, upwind, c, w), 3.25, rtol=0, atol=1e-14)
    # linear data: all smoothness indicators equal, every candidate exact -> face value exact
    f, x = line_of(lambda x: 2.0 * x - 1.0, n, w)
    faces = np.arange(n + 1) * 1.0
    np.testing.assert_allclose(O.recon_line(scheme, upwind, f, w), 2.0 * faces - 1.0, rtol=0, atol=1e-12)


def test_weno5_linear_weights_closed_form():
    # equal beta -> omega = (0.1,0.6,0.3) -> (2 f0 - 13 f1 + 47 f2 + 27 f3 - 3 f4)/60 at face 3
    n, w = 12, 3
    f = 1.0 + 0.5 * np.arange(n + 2 * w)
    out = O.recon_line(O.WENO5JS, 1, f, w)
    m = 6
    s = f[w + m - 3: w + m + 2]
    np.testing.assert_allclose(out[m], (2 * s[0] - 13 * s[1] + 47 * s[2] + 27 * s[3] - 3 * s[4]) / 60, rtol=1e-15)


@pytest.mark.parametrize("scheme,zquirk,order", [(O.WENO5JS, 1, 4.6), (O.WENO5Z, 0, 4.6), (O.CRWENO5, 1, 4.6)])
def test_convergence_order_smooth(scheme, zquirk, order):
    # cell averages of sin -> face point values, 5th order (weights near optimal)
    errs = []
    for n in (32, 64, 128):
        h = 2 * np.pi / n
        w = 3
        xc = (np.arange(-w, n + w) + 0.5) * h
        avg = (np.cos(xc - h / 2) - np.cos(xc + h / 2)) / h  # exact cell averages of sin
        out = O.recon_line(scheme, 1, avg, w, zquirk=zquirk)
        xf = np.arange(n + 1) * h
        errs.append(np.abs(out - np.sin(xf)).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] >= order, (errs, rates)


def _poly_averages(coef, n, w, h=1.0):
    """exact cell averages of sum_k coef[k] x^k on cells [-w, n+w) of width h"""
    xl = (np.arange(-w, n + w)) * h
    xr = xl + h
    avg = np.zeros(n + 2 * w)
    for k, a in enumerate(coef):
        avg += a * (xr ** (k + 1) - xl ** (k + 1)) / ((k + 1) * h)
    return avg


@pytest.mark.parametrize("upwind", [1, -1])
def test_central8_polynomial_exactness(upwind):
    # the eight-cell face value is the point value of the degree-7 polynomial
    # through eight cell averages: exact up to degree 7, not at degree 8 (a
    # wrong weight or a shifted stencil breaks the first, a lower-order
    # formula the second)
    n, w = 12, 4
    rng = np.random.default_rng(3)
    for deg in range(8):
        coef = rng.standard_normal(deg + 1)
        avg = _poly_averages(coef, n, w, 0.1)   # cell i spans [0.1 i, 0.1 (i+1)]: face m at 0.1 m
        got = O.recon_line(O.CENTRAL8, upwind, avg, w)
        want = np.polyval(coef[::-1], np.arange(n + 1) * 0.1)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * max(1.0, np.abs(want).max()))
    coef = np.zeros(9)
    coef[8] = 1.0
    avg = _poly_averages(coef, n, w, 0.1)
    got = O.recon_line(O.CENTRAL8, upwind, avg, w)
    want = (np.arange(n + 1) * 0.1) ** 8
    # degree 8 leaves the truncation error 64 h^8 = 6.4e-7 at every face
    assert np.abs(got - want).max() > 1e-10


def test_central8_closed_form_and_eighth_order():
    n, w = 14, 4
    f = np.random.default_rng(4).standard_normal(n + 2 * w)
    out = O.recon_line(O.CENTRAL8, 1, f, w)
    m = 6
    s = f[w + m - 4: w + m + 4]
    want = (-3 * s[0] + 29 * s[1] - 139 * s[2] + 533 * s[3] + 533 * s[4] - 139 * s[5] + 29 * s[6] - 3 * s[7]) / 840
    np.testing.assert_allclose(out[m], want, rtol=1e-14)
    # both split parts use the same symmetric stencil
    np.testing.assert_array_equal(out, O.recon_line(O.CENTRAL8, -1, f, w))
    errs = []
    for n in (16, 32, 64):
        h = 2 * np.pi / n
        xc = (np.arange(-w, n + w) + 0.5) * h
        avg = (np.cos(xc - h / 2) - np.cos(xc + h / 2)) / h
        errs.append(np.abs(O.recon_line(O.CENTRAL8, 1, avg, w) - np.sin(np.arange(n + 1) * h)).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] >= 7.7, (errs, rates)


@pytest.mark.parametrize("upwind", [1, -1])
def test_central8_keeps_physical_face_closures(upwind):
    n, w = 14, 4
    f = np.random.default_rng(5).standard_normal(n + 2 * w) + 2.0
    c8 = O.recon_line(O.CENTRAL8, upwind, f, w, -1, -1)
    js = O.recon_line(O.WENO5JS, upwind, f, w, -1, -1)
    np.testing.assert_array_equal(c8[:w], js[:w])
    np.testing.assert_array_equal(c8[n - w + 1:], js[n - w + 1:])
    inner = O.recon_line(O.CENTRAL8, upwind, f, w)
    np.testing.assert_array_equal(c8[w:n - w + 1], inner[w:n - w + 1])


def test_weno5z_quirk_degrades_order():
    # with the +inner beta2 of weno5z.cpp:80, beta2 no longer measures the
    # curvature of its stencil and WENO5-Z drops to about third order
    errs = []
    for n in (64, 128, 256):
        h = 2 * np.pi / n
        xc = (np.arange(-3, n + 3) + 0.5) * h
        avg = (np.cos(xc - h / 2) - np.cos(xc + h / 2)) / h
        out = O.recon_line(O.WENO5Z, 1, avg, 3, zquirk=1)
        errs.append(np.abs(out - np.sin(np.arange(n + 1) * h)).max())
    rate = np.log2(errs[-2] / errs[-1])
    assert 2.5 < rate < 3.6, (errs, rate)


def test_crweno5_more_accurate_than_weno5():
    n, w = 64, 3
    h = 2 * np.pi / n
    xc = (np.arange(-w, n + w) + 0.5) * h
    avg = (np.cos(xc - h / 2) - np.cos(xc + h / 2)) / h
    xf = np.arange(n + 1) * h
    e_js = np.abs(O.recon_line(O.WENO5JS, 1, avg, w) - np.sin(xf))[w + 1:n - w].max()
    e_cr = np.abs(O.recon_line(O.CRWENO5, 1, avg, w) - np.sin(xf))[w + 1:n - w].max()
    assert e_cr < 0.5 * e_js


@pytest.mark.parametrize("scheme", [O.WENO5JS, O.CRWENO5])
def test_minus_is_mirror_of_plus(scheme):
    rng = np.random.default_rng(0)
    n, w = 20, 3
    f = rng.standard_normal(n + 2 * w)
    plus_rev = O.recon_line(scheme, 1, f[::-1].copy(), w)[::-1]
    minus = O.recon_line(scheme, -1, f, w)
    np.testing.assert_allclose(minus, plus_rev, rtol=1e-14, atol=1e-14)


def test_weno5z_quirk_only_on_plus_inner_faces():
    rng = np.random.default_rng(1)
    n, w = 20, 3
    f = rng.standard_normal(n + 2 * w)
    q = O.recon_line(O.WENO5Z, 1, f, w, zquirk=1)
    t = O.recon_line(O.WENO5Z, 1, f, w, zquirk=0)
    np.testing.assert_array_equal(q[:w], t[:w])
    np.testing.assert_array_equal(q[n - w + 1:], t[n - w + 1:])
    assert np.all(q[w:n - w + 1] != t[w:n - w + 1])
    np.testing.assert_array_equal(O.recon_line(O.WENO5Z, -1, f, w, zquirk=1), O.recon_line(O.WENO5Z, -1, f, w, zquirk=0))


@pytest.mark.parametrize("scheme,deg,w", [(O.CD6, 6, 3), (O.CD8, 8, 4)])
def test_central_difference_polynomial_exact(scheme, deg, w):
    n = 20
    x = np.arange(-w, n + w) * 1.0
    coef = np.arange(1, deg + 2) / 7.0
    f = np.polyval(coef, x / n)
    df = np.polyval(np.polyder(coef), x / n) / n
    out = O.diff_line(scheme, f, w)
    np.testing.assert_allclose(out, df[w:n + w], rtol=0, atol=1e-12)


def test_exterior_derivative_closure_orders():
    # one-sided 1st (linear exact), central 2nd (quadratic exact), central 4th (quartic exact)
    n, w = 16, 3
    x = np.arange(-w, n + w) * 1.0
    for deg, k in ((1, 0), (2, 1), (4, 2)):
        coef = np.ones(deg + 1)
        f = np.polyval(coef, x / 10.0)
        out = O.diff_line(O.CD6, f, w, typeL=-1, typeR=-1)
        df = np.polyval(np.polyder(coef), x / 10.0) / 10.0
        np.testing.assert_allclose(out[k], df[w + k], rtol=1e-12)
        np.testing.assert_allclose(out[n - 1 - k], df[w + n - 1 - k], rtol=1e-12)


@pytest.mark.parametrize("riemann", [O.LF, O.AUSM, O.AUSMP])
def test_flux_splitting_sums_to_flux(riemann):
    p = O.Params(model=O.EULER, mach=1.0, riemann=riemann)
    rng = np.random.default_rng(2)
    for _ in range(20):
        rho, u, v, w_, pr = 0.5 + rng.random(), *rng.standard_normal(3), 0.5 + rng.random()
        jac = 0.3 + rng.random()
        m = rng.standard_normal(3)
        fl, fr = O.split_point(p, [rho, u, v, w_, pr], jac, m)
        U = m @ np.array([u, v, w_])
        E = pr / 0.4 + 0.5 * rho * (u * u + v * v + w_ * w_)
        F = np.array([U * rho, U * rho * u + m[0] * pr, U * rho * v + m[1] * pr, U * rho * w_ + m[2] * pr,
                      U * (E + pr)]) / jac
        np.testing.assert_allclose(fl + fr, F, rtol=1e-12, atol=1e-12)
        if riemann == O.LF:
            lam = abs(U) + np.sqrt(1.4 * pr / rho) * np.linalg.norm(m)
            np.testing.assert_allclose(fl - fr, lam * np.array([rho, rho * u, rho * v, rho * w_, E]) / jac,
                                       rtol=1e-12)


def test_rk3_amplification_factor():
    # u' = -R with R = a u: one TVD-RK3 step multiplies u by 1 + z + z^2/2 + z^3/6, z = -a dt
    a, dt = 1.7, 0.3
    u0 = np.array([2.0])
    un = u0.copy()
    ui = u0.copy()
    for s in range(3):
        ui = O.rk3(s, a * ui, un, ui, dt)
    z = -a * dt
    np.testing.assert_allclose(ui, u0 * (1 + z + z * z / 2 + z ** 3 / 6), rtol=1e-15)


@pytest.mark.parametrize("model", [O.EULER, O.NS])
@pytest.mark.parametrize("scheme", [O.WENO5JS, O.WENO5Z, O.CRWENO5, O.CENTRAL8])
def test_periodic_rhs_telescopes(model, scheme):
    blocks = cases.taylor_green(dims=(16, 12, 14), noise=1e-2, seed=8)
    w = 4 if scheme == O.CENTRAL8 else 3
    oc = O.OracleCase(O.Params(model=model, scheme=(scheme,) * 3, width=(w,) * 3), blocks)
    oc.rhs()
    r = oc.get_rhs(0)
    for c in range(5):
        assert abs(r[c].sum()) <= 1e-13 * np.abs(r[c]).sum() + 1e-300


@pytest.mark.parametrize("model", [O.EULER, O.NS])
def test_free_stream_preserved(model):
    n = 12
    jac, met, dxyz = cases.cartesian_metrics((n, n, n), (0.3, 0.2, 0.25))
    one = np.ones((n, n, n))
    q = cases.conserved_over_j(one, 0.3 * one, -0.2 * one, 0.1 * one, 71.0 * one, 1.4, jac)
    b = cases.BlockSpec((n, n, n), cases.periodic_faces(), jac, met, dxyz, q)
    oc = O.OracleCase(O.Params(model=model), [b])
    oc.rhs()
    r = oc.get_rhs(0)
    scale = np.abs(q).max() / 0.2
    assert np.abs(r).max() <= 1e-13 * scale


def test_multiblock_equals_single_block():
    # cutting a periodic block into slabs must not change anything (halo = periodic wrap)
    spec = cases.taylor_green(dims=(24, 10, 12), noise=1e-3, seed=9)[0]
    p = O.Params(model=O.NS)
    one = O.OracleCase(p, [spec])
    one.steps(1, 0.5)
    parts = cases.split_block_x(spec, 3)
    many = O.OracleCase(p, parts)
    many.steps(1, 0.5)
    got = np.concatenate([many.get_q(b) for b in range(3)], axis=3)
    np.testing.assert_array_equal(got, one.get_q(0))


def quartic_nodes(dims, seed):
    """Node coordinates that are quartic in each index direction inside every
    4-cell tile (one global quartic per direction plus cross terms of degree
    <= 4 per direction), so the tiled 5-point interpolation is exact."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    K, J, I = np.meshgrid(np.arange(nz + 1.0), np.arange(ny + 1.0), np.arange(nx + 1.0), indexing="ij")
    a = rng.uniform(-1, 1, 6) * np.array([1e-2, 1e-3, 1e-4, 1e-3, 1e-3, 1e-4])

    def f(i, j, k):
        return (i + a[0] * i ** 2 + a[1] * i ** 3 + a[2] * i ** 4 + a[3] * i * j + a[4] * j * k ** 2
                + a[5] * i ** 2 * j ** 2 * k)

    def grad(i, j, k):
        return (1 + 2 * a[0] * i + 3 * a[1] * i ** 2 + 4 * a[2] * i ** 3 + a[3] * j + 2 * a[5] * i * j ** 2 * k,
                a[3] * i + a[4] * k ** 2 + 2 * a[5] * i ** 2 * j * k,
                2 * a[4] * j * k + a[5] * i ** 2 * j ** 2)

    xn = 0.3 * f(I, J, K)
    yn = 0.2 * f(J, K, I) + 0.05 * I
    zn = 0.25 * f(K, I, J)
    return (xn, yn, zn), f, grad


def test_mesh_metrics_exact_on_quartic_meshes():
    # the 5-point Lagrange weights (value and derivative at the cell centres)
    # are fixed uniquely by exactness on polynomials of degree 4
    dims = (8, 12, 8)
    (xn, yn, zn), f, grad = quartic_nodes(dims, 3)
    centre, dxyz, jac, met = O.mesh_metrics(xn, yn, zn)
    nx, ny, nz = dims
    K, J, I = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    np.testing.assert_allclose(centre[0], 0.3 * f(I, J, K), rtol=1e-13)
    np.testing.assert_allclose(centre[1], 0.2 * f(J, K, I) + 0.05 * I, rtol=1e-13)
    gi, gj, gk = grad(I, J, K)
    np.testing.assert_allclose(dxyz[0], 0.3 * gi, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(dxyz[1], 0.3 * gj, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(dxyz[2], 0.3 * gk, rtol=1e-12, atol=1e-14)
    # y = 0.2 f(J,K,I) + 0.05 I: d/dxi = 0.2 df/dk(J,K,I) + 0.05
    gj2, gk2, gi2 = grad(J, K, I)
    np.testing.assert_allclose(dxyz[3], 0.2 * gi2 + 0.05, rtol=1e-12, atol=1e-14)
    # J and metrics are the inverse of the derivative matrix
    D = dxyz.reshape(3, 3, -1).transpose(2, 0, 1)
    M = met.reshape(3, 3, -1).transpose(2, 0, 1)
    eye = np.einsum("nij,njk->nik", M, D)
    np.testing.assert_allclose(eye, np.broadcast_to(np.eye(3), eye.shape), atol=1e-12)
    np.testing.assert_allclose(jac.ravel(), 1.0 / np.linalg.det(D), rtol=1e-12)


def test_mesh_metrics_rejects_partial_tiles_and_folded_cells():
    xn, yn, zn = quartic_nodes((8, 8, 8), 4)[0]
    with pytest.raises(ValueError, match="-1"):
        O.mesh_metrics(xn[:, :, :8], yn[:, :, :8], zn[:, :, :8])  # nx = 7
    with pytest.raises(ValueError, match="-2"):
        O.mesh_metrics(-xn, yn, zn)  # mirrored: negative Jacobian


def test_diagnostics_taylor_green_vorticity_and_energy():
    # TGV on a Cartesian 32^3 grid: vorticity from 6th-order differences is the
    # analytic curl to O(h^6); the kinetic sum is the cell sum of rho|u|^2/2
    n = 32
    blocks = cases.taylor_green(n)
    oc = O.OracleCase(O.Params(model=O.NS), blocks)
    ens, kin, f = oc.diagnostics()
    h = 2 * np.pi / n
    X, Y, Z = cases.cell_centres((n, n, n), (0.0, 0.0, 0.0), (h, h, h))
    w0 = -np.cos(X) * np.sin(Y) * np.sin(Z)
    w1 = -np.sin(X) * np.cos(Y) * np.sin(Z)
    w2 = 2.0 * np.sin(X) * np.sin(Y) * np.cos(Z)
    for got, want in zip(f[:3], (w0, w1, w2)):
        np.testing.assert_allclose(got, want, atol=2e-5)
    np.testing.assert_allclose(f[3], 0.5 * (f[0] ** 2 + f[1] ** 2 + f[2] ** 2), rtol=1e-14)
    assert abs(ens - f[3].sum()) <= 1e-12 * ens
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    assert abs(kin - (0.5 * (u * u + v * v)).sum()) <= 1e-12 * kin


def test_square_conventions():
    """The three square conventions of the restatement: identical where every
    square is exact (integer data: libm pow(x, 2) is then x*x), and the libm
    one differs from the correctly rounded squares on general data, as glibc's
    pow(x, 2.0) is not correctly rounded for every argument."""
    ext_int = np.array([3.0, -1.0, 4.0, 1.0, -5.0, 9.0, 2.0, -6.0, 5.0, 3.0, 5.0, -8.0, 9.0, 7.0], dtype=np.float64)
    rng = np.random.default_rng(5)
    ext_rnd = rng.standard_normal(4000)
    outs = {}
    for mode in (O.SQ_LIBM, O.SQ_O2, O.SQ_EXACT):
        with O.square_mode(mode):
            outs[mode] = (O.recon_line(O.WENO5JS, 1, ext_int, 3), O.recon_line(O.WENO5Z, -1, ext_rnd, 3))
    for mode in (O.SQ_O2, O.SQ_EXACT):
        np.testing.assert_array_equal(outs[mode][0], outs[O.SQ_LIBM][0])
    # general data: the conventions differ in the last bits of some faces
    d = np.abs(outs[O.SQ_EXACT][1] - outs[O.SQ_LIBM][1])
    assert 0 < np.count_nonzero(d) < d.size
    assert d.max() <= 1e-14 * np.abs(outs[O.SQ_LIBM][1]).max()
    assert O.lib().orc_get_square_mode() == O.SQ_LIBM


@pytest.mark.parametrize("kind", ["density", "pressure"])
def test_positivity_counts(kind):
    """The restatement counts the con2prim inputs the GPU path reports as a
    status: density <= 0, and NaN pressure after the e<0 fix-up."""
    from nuc3d_amd import cases
    b = cases.taylor_green(dims=(12, 10, 8), noise=1e-3, seed=41)[0]
    if kind == "density":
        b.q[0, 4, 5, 6] = -b.q[0, 4, 5, 6]
    else:
        b.q[4, 4, 5, 6] = np.nan
    o = O.OracleCase(O.Params(), [b])
    assert o.positivity() == (0, 0)
    o.rhs()
    assert o.positivity() == ((1, 0) if kind == "density" else (0, 1))
    clean = O.OracleCase(O.Params(), cases.taylor_green(dims=(12, 10, 8), noise=1e-3, seed=41))
    clean.steps(2, 0.5)
    assert clean.positivity() == (0, 0)

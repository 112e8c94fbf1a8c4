"""Synthetic workloads for the per-block RHS / TVD-RK3 path.

Every case is built as a list of `BlockSpec`s: structured blocks with their
face topology (the six lines of bc_topo/BC_Topo_<rank>.in,
boundaryConditions.cpp:69-93), cell-centred metrics and the conserved state
divided by the Jacobian (the reference's Q_work, block.cpp:523-567).

Arrays are float64, i-fastest: numpy shape (nz, ny, nx) per field.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

# Face ids (boundaryConditions.hpp:56-61); a face tuple is (type, id, values)
# with type -1 for a physical face and the neighbour block index otherwise.
BC_INLET, BC_OUTLET, BC_WALL, BC_SYMM = 0, 1, 2, 3


@dataclass
class BlockSpec:
    dims: Tuple[int, int, int]          # (nx, ny, nz) cells
    faces: List[tuple]                  # 6 x (type, id, values or None)
    jac: np.ndarray                     # (nz, ny, nx)
    met: np.ndarray                     # (9, nz, ny, nx): xi_x..zeta_z
    dxyz: np.ndarray                    # (9, nz, ny, nx): x_xi,x_eta,x_zeta,y_xi..z_zeta
    q: np.ndarray                       # (5, nz, ny, nx): conserved / J
    # node coordinates (xn, yn, zn), each (nz+1, ny+1, nx+1), for blocks whose
    # metrics are computed on the GPU from the mesh (jac, met, dxyz and q may
    # then be None until the state is set)
    nodes: Optional[tuple] = None


@dataclass
class Physics:
    gamma: float = 1.4
    mach: float = 0.1
    reynolds: float = 1600.0
    prandtl: float = 0.72
    t_ref: float = 110.4
    t_inf: float = 288.0
    t_wall: float = 1.0


def metrics_from_jacobian_matrix(dx: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """J and the nine metric terms from the nine derivatives x_xi..z_zeta, in
    the operation order of block::getJacobians (block.cpp:285-300).

    dx: (9, ...) ordered x_xi, x_eta, x_zeta, y_xi, y_eta, y_zeta, z_xi, z_eta, z_zeta.
    """
    x_xi, x_eta, x_zeta, y_xi, y_eta, y_zeta, z_xi, z_eta, z_zeta = dx
    jacob = 1.0 / (x_xi * (y_eta * z_zeta - y_zeta * z_eta)
                   + x_eta * (y_zeta * z_xi - y_xi * z_zeta)
                   + x_zeta * (y_xi * z_eta - y_eta * z_xi))
    met = np.stack([
        (y_eta * z_zeta - y_zeta * z_eta) * jacob,
        (x_zeta * z_eta - x_eta * z_zeta) * jacob,
        (x_eta * y_zeta - x_zeta * y_eta) * jacob,
        (y_zeta * z_xi - y_xi * z_zeta) * jacob,
        (x_xi * z_zeta - x_zeta * z_xi) * jacob,
        (x_zeta * y_xi - x_xi * y_zeta) * jacob,
        (y_xi * z_eta - y_eta * z_xi) * jacob,
        (x_eta * z_xi - x_xi * z_eta) * jacob,
        (x_xi * y_eta - x_eta * y_xi) * jacob,
    ])
    return jacob, met


def cartesian_metrics(dims, h) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Uniform Cartesian cells of size h=(hx,hy,hz): returns jac, met, dxyz."""
    nx, ny, nz = dims
    shape = (nz, ny, nx)
    dx = np.zeros((9,) + shape)
    dx[0] = h[0]
    dx[4] = h[1]
    dx[8] = h[2]
    jac, met = metrics_from_jacobian_matrix(dx)
    return jac, met, dx


def conserved_over_j(rho, u, v, w, p, gamma, jac) -> np.ndarray:
    """Q_work from primitives, as block::initialQ (block.cpp:554-563)."""
    rhou = rho * u
    rhov = rho * v
    rhow = rho * w
    rhoe = (p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w))
    return np.stack([rho / jac, rhou / jac, rhov / jac, rhow / jac, rhoe / jac])


def periodic_faces(self_index: int = 0) -> List[tuple]:
    """Single block periodic in all directions: each face is connected to the
    opposite face of the same block (self-messaging, MPICommunicator.cpp:55-81)."""
    b = self_index
    return [(b, 1, None), (b, 0, None), (b, 3, None), (b, 2, None), (b, 5, None), (b, 4, None)]


def cell_centres(dims, lo, h):
    nx, ny, nz = dims
    x = lo[0] + (np.arange(nx) + 0.5) * h[0]
    y = lo[1] + (np.arange(ny) + 0.5) * h[1]
    z = lo[2] + (np.arange(nz) + 0.5) * h[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def taylor_green(n=64, phys: Optional[Physics] = None, pressure="baseline", noise=0.0, seed=0,
                 dims=None) -> List[BlockSpec]:
    """Taylor-Green vortex on [0,2pi]^3 (block.cpp:683-690).

    pressure="baseline": p = 1/(gamma Ma^2) + (cos2x+cos2y)(cos2z+2)/16;
    pressure="reference": the reference's extra -2/16 term (block.cpp:689).
    noise > 0 adds seeded uniform perturbations of relative size `noise` to rho and p.
    """
    phys = phys or Physics()
    dims = tuple(dims) if dims is not None else (n, n, n)
    h = tuple(2.0 * np.pi / d for d in dims)
    X, Y, Z = cell_centres(dims, (0.0, 0.0, 0.0), h)
    g, ma = phys.gamma, phys.mach
    rho = np.ones_like(X)
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(X)
    if pressure == "reference":
        p = 1.0 / (g * ma * ma) + ((np.cos(2.0 * Z) + 2.0) * (np.cos(2.0 * X) + np.cos(2.0 * Y)) - 2.0) / 16.0
    else:
        p = 1.0 / (g * ma * ma) + (np.cos(2.0 * X) + np.cos(2.0 * Y)) * (np.cos(2.0 * Z) + 2.0) / 16.0
    if noise:
        rng = np.random.default_rng(seed)
        rho = rho * (1.0 + noise * rng.uniform(-1, 1, rho.shape))
        p = p * (1.0 + noise * rng.uniform(-1, 1, p.shape))
    jac, met, dxyz = cartesian_metrics(dims, h)
    q = conserved_over_j(rho, u, v, w, p, g, jac)
    return [BlockSpec(dims, periodic_faces(), jac, met, dxyz, q)]


def wavy_metrics(dims, amp=0.15) -> Tuple[np.ndarray, np.ndarray, np.ndarray, Tuple[np.ndarray, ...]]:
    """A smoothly deformed, triply periodic curvilinear mesh on [0,2pi]^3:
        x = X + a sin(Y) sin(Z),  y = Y + a sin(Z) sin(X),  z = Z + a sin(X) sin(Y)
    with (X, Y, Z) the uniform cell centres.  Every metric term, cross terms
    included, is non-zero and periodic.  Returns jac, met, dxyz (the nine
    derivatives per unit index, as block::getJacobians expects them,
    block.cpp:237-325) and the physical cell centres (x, y, z)."""
    nx, ny, nz = dims
    h = (2.0 * np.pi / nx, 2.0 * np.pi / ny, 2.0 * np.pi / nz)
    X, Y, Z = cell_centres(dims, (0.0, 0.0, 0.0), h)
    sX, cX, sY, cY, sZ, cZ = np.sin(X), np.cos(X), np.sin(Y), np.cos(Y), np.sin(Z), np.cos(Z)
    x = X + amp * sY * sZ
    y = Y + amp * sZ * sX
    z = Z + amp * sX * sY
    one = np.ones_like(X)
    dx = np.stack([
        h[0] * one, h[1] * amp * cY * sZ, h[2] * amp * sY * cZ,       # x_xi, x_eta, x_zeta
        h[0] * amp * sZ * cX, h[1] * one, h[2] * amp * cZ * sX,       # y_xi, y_eta, y_zeta
        h[0] * amp * cX * sY, h[1] * amp * sX * cY, h[2] * one,       # z_xi, z_eta, z_zeta
    ])
    jac, met = metrics_from_jacobian_matrix(dx)
    return jac, met, dx, (x, y, z)


def wavy_nodes(dims, amp=0.15, lo=(0.0, 0.0, 0.0), extent=(2.0 * np.pi,) * 3):
    """Node coordinates ((nz+1, ny+1, nx+1) each) of a smoothly deformed box:
    X + a sin(Y) sin(Z), Y + a sin(Z) sin(X), Z + a sin(X) sin(Y) on the uniform
    nodes (X, Y, Z) of [lo, lo + extent] -- the mesh-file input of
    block::initial (block.cpp:42-70)."""
    nx, ny, nz = dims
    x = lo[0] + np.arange(nx + 1) * (extent[0] / nx)
    y = lo[1] + np.arange(ny + 1) * (extent[1] / ny)
    z = lo[2] + np.arange(nz + 1) * (extent[2] / nz)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return (X + amp * np.sin(Y) * np.sin(Z), Y + amp * np.sin(Z) * np.sin(X), Z + amp * np.sin(X) * np.sin(Y))


def taylor_green_wavy(dims=(24, 20, 18), amp=0.15, phys: Optional[Physics] = None, noise=1e-3,
                      seed=0) -> List[BlockSpec]:
    """The Taylor-Green field (reference pressure form, block.cpp:683-690)
    evaluated at the physical centres of the wavy periodic mesh of
    wavy_metrics, with seeded noise on rho and p."""
    phys = phys or Physics()
    jac, met, dxyz, (x, y, z) = wavy_metrics(dims, amp)
    g, ma = phys.gamma, phys.mach
    rho = np.ones_like(x)
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = 0.1 * np.sin(x + y) * np.cos(z)
    p = 1.0 / (g * ma * ma) + ((np.cos(2.0 * z) + 2.0) * (np.cos(2.0 * x) + np.cos(2.0 * y)) - 2.0) / 16.0
    if noise:
        rng = np.random.default_rng(seed)
        rho = rho * (1.0 + noise * rng.uniform(-1, 1, rho.shape))
        p = p * (1.0 + noise * rng.uniform(-1, 1, p.shape))
    q = conserved_over_j(rho, u, v, w, p, g, jac)
    return [BlockSpec(tuple(dims), periodic_faces(), jac, met, dxyz, q)]


def shu_osher(nx=512, ny=64, nz=64, seed=0, phys: Optional[Physics] = None) -> List[BlockSpec]:
    """3-D Shu-Osher shock/entropy-wave problem on x in [-5,5], periodic in y and
    z with h_y = h_z = h_x; inflow (post-shock state) at x-, outflow at x+."""
    phys = phys or Physics(mach=1.0)
    hx = 10.0 / nx
    h = (hx, hx, hx)
    Ly, Lz = ny * hx, nz * hx
    rng = np.random.default_rng(seed)
    phi1, phi2 = rng.uniform(0.0, 2.0 * np.pi, 2)
    X, Y, Z = cell_centres((nx, ny, nz), (-5.0, 0.0, 0.0), h)
    post = (3.857143, 2.629369, 0.0, 0.0, 10.3333)
    pre_rho = 1.0 + 0.2 * np.sin(5.0 * X + 0.1 * np.sin(2 * np.pi * Y / Ly + phi1) * np.sin(2 * np.pi * Z / Lz + phi2))
    shocked = X < -4.0
    rho = np.where(shocked, post[0], pre_rho)
    u = np.where(shocked, post[1], 0.0)
    v = np.zeros_like(X)
    w = np.zeros_like(X)
    p = np.where(shocked, post[4], 1.0)
    jac, met, dxyz = cartesian_metrics((nx, ny, nz), h)
    q = conserved_over_j(rho, u, v, w, p, phys.gamma, jac)
    faces = [(-1, BC_INLET, list(post)), (-1, BC_OUTLET, None), (0, 3, None), (0, 2, None), (0, 5, None),
             (0, 4, None)]
    return [BlockSpec((nx, ny, nz), faces, jac, met, dxyz, q)]


def isotropic_turbulence(n=256, seed=0, mt=0.6, k0=4.0, re_lambda=100.0, phys: Optional[Physics] = None):
    """Decaying compressible isotropic turbulence on [0,2pi]^3: uniform rho and T,
    solenoidal velocity with E(k) ~ k^4 exp(-2 (k/k0)^2), seeded random phases,
    scaled to turbulent Mach number mt.  Returns (blocks, physics) where the
    Reynolds number is set so that the Taylor-microscale Reynolds number of the
    initial field equals re_lambda."""
    phys = phys or Physics(mach=1.0, reynolds=1.0)
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n, d=1.0 / n)
    kr = np.fft.rfftfreq(n, d=1.0 / n)
    KZ, KY, KX = np.meshgrid(k, k, kr, indexing="ij")
    K2 = KX * KX + KY * KY + KZ * KZ
    Kmag = np.sqrt(K2)
    Kmag[0, 0, 0] = 1.0
    amp = np.sqrt(Kmag ** 4 * np.exp(-2.0 * (Kmag / k0) ** 2) / (4.0 * np.pi * Kmag ** 2))
    amp[0, 0, 0] = 0.0
    uh = []
    for _ in range(3):
        ph = rng.uniform(0.0, 2.0 * np.pi, K2.shape)
        uh.append(amp * np.exp(1j * ph))
    # project onto divergence-free fields: u_hat - k (k.u_hat)/|k|^2
    kd = (KX * uh[0] + KY * uh[1] + KZ * uh[2]) / np.where(K2 > 0, K2, 1.0)
    uh = [uh[0] - KX * kd, uh[1] - KY * kd, uh[2] - KZ * kd]
    vel = [np.fft.irfftn(a, s=(n, n, n), axes=(0, 1, 2)) for a in uh]
    g = phys.gamma
    # c = sqrt(gamma p / rho) = 1/Ma for p = 1/(gamma Ma^2), rho = 1
    c = 1.0 / phys.mach
    urms = np.sqrt(sum(np.mean(a * a) for a in vel))
    scale = mt * c / urms
    vel = [a * scale for a in vel]
    u, v, w = vel
    # Taylor microscale from <u^2>/<(du/dx)^2> and Re = Re_lambda / (u' lambda)
    dudx = np.fft.irfftn(1j * KX * np.fft.rfftn(u), s=(n, n, n), axes=(0, 1, 2))
    uprime = np.sqrt(np.mean(u * u))
    lam = np.sqrt(np.mean(u * u) / np.mean(dudx * dudx))
    phys.reynolds = re_lambda / (uprime * lam)
    rho = np.ones((n, n, n))
    p = np.full((n, n, n), 1.0 / (g * phys.mach * phys.mach))
    h = (2.0 * np.pi / n,) * 3
    jac, met, dxyz = cartesian_metrics((n, n, n), h)
    q = conserved_over_j(rho, u, v, w, p, g, jac)
    return [BlockSpec((n, n, n), periodic_faces(), jac, met, dxyz, q)], phys


def spherical_blast(n=64, dims=None, phys: Optional[Physics] = None) -> List[BlockSpec]:
    """3-D spherical blast on a periodic cube of side L=1: p=10, rho=1 inside
    r<0.2L, p=1, rho=0.125 outside."""
    phys = phys or Physics(mach=1.0)
    dims = tuple(dims) if dims is not None else (n, n, n)
    h = tuple(1.0 / d for d in dims)
    X, Y, Z = cell_centres(dims, (0.0, 0.0, 0.0), h)
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
    inside = r < 0.2
    rho = np.where(inside, 1.0, 0.125)
    p = np.where(inside, 10.0, 1.0)
    z0 = np.zeros_like(X)
    jac, met, dxyz = cartesian_metrics(dims, h)
    q = conserved_over_j(rho, z0, z0, z0, p, phys.gamma, jac)
    return [BlockSpec(dims, periodic_faces(), jac, met, dxyz, q)]


def split_block_x(spec: BlockSpec, nparts: int) -> List[BlockSpec]:
    """Cut one block into nparts slabs along x, connected face to face, keeping
    the original x-faces at the ends.  The result is the multi-block form of
    the same problem (one block per rank in the reference)."""
    nx, ny, nz = spec.dims
    cuts = np.linspace(0, nx, nparts + 1).astype(int)
    out = []
    for p in range(nparts):
        a, b = cuts[p], cuts[p + 1]
        faces = list(spec.faces)
        # left face
        if p == 0:
            f0 = spec.faces[0]
            faces[0] = (nparts - 1, 1, None) if f0[0] >= 0 else f0
        else:
            faces[0] = (p - 1, 1, None)
        if p == nparts - 1:
            f1 = spec.faces[1]
            faces[1] = (0, 0, None) if f1[0] >= 0 else f1
        else:
            faces[1] = (p + 1, 0, None)
        for f in range(2, 6):
            t, i, v = spec.faces[f]
            faces[f] = (p, i, v) if t >= 0 else (t, i, v)
        out.append(BlockSpec((b - a, ny, nz), faces, spec.jac[:, :, a:b].copy(), spec.met[:, :, :, a:b].copy(),
                             spec.dxyz[:, :, :, a:b].copy(), spec.q[:, :, :, a:b].copy()))
    return out


@dataclass
class BatchMember:
    block: BlockSpec
    group: int
    mach: float
    reynolds: float
    kind: str


def batch_cases(n_cases=48, lo=32, hi=160, seed=0, step=4) -> List[BatchMember]:
    """A batch of independent cases of seeded uneven sizes: cubes of n^3 cells
    with n in [lo, hi] (multiples of `step`), half Taylor-Green vortices on
    [0,2pi]^3 at Ma 0.1 with Re drawn from [400, 3200], half spherical blasts
    on the unit cube (p=10, rho=1 inside r<0.2; p=1, rho=0.125 outside) at the
    blast's own normalisation (Ma 1) and Re 1e4.  Every case is one periodic
    block of its own group (its own dt)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo // step, hi // step + 1, n_cases) * step
    out = []
    for c in range(n_cases):
        n = int(sizes[c])
        if c % 2 == 0:
            re = float(rng.uniform(400.0, 3200.0))
            phys = Physics(mach=0.1, reynolds=re)
            b = taylor_green(n, phys=phys)[0]
            kind = "tgv"
            mach = 0.1
        else:
            re = 1.0e4
            b = spherical_blast(n)[0]
            kind = "blast"
            mach = 1.0
        b.faces = periodic_faces(c)
        out.append(BatchMember(b, c, mach, re, kind))
    return out


def flat_plate(sizes_x, sizes_z, ny, delta=1.0, lx=None, ly=None, lz=None, bump_height=0.5, bump_width=2.0,
               beta=2.5, mach=2.9):
    """Multiblock supersonic flat-plate boundary layer over a Gaussian bump:
    len(sizes_x) x len(sizes_z) blocks (x index fastest) of sizes_x[i] x ny x
    sizes_z[k] cells on a curvilinear mesh.  The wall y = h(x, z) carries the
    bump h = bump_height * delta * exp(-((x-x0)^2 + (z-z0)^2) / (bump_width *
    delta)^2) at x0 = 0.3 lx, z0 = lz / 2; wall-normal nodes are clustered at
    the wall by tanh stretching, y = h + (ly - h) (1 - tanh(beta (1 - eta)) /
    tanh(beta)).  Faces: x- inflow (free stream rho = 1, u = 1, p = 1/(gamma
    Ma^2)), x+ outflow, y- isothermal wall, y+ outflow, z periodic; blocks are
    joined face to face.  Every block carries its nodes; metrics and state are
    set by setup_flat_plate."""
    nbx, nbz = len(sizes_x), len(sizes_z)
    nxt, nzt = int(sum(sizes_x)), int(sum(sizes_z))
    lx = lx if lx is not None else 20.0 * delta
    ly = ly if ly is not None else 5.0 * delta
    lz = lz if lz is not None else 10.0 * delta
    gamma = 1.4
    inflow = [1.0, 1.0, 0.0, 0.0, 1.0 / (gamma * mach * mach)]
    x0, z0, w = 0.3 * lx, 0.5 * lz, bump_width * delta
    xs = np.concatenate([[0], np.cumsum(sizes_x)])
    zs = np.concatenate([[0], np.cumsum(sizes_z)])
    eta = np.arange(ny + 1) / ny
    stretch = 1.0 - np.tanh(beta * (1.0 - eta)) / np.tanh(beta)
    blocks = []
    for k in range(nbz):
        for i in range(nbx):
            b = k * nbx + i
            xn = (xs[i] + np.arange(sizes_x[i] + 1)) * (lx / nxt)
            zn = (zs[k] + np.arange(sizes_z[k] + 1)) * (lz / nzt)
            Z, E, X = np.meshgrid(zn, stretch, xn, indexing="ij")
            h = bump_height * delta * np.exp(-((X - x0) ** 2 + (Z - z0) ** 2) / (w * w))
            Y = h + (ly - h) * E
            left = (-1, BC_INLET, inflow) if i == 0 else (b - 1, 1, None)
            right = (-1, BC_OUTLET, None) if i == nbx - 1 else (b + 1, 0, None)
            back = (k - 1) % nbz * nbx + i
            front = (k + 1) % nbz * nbx + i
            faces = [left, right, (-1, BC_WALL, None), (-1, BC_OUTLET, None), (back, 5, None), (front, 4, None)]
            blocks.append(BlockSpec((int(sizes_x[i]), int(ny), int(sizes_z[k])), faces, None, None, None, None,
                                    nodes=(X, Y, Z)))
    return blocks


def flat_plate_blocks(seed=3, lo=96, hi=192, ny=192, nbx=4, nbz=4):
    """The 16-block flat plate of seeded uneven sizes (multiples of 4 in
    [lo, hi] per x column and z row, ny wall-normal cells): 60.5 M cells with
    the defaults (seed 3: x columns 176, 104, 112, 116; z rows 112, 176, 180,
    152)."""
    rng = np.random.default_rng(seed)
    sx = rng.integers(lo // 4, hi // 4 + 1, nbx) * 4
    sz = rng.integers(lo // 4, hi // 4 + 1, nbz) * 4
    return flat_plate(sx, sz, ny)


def flat_plate_state(centre, jac, delta=1.0, mach=2.9, t_wall=2.5, gamma=1.4, y_wall=None):
    """Initial conserved/J of a flat-plate block from its cell centres
    (3, nz, ny, nx) and J: u = tanh(3 d / delta) with d the height above the
    wall at the cell's column, T = T_w + (1 - T_w) u (the wall at T_w, the
    free stream at 1), uniform p = 1/(gamma Ma^2), rho = gamma Ma^2 p / T."""
    y = centre[1]
    d = y - (y_wall if y_wall is not None else y[:, :1, :])
    d = np.maximum(d, 0.0)
    u = np.tanh(3.0 * d / delta)
    T = t_wall + (1.0 - t_wall) * u
    p = np.full_like(u, 1.0 / (gamma * mach * mach))
    rho = gamma * mach * mach * p / T
    z0 = np.zeros_like(u)
    return conserved_over_j(rho, u, z0, z0, p, gamma, jac)


def setup_flat_plate(case, blocks, mach=2.9, t_wall=2.5):
    """Metrics of every block from its nodes on the GPU (nuc3d_upload_mesh),
    then the initial state from the cell centres the GPU returns.  The wall
    height of each column is the height of the node below its first cell."""
    for b, spec in enumerate(blocks):
        centre = case.upload_mesh(b, *spec.nodes)
        jac = case.get_metrics(b)[0]
        yn = spec.nodes[1]
        wall = 0.25 * (yn[:-1, :1, :-1] + yn[1:, :1, :-1] + yn[:-1, :1, 1:] + yn[1:, :1, 1:])
        case.set_q(b, flat_plate_state(centre, jac, mach=mach, t_wall=t_wall, y_wall=wall))

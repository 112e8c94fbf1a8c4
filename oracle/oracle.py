"""ctypes front end of the CPU oracle (nuc3d_oracle.c) and of the reference
harness (ref_harness.cpp compiled with the reference's sources).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker, never as the product.
"""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnuc3d_oracle.so")
REF_BIN = os.path.join(HERE, "_ref", "ref_harness")
REF_BIN_O2 = os.path.join(HERE, "_ref", "ref_harness_O2")
# the reference's driver code with its operator units replaced by the drop-in
# GPU definitions (nuc3d_amd/shim/): the caller of the operator-level ABI under
# test, checked against REF_BIN
REF_BIN_GPUOPS = os.path.join(HERE, "_ref", "ref_harness_gpuops")
# the reference's block::initial on a mesh file (metrics and initial state)
REF_BIN_METRICS = os.path.join(HERE, "_ref", "ref_metrics")

WENO5JS, WENO5Z, CRWENO5 = 0, 1, 2
CENTRAL8 = 3   # not a reference scheme: the smooth-region face value of the learned hybrid (width 4)
HYBRID = 4     # not a reference scheme: per-face WENO5-JS or CENTRAL8 from a decision mask (width 4)
CD6, CD8 = 0, 1
LF, AUSM, AUSMP = 0, 1, 2
EULER, NS = 0, 1
BC_INLET, BC_OUTLET, BC_WALL, BC_SYMM = 0, 1, 2, 3
# how the reference's pow(x, 2) squares are restated: libm pow (the makefile's
# -O0 build), literal-exponent squares folded to x*x (the same sources at -O2),
# or every square as the correctly rounded x*x (nuc3d_oracle.c)
SQ_LIBM, SQ_O2, SQ_EXACT = 0, 1, 2

SCHEME_NAMES = {WENO5JS: "weno5js", WENO5Z: "weno5z", CRWENO5: "crweno5"}
DIFF_NAMES = {CD6: "cd6", CD8: "cd8"}
RIEMANN_NAMES = {LF: "LF", AUSM: "AUSM", AUSMP: "AUSMp"}


class _Params(ctypes.Structure):
    _fields_ = [
        ("gamma", ctypes.c_double), ("mach", ctypes.c_double), ("reynolds", ctypes.c_double),
        ("prandtl", ctypes.c_double), ("t_ref", ctypes.c_double), ("t_inf", ctypes.c_double),
        ("t_wall", ctypes.c_double), ("model", ctypes.c_int), ("riemann", ctypes.c_int),
        ("scheme", ctypes.c_int * 3), ("diff", ctypes.c_int * 3), ("width", ctypes.c_int * 3),
        ("zquirk", ctypes.c_int),
    ]


@dataclass
class Params:
    """Physical and numerical parameters (inp/PhysModel.in + inp/Method.in)."""
    gamma: float = 1.4
    mach: float = 0.1
    reynolds: float = 1600.0
    prandtl: float = 0.72
    t_ref: float = 110.4
    t_inf: float = 288.0
    t_wall: float = 1.0
    model: int = NS
    riemann: int = LF
    scheme: Sequence[int] = (WENO5JS, WENO5JS, WENO5JS)
    diff: Sequence[int] = (CD6, CD6, CD6)
    width: Sequence[int] = (3, 3, 3)
    zquirk: int = 1

    def to_c(self) -> _Params:
        p = _Params()
        for k in ("gamma", "mach", "reynolds", "prandtl", "t_ref", "t_inf", "t_wall", "model", "riemann", "zquirk"):
            setattr(p, k, getattr(self, k))
        for i in range(3):
            p.scheme[i] = self.scheme[i]
            p.diff[i] = self.diff[i]
            p.width[i] = self.width[i]
        return p


@dataclass
class Block:
    """One structured block: dims (nx,ny,nz), face topology, metrics and state.

    Arrays are float64 and i-fastest, i.e. numpy shape (nz, ny, nx).
    faces: 6 tuples (type, id, values[5]) as in bc_topo/BC_Topo_<rank>.in.
    """
    dims: Sequence[int]
    faces: List[tuple]
    jac: np.ndarray
    met: np.ndarray          # (9, nz, ny, nx)
    dxyz: np.ndarray         # (9, nz, ny, nx)
    q: np.ndarray            # (5, nz, ny, nx), conserved / J


def _lib():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
    lib = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.orc_case_create.argtypes = [ctypes.POINTER(_Params), ctypes.c_int, ip, ctypes.POINTER(ctypes.c_void_p)]
    lib.orc_case_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_set_faces.argtypes = [ctypes.c_void_p, ctypes.c_int, ip, ip, dp]
    lib.orc_set_metrics.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp]
    lib.orc_set_q.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
    lib.orc_get_q.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
    lib.orc_get_rhs.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
    lib.orc_get_prim.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
    lib.orc_get_dt_local.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_get_dt_local.restype = ctypes.c_double
    lib.orc_get_dt_global.argtypes = [ctypes.c_void_p]
    lib.orc_get_dt_global.restype = ctypes.c_double
    lib.orc_get_max_eigen.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
    lib.orc_get_fixups.argtypes = [ctypes.c_void_p]
    lib.orc_get_fixups.restype = ctypes.c_long
    lib.orc_rhs.argtypes = [ctypes.c_void_p]
    lib.orc_stage.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
    lib.orc_steps.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, dp]
    lib.orc_recon_line.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, dp]
    lib.orc_diff_line.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, dp]
    lib.orc_split_point.argtypes = [ctypes.POINTER(_Params), dp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, dp, dp]
    lib.orc_split_point.restype = None
    lib.orc_rk3.argtypes = [ctypes.c_int, ctypes.c_long, dp, dp, dp, ctypes.c_double]
    lib.orc_rk3.restype = None
    lib.orc_mesh_metrics.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, dp, dp, dp]
    lib.orc_diagnostics.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_int, dp]
    lib.orc_get_residual.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_get_residual.restype = ctypes.c_double
    lib.orc_set_hybrid_mask.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.orc_get_positivity.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
    lib.orc_get_positivity.restype = None
    lib.orc_set_square_mode.argtypes = [ctypes.c_int]
    lib.orc_set_square_mode.restype = None
    lib.orc_get_square_mode.argtypes = []
    # NUC3D_ORACLE_SQ selects the square convention for a whole run (default libm)
    env = os.environ.get("NUC3D_ORACLE_SQ")
    if env:
        lib.orc_set_square_mode(int(env))
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _lib()
    return _LIB


class square_mode:
    """Context manager: evaluate the oracle with another square convention
    (SQ_LIBM, SQ_O2 or SQ_EXACT), restoring the previous one on exit."""

    def __init__(self, mode: int):
        self.mode = mode
        self.prev = None

    def __enter__(self):
        self.prev = lib().orc_get_square_mode()
        lib().orc_set_square_mode(self.mode)
        return self

    def __exit__(self, *exc):
        lib().orc_set_square_mode(self.prev)
        return False


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleCase:
    """A multi-block case evaluated by the C restatement."""

    def __init__(self, params: Params, blocks: Sequence[Block]):
        self.params = params
        self.blocks = list(blocks)
        self._cp = params.to_c()
        dims = (ctypes.c_int * (3 * len(blocks)))(*[int(d) for b in blocks for d in b.dims])
        h = ctypes.c_void_p()
        rc = lib().orc_case_create(ctypes.byref(self._cp), len(blocks), dims, ctypes.byref(h))
        if rc:
            raise ValueError(f"orc_case_create failed ({rc})")
        self.h = h
        for bi, b in enumerate(blocks):
            t = (ctypes.c_int * 6)(*[int(f[0]) for f in b.faces])
            i = (ctypes.c_int * 6)(*[int(f[1]) for f in b.faces])
            vals = np.zeros(30)
            for fi, f in enumerate(b.faces):
                if len(f) > 2 and f[2] is not None:
                    vals[5 * fi:5 * fi + len(f[2])] = f[2]
            self._check(lib().orc_set_faces(self.h, bi, t, i, _dp(vals)), "set_faces")
            jac = np.ascontiguousarray(b.jac, dtype=np.float64)
            met = np.ascontiguousarray(b.met, dtype=np.float64)
            dxyz = np.ascontiguousarray(b.dxyz, dtype=np.float64)
            self._check(lib().orc_set_metrics(self.h, bi, _dp(jac), _dp(met), _dp(dxyz)), "set_metrics")
            self.set_q(bi, b.q)

    @staticmethod
    def _check(rc, what):
        if rc:
            raise RuntimeError(f"oracle {what} failed ({rc})")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_case_destroy(self.h)
            self.h = None

    def _shape(self, bi, ncomp):
        nx, ny, nz = self.blocks[bi].dims
        return (ncomp, nz, ny, nx)

    def set_q(self, bi, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self._check(lib().orc_set_q(self.h, bi, _dp(q)), "set_q")

    def set_hybrid_mask(self, bi, d, mask):
        """Decisions of the hybrid scheme for direction d of block bi: uint8 in the
        layout of that direction's face array (n_{d+2}, n_{d+1}, n_d + 1); 1 or 2
        take WENO5-JS, 0 the central value.  None: all WENO5-JS."""
        if mask is None:
            self._check(lib().orc_set_hybrid_mask(self.h, bi, d, None), "set_hybrid_mask")
            return
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        n = self.blocks[bi].dims
        assert mask.size == (n[d] + 1) * n[(d + 1) % 3] * n[(d + 2) % 3]
        self._check(lib().orc_set_hybrid_mask(self.h, bi, d, mask.ctypes.data_as(ctypes.c_void_p)),
                    "set_hybrid_mask")

    def get_q(self, bi=0):
        out = np.empty(self._shape(bi, 5))
        lib().orc_get_q(self.h, bi, _dp(out))
        return out

    def get_rhs(self, bi=0):
        out = np.empty(self._shape(bi, 5))
        lib().orc_get_rhs(self.h, bi, _dp(out))
        return out

    def get_prim(self, bi=0):
        out = np.empty(self._shape(bi, 8))
        lib().orc_get_prim(self.h, bi, _dp(out))
        return out

    def dt_local(self, bi=0):
        return lib().orc_get_dt_local(self.h, bi)

    def dt_global(self):
        return lib().orc_get_dt_global(self.h)

    def max_eigen(self, bi=0):
        out = np.empty(3)
        lib().orc_get_max_eigen(self.h, bi, _dp(out))
        return out

    def fixups(self):
        return lib().orc_get_fixups(self.h)

    def positivity(self):
        """con2prim inputs so far with density <= 0 (or NaN) and with NaN pressure."""
        v = (ctypes.c_long * 2)()
        lib().orc_get_positivity(self.h, v)
        return int(v[0]), int(v[1])

    def rhs(self):
        self._check(lib().orc_rhs(self.h), "rhs")

    def stage(self, stage, cfl):
        self._check(lib().orc_stage(self.h, stage, cfl), "stage")

    def steps(self, n, cfl):
        dts = np.empty(max(n, 1))
        self._check(lib().orc_steps(self.h, n, cfl, _dp(dts)), "steps")
        return dts[:n]

    def residual(self, bi=-1):
        """Residual of the last step (PDEData3d::setRES): block bi or, for
        bi < 0, the maximum over blocks."""
        return lib().orc_get_residual(self.h, bi)

    def diagnostics(self, bi=0):
        """(enstrophy, kinetic, fields5 of block bi) of postproc::solveQ."""
        e, k = ctypes.c_double(), ctypes.c_double()
        f = np.empty(self._shape(bi, 5))
        self._check(lib().orc_diagnostics(self.h, ctypes.byref(e), ctypes.byref(k), bi, _dp(f)), "diagnostics")
        return e.value, k.value, f


def recon_line(scheme, upwind, ext, w, typeL=0, typeR=0, zquirk=1):
    ext = np.ascontiguousarray(ext, dtype=np.float64)
    n = ext.size - 2 * w
    out = np.empty(n + 1)
    rc = lib().orc_recon_line(scheme, upwind, n, w, _dp(ext), typeL, typeR, zquirk, _dp(out))
    if rc:
        raise ValueError(f"orc_recon_line failed ({rc})")
    return out


def diff_line(scheme, ext, w, typeL=0, typeR=0):
    ext = np.ascontiguousarray(ext, dtype=np.float64)
    n = ext.size - 2 * w
    out = np.empty(n)
    rc = lib().orc_diff_line(scheme, n, w, _dp(ext), typeL, typeR, _dp(out))
    if rc:
        raise ValueError(f"orc_diff_line failed ({rc})")
    return out


def split_point(params: Params, prim, jac, m):
    cp = params.to_c()
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    fl, fr = np.empty(5), np.empty(5)
    lib().orc_split_point(ctypes.byref(cp), _dp(prim), jac, m[0], m[1], m[2], _dp(fl), _dp(fr))
    return fl, fr


def rk3(stage, rhs, un, ui, dt):
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    un = np.ascontiguousarray(un, dtype=np.float64)
    ui = np.array(ui, dtype=np.float64, order="C")
    lib().orc_rk3(stage, rhs.size, _dp(rhs), _dp(un), _dp(ui), dt)
    return ui


def mesh_metrics(xn, yn, zn):
    """Cell centres, coordinate derivatives, J and metrics from node
    coordinates of shape (nz+1, ny+1, nx+1) (block.cpp:147-477).
    Returns centre (3,nz,ny,nx), dxyz (9,...), jac (...), met (9,...)."""
    xn, yn, zn = (np.ascontiguousarray(a, dtype=np.float64) for a in (xn, yn, zn))
    nz, ny, nx = (d - 1 for d in xn.shape)
    shape = (nz, ny, nx)
    centre, dxyz, jac, met = np.empty((3,) + shape), np.empty((9,) + shape), np.empty(shape), np.empty((9,) + shape)
    rc = lib().orc_mesh_metrics(nx, ny, nz, _dp(xn), _dp(yn), _dp(zn), _dp(centre), _dp(dxyz), _dp(jac), _dp(met))
    if rc:
        raise ValueError("orc_mesh_metrics failed (%d)" % rc)
    return centre, dxyz, jac, met


# --------------------------------------------------------------------------
# reference harness (oracle/_ref), only where it was built from the reference


def ref_available(optimised=False) -> bool:
    return os.path.exists(REF_BIN_O2 if optimised else REF_BIN)


def _model_name(p: Params):
    return "NavierStokes3d" if p.model == NS else "Euler3d"


def write_ref_case(path: str, params: Params, blocks: Sequence[Block]):
    """Write inp/PhysModel.in, inp/Method.in, bc_topo/ and in_<rank>.bin in the
    formats the reference reads (physicsModel.cpp:60-112, fieldOperator.cpp:33-64,
    boundaryConditions.cpp:25-93)."""
    os.makedirs(os.path.join(path, "inp"), exist_ok=True)
    os.makedirs(os.path.join(path, "bc_topo"), exist_ok=True)
    with open(os.path.join(path, "inp", "PhysModel.in"), "w") as f:
        f.write("neqs 5\nEoS IdealGas\nRiemann %s\nModel %s\nVisModel Sutherland\n"
                % (RIEMANN_NAMES[params.riemann], _model_name(params)))
        for k, v in (("Reynolds", params.reynolds), ("Mach", params.mach), ("Pt", params.prandtl),
                     ("Gamma", params.gamma), ("T_ref", params.t_ref), ("T_inf", params.t_inf),
                     ("T_wall", params.t_wall)):
            f.write("%s %.17g\n" % (k, v))
    with open(os.path.join(path, "inp", "Method.in"), "w") as f:
        f.write("scheme_time RK3\n")
        for ax, d in zip("xyz", range(3)):
            f.write("scheme_%s_ivs %s\n" % (ax, SCHEME_NAMES[params.scheme[d]]))
        for ax, d in zip("xyz", range(3)):
            f.write("scheme_%s_vis %s\n" % (ax, DIFF_NAMES[params.diff[d]]))
    for r, b in enumerate(blocks):
        lines = []
        for (t, i, *rest) in b.faces:
            vals = rest[0] if rest and rest[0] is not None else []
            lines.append(" ".join([str(int(t)), str(int(i))] + ["%.17g" % v for v in vals]))
        with open(os.path.join(path, "bc_topo", "BC_Topo_%d.in" % r), "w") as f:
            f.write("\n".join(lines) + "\n")
        with open(os.path.join(path, "in_%d.bin" % r), "wb") as f:
            np.asarray(b.dims, dtype=np.int32).tofile(f)
            for a in (b.jac[None], b.met, b.dxyz, b.q):
                np.ascontiguousarray(a, dtype=np.float64).tofile(f)


def run_ref(path: str, nranks: int, mode: str = "rhs", nsteps: int = 0, cfl: float = 0.5,
            optimised: bool = False, timeout: float = 600.0, exe: str = None):
    """Run the reference harness in `path` (or `exe`, another build of it);
    return per-rank output dicts."""
    exe = exe or (REF_BIN_O2 if optimised else REF_BIN)
    args = [exe, mode] + ([str(nsteps), repr(cfl)] if mode == "steps" else [])
    env = dict(os.environ)
    if nranks == 1:
        cmd = args
    else:
        mpiexec = shutil.which("mpiexec", path="/opt/conda/bin") or "/opt/conda/bin/mpiexec"
        cmd = [mpiexec, "-n", str(nranks)] + args
    subprocess.run(cmd, cwd=path, check=True, env=env, timeout=timeout,
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    return [read_ref_output(path, r) for r in range(nranks)]


def read_ref_output(path: str, rank: int):
    dims = np.fromfile(os.path.join(path, "in_%d.bin" % rank), dtype=np.int32, count=3)
    nx, ny, nz = (int(d) for d in dims)
    N = nx * ny * nz
    raw = np.fromfile(os.path.join(path, "out_%d.bin" % rank), dtype=np.float64)
    sh = lambda a, c: a.reshape(c, nz, ny, nx)
    o = 0
    rhs = sh(raw[o:o + 5 * N], 5); o += 5 * N
    q = sh(raw[o:o + 5 * N], 5); o += 5 * N
    w = sh(raw[o:o + 5 * N], 5); o += 5 * N
    w0 = sh(raw[o:o + 3 * N], 3); o += 3 * N
    dt_local, dt_global, e0, e1, e2, ns = raw[o:o + 6]; o += 6
    dts = raw[o:o + int(ns)]; o += int(ns)
    walls = raw[o:o + int(ns)]; o += int(ns)
    res = raw[o:o + int(ns)]
    return dict(rhs=rhs, q=q, prim=np.concatenate([w, w0]), dt_local=dt_local, dt_global=dt_global,
                max_eigen=np.array([e0, e1, e2]), dts=dts, walls=walls, res=res)


def write_mesh_case(path: str, params: Params, nodes, faces, benchmark: int = 2, start_step: int = 0,
                    tecplot: bool = False):
    """A case the reference builds from mesh files: mesh/mesh_<rank>.x (node
    counts, then x, y, z one value per line, block.cpp:42-70), the inp/ files
    of write_ref_case, inp/IOController.in with `Benchmark` (the initial
    condition, block.cpp:523-567), inp/PostProc.in with no statistics, and the
    flowData/ directory its geometry output goes to.  With start_step > 0 the
    reference reads its initial state from
    flowData/NUC3d_step_<start_step>_ID_<rank>.bin instead (block.cpp:92-95).
    `tecplot` switches on the Tecplot flow output of block::Output.
    nodes: per block (xn, yn, zn) of shape (nz+1, ny+1, nx+1)."""
    dummy = []
    for (xn, yn, zn), fc in zip(nodes, faces):
        nz, ny, nx = (d - 1 for d in xn.shape)
        z = np.zeros((nz, ny, nx))
        dummy.append(Block((nx, ny, nz), fc, z, np.zeros((9, nz, ny, nx)), np.zeros((9, nz, ny, nx)),
                           np.zeros((5, nz, ny, nx))))
    write_ref_case(path, params, dummy)
    os.makedirs(os.path.join(path, "mesh"), exist_ok=True)
    os.makedirs(os.path.join(path, "flowData"), exist_ok=True)
    with open(os.path.join(path, "inp", "IOController.in"), "w") as f:
        f.write("startStep %d\nendStep %d\nsaveStep 1000\npostStep 1000\nBenchmark %d\n"
                "cfl 0.5\ntl 0.0\ndt 0.01\nBinary no\nTecplot %s\nPostProc no\n"
                % (start_step, start_step + 1, benchmark, "yes" if tecplot else "no"))
    with open(os.path.join(path, "inp", "PostProc.in"), "w") as f:  # postproc.cpp:29-72: no statistics
        f.write("AverageType 0\nAverageStartStep 0\nSimutaneousData 0\nAveragedSaveType 0\n"
                "SimutaneousSaveType 0\n")
    for r, (xn, yn, zn) in enumerate(nodes):
        nz1, ny1, nx1 = xn.shape
        with open(os.path.join(path, "mesh", "mesh_%d.x" % r), "w") as f:
            f.write("%d %d %d\n" % (nx1, ny1, nz1))
            for a in (xn, yn, zn):
                f.write("\n".join(repr(float(v)) for v in np.asarray(a, dtype=np.float64).ravel()))
                f.write("\n")


def run_ref_metrics(path: str, nranks: int, timeout: float = 600.0, output: bool = False):
    """Run the reference's block::initial (oracle/_ref/ref_metrics) in `path`;
    per rank: dict(jac, met, dxyz, centre, q).  block::initial also writes
    flowData/GEO__id_<rank>.dat; with `output`, block::Output then writes the
    Tecplot flow file flowData/NUC3d_step_<step>_id_<rank>.dat."""
    args = [REF_BIN_METRICS] + (["output"] if output else [])
    if nranks == 1:
        cmd = args
    else:
        mpiexec = shutil.which("mpiexec", path="/opt/conda/bin") or "/opt/conda/bin/mpiexec"
        cmd = [mpiexec, "-n", str(nranks)] + args
    subprocess.run(cmd, cwd=path, check=True, timeout=timeout, stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    out = []
    for r in range(nranks):
        fn = os.path.join(path, "metrics_%d.bin" % r)
        dims = np.fromfile(fn, dtype=np.int32, count=3)
        nx, ny, nz = (int(d) for d in dims)
        raw = np.fromfile(fn, dtype=np.float64, offset=12)
        N = nx * ny * nz
        sh = lambda a, c: a.reshape(c, nz, ny, nx)
        o = 0
        jac = raw[o:o + N].reshape(nz, ny, nx); o += N
        met = sh(raw[o:o + 9 * N], 9); o += 9 * N
        dxyz = sh(raw[o:o + 9 * N], 9); o += 9 * N
        centre = sh(raw[o:o + 3 * N], 3); o += 3 * N
        q = sh(raw[o:o + 5 * N], 5)
        out.append(dict(jac=jac, met=met, dxyz=dxyz, centre=centre, q=q))
    return out


def run_ref_post(path: str, nranks: int, timeout: float = 600.0):
    """The reference's postproc::solveQ on the case written by write_ref_case in
    `path` (ref_harness post): returns (enstrophy, kinetic) as it prints them
    to enstrophy.dat (12 significant digits, postproc.cpp:833-842)."""
    with open(os.path.join(path, "inp", "PostProc.in"), "w") as f:
        f.write("AverageType 0\nAverageStartStep 0\nSimutaneousData 1\nAveragedSaveType 0\n"
                "SimutaneousSaveType 0\n")
    with open(os.path.join(path, "inp", "IOController.in"), "w") as f:  # IOcontroller.cpp:39-62 exits without it
        f.write("startStep 0\nendStep 1\nsaveStep 1000\npostStep 1\nBenchmark 0\n"
                "cfl 0.5\ntl 0.0\ndt 0.01\nBinary no\nTecplot no\nPostProc yes\n")
    fn = os.path.join(path, "enstrophy.dat")
    if os.path.exists(fn):
        os.remove(fn)
    run_ref(path, nranks, "post", timeout=timeout)
    last = open(fn).read().split("\n")[-2].split()
    return float(last[2]), float(last[3])

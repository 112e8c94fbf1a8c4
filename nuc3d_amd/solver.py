"""Python view of the C ABI (include/nuc3d_hip.h).

`Case` holds every block of a case on the current GPU and mirrors the calls
the reference's driver makes per time step (block::solve, block.cpp:569-594):
`rhs()` is EulerData3D::solve / NaiverStokesData3d::solve, `stage(s, cfl)`
adds PDEData3d::solve for RK stage s, `steps(n, cfl)` runs whole TVD-RK3
steps.  Errors of the library raise `Nuc3dError` with its message.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import LIB_PATH

WENO5JS, WENO5Z, CRWENO5 = 0, 1, 2
CENTRAL8 = 3   # eight-cell central face value, width 4; no counterpart in the reference (the hybrid's smooth scheme)
HYBRID = 4     # per-face CENTRAL8 or WENO5-JS chosen by the learned sensor, width 4; no counterpart in the reference
SENSOR_NPARAM = 2145
CD6, CD8 = 0, 1
RECON_BOUNDARY = 3   # physical-face reconstruction closure (operator level)
DIFF_BOUNDARY = 2    # physical-face derivative closure (operator level)
PART_INNER, PART_LEFT, PART_RIGHT = 0, 1, 2
LF, AUSM, AUSMP = 0, 1, 2
EULER3D, NAVIERSTOKES3D = 0, 1


class Nuc3dError(RuntimeError):
    pass


class _CParams(ctypes.Structure):
    _fields_ = [
        ("gamma", ctypes.c_double), ("mach", ctypes.c_double), ("reynolds", ctypes.c_double),
        ("prandtl", ctypes.c_double), ("t_ref", ctypes.c_double), ("t_inf", ctypes.c_double),
        ("t_wall", ctypes.c_double), ("model", ctypes.c_int), ("riemann", ctypes.c_int),
        ("scheme", ctypes.c_int * 3), ("diff", ctypes.c_int * 3), ("width", ctypes.c_int * 3),
        ("zquirk", ctypes.c_int), ("exact", ctypes.c_int),
    ]


@dataclass
class Params:
    """inp/PhysModel.in + inp/Method.in of the reference (physicsModel.cpp:41-69,
    fieldOperator.cpp:20-31), plus `exact`: 1 (default) keeps the reference's
    operation order everywhere; 0 contracts the WENO arithmetic and uses a
    one-division form of the nonlinear weights (about a third faster per
    sweep, within 1e-12 on well-conditioned fields only)."""
    gamma: float = 1.4
    mach: float = 0.1
    reynolds: float = 1600.0
    prandtl: float = 0.72
    t_ref: float = 110.4
    t_inf: float = 288.0
    t_wall: float = 1.0
    model: int = NAVIERSTOKES3D
    riemann: int = LF
    scheme: Sequence[int] = (WENO5JS, WENO5JS, WENO5JS)
    diff: Sequence[int] = (CD6, CD6, CD6)
    width: Sequence[int] = (3, 3, 3)
    zquirk: int = 1
    exact: int = 1

    def to_c(self) -> _CParams:
        p = _CParams()
        for k in ("gamma", "mach", "reynolds", "prandtl", "t_ref", "t_inf", "t_wall", "model", "riemann", "zquirk",
                  "exact"):
            setattr(p, k, getattr(self, k))
        for i in range(3):
            p.scheme[i] = int(self.scheme[i])
            p.diff[i] = int(self.diff[i])
            p.width[i] = int(self.width[i])
        return p


_LIB = None


def lib():
    """Load libnuc3d_hip.so; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise Nuc3dError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    fp, llp = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_longlong)
    u16p = ctypes.POINTER(ctypes.c_uint16)
    sig = {
        "nuc3d_case_create": [ctypes.POINTER(_CParams), ctypes.c_int, ip, ctypes.POINTER(vp)],
        "nuc3d_case_destroy": [vp],
        "nuc3d_set_faces": [vp, ctypes.c_int, ip, ip, dp],
        "nuc3d_upload_metrics": [vp, ctypes.c_int, dp, dp, dp],
        "nuc3d_upload_q": [vp, ctypes.c_int, dp],
        "nuc3d_upload_mesh": [vp, ctypes.c_int, dp, dp, dp, dp],
        "nuc3d_save_q": [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_double],
        "nuc3d_diagnostics": [vp, dp, dp, ctypes.c_int, dp, vp],
        "nuc3d_get_residual": [vp, dp, dp],
        "nuc3d_run": [vp, ctypes.POINTER(_RunConfig), ctypes.POINTER(_RunStatus), vp],
        "nuc3d_load_q": [vp, ctypes.c_int, ctypes.c_char_p, dp],
        "nuc3d_download_metrics": [vp, ctypes.c_int, dp, dp, dp],
        "nuc3d_download_q": [vp, ctypes.c_int, dp],
        "nuc3d_download_rhs": [vp, ctypes.c_int, dp],
        "nuc3d_rhs": [vp, vp],
        "nuc3d_stage": [vp, ctypes.c_int, ctypes.c_double, vp],
        "nuc3d_steps": [vp, ctypes.c_int, ctypes.c_double, vp],
        "nuc3d_step_graph": [vp, ctypes.c_int, ctypes.c_double, vp],
        "nuc3d_sync": [vp],
        "nuc3d_get_dt": [vp, dp],
        "nuc3d_get_max_eigen": [vp, ctypes.c_int, dp],
        "nuc3d_set_block_group": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double],
        "nuc3d_get_dt_group": [vp, ctypes.c_int, dp],
        "nuc3d_autotune": [vp, vp, ip],
        "nuc3d_get_fixups": [vp, ctypes.POINTER(ctypes.c_longlong)],
        "nuc3d_get_positivity": [vp, llp],
        "nuc3d_set_division": [vp, ctypes.c_int],
        "nuc3d_diagnostics_noalloc": [vp, dp, dp, vp],
        "nuc3d_write_tecplot": [vp, ctypes.c_int, ctypes.c_char_p, dp, dp, dp],
        "nuc3d_get_sweep_tiles": [vp, ctypes.c_int, ip],
        "nuc3d_set_sweep_tiles": [vp, ctypes.c_int, ip],
        "nuc3d_write_tecplot_geo": [vp, ctypes.c_int, ctypes.c_char_p, dp, dp, dp],
        "nuc3d_selftest_division": [ctypes.c_longlong, ctypes.c_ulonglong, llp],
        "nuc3d_timing": [vp, ctypes.c_int],
        "nuc3d_timing_read": [vp, dp, ctypes.POINTER(ctypes.c_int)],
        "nuc3d_sensor_create": [ctypes.c_longlong, ctypes.POINTER(vp)],
        "nuc3d_sensor_destroy": [vp],
        "nuc3d_sensor_set_params": [vp, fp],
        "nuc3d_sensor_get_params": [vp, fp],
        "nuc3d_sensor_sample": [vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, llp],
        "nuc3d_sensor_count": [vp, llp, llp],
        "nuc3d_sensor_clear": [vp],
        "nuc3d_sensor_upload": [vp, u16p, fp, ctypes.c_longlong],
        "nuc3d_sensor_download": [vp, ctypes.c_longlong, ctypes.c_longlong, u16p, fp],
        "nuc3d_sensor_grad": [vp, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_float, fp, dp],
        "nuc3d_sensor_infer": [vp, ctypes.c_longlong, ctypes.c_longlong, fp],
        "nuc3d_sensor_train": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_ulonglong,
                               fp],
        "nuc3d_set_sensor": [vp, fp, ctypes.c_float],
        "nuc3d_download_hybrid_flags": [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_ubyte)],
        "nuc3d_get_hybrid_counts": [vp, llp],
        "nuc3d_set_sensor_interval": [vp, ctypes.c_int],
        "nuc3d_set_uniform_mesh": [vp, ctypes.c_int],
        "nuc3d_set_overlap": [vp, ctypes.c_int],
        "nuc3d_set_persistent": [vp, ctypes.c_int],
        "nuc3d_get_uniform_mesh": [vp, ctypes.c_int, ip],
        "nuc3d_op_interpolate": [ctypes.c_int] * 8 + [dp, dp, dp, vp],
        "nuc3d_op_differentiate": [ctypes.c_int] * 6 + [dp, dp, dp, vp],
        "nuc3d_op_rk3": [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_longlong,
                         ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.POINTER(dp), vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.nuc3d_has_crweno5.argtypes = []
    L.nuc3d_has_crweno5.restype = ctypes.c_int
    L.nuc3d_last_error.argtypes = []
    L.nuc3d_last_error.restype = ctypes.c_char_p
    _LIB = L
    return L


def selftest_division(n, seed=0):
    """Quotients of the reduced FP64 division that differ from a / b over n
    hashed operand pairs (x7), and square roots of |a / b| by the shortened
    square root that differ from sqrt, on the device."""
    v = ctypes.c_longlong()
    _check(lib().nuc3d_selftest_division(int(n), int(seed), ctypes.byref(v)), "selftest_division")
    return v.value


def crweno5_available():
    """Whether the library implements the CRWENO5 line solves."""
    return hasattr(lib(), "nuc3d_has_crweno5") and lib().nuc3d_has_crweno5() == 1


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc, what):
    if rc != 0:
        msg = lib().nuc3d_last_error().decode(errors="replace")
        raise Nuc3dError(f"{what} failed ({rc}): {msg}")


class _RunConfig(ctypes.Structure):
    _fields_ = [("start_step", ctypes.c_int), ("end_step", ctypes.c_int), ("save_step", ctypes.c_int),
                ("post_step", ctypes.c_int), ("cfl", ctypes.c_double), ("tl", ctypes.c_double),
                ("time", ctypes.c_double), ("dir", ctypes.c_char_p), ("verbose", ctypes.c_int)]


class _RunStatus(ctypes.Structure):
    _fields_ = [("step", ctypes.c_int), ("time", ctypes.c_double), ("dt", ctypes.c_double),
                ("residual", ctypes.c_double)]


class Case:
    """All blocks of a case resident on the current GPU."""

    def __init__(self, params: Params, blocks, stream: Optional[int] = None):
        L = lib()
        self.params = params
        self.blocks = list(blocks)
        self.stream = ctypes.c_void_p(stream) if stream else None
        self._cp = params.to_c()
        dims = (ctypes.c_int * (3 * len(self.blocks)))(*[int(d) for b in self.blocks for d in b.dims])
        h = ctypes.c_void_p()
        _check(L.nuc3d_case_create(ctypes.byref(self._cp), len(self.blocks), dims, ctypes.byref(h)), "case_create")
        self.h = h
        for bi, b in enumerate(self.blocks):
            t = (ctypes.c_int * 6)(*[int(f[0]) for f in b.faces])
            i = (ctypes.c_int * 6)(*[int(f[1]) for f in b.faces])
            vals = np.zeros(30)
            for fi, f in enumerate(b.faces):
                if len(f) > 2 and f[2] is not None:
                    vals[5 * fi:5 * fi + len(f[2])] = f[2]
            _check(L.nuc3d_set_faces(self.h, bi, t, i, _dp(vals)), "set_faces")
            # blocks given by their mesh nodes get metrics later (upload_mesh)
            if b.jac is not None:
                jac = np.ascontiguousarray(b.jac, dtype=np.float64)
                met = np.ascontiguousarray(b.met, dtype=np.float64)
                dxyz = np.ascontiguousarray(b.dxyz, dtype=np.float64)
                _check(L.nuc3d_upload_metrics(self.h, bi, _dp(jac), _dp(met), _dp(dxyz)), "upload_metrics")
            if b.q is not None:
                self.set_q(bi, b.q)

    def close(self):
        if getattr(self, "h", None):
            lib().nuc3d_case_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, bi, nc):
        nx, ny, nz = self.blocks[bi].dims
        return (nc, nz, ny, nx)

    def set_q(self, bi, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        _check(lib().nuc3d_upload_q(self.h, bi, _dp(q)), "upload_q")

    def get_q(self, bi=0):
        out = np.empty(self._shape(bi, 5))
        _check(lib().nuc3d_download_q(self.h, bi, _dp(out)), "download_q")
        return out

    def get_rhs(self, bi=0):
        out = np.empty(self._shape(bi, 5))
        _check(lib().nuc3d_download_rhs(self.h, bi, _dp(out)), "download_rhs")
        return out

    def rhs(self):
        _check(lib().nuc3d_rhs(self.h, self.stream), "rhs")

    def stage(self, s, cfl):
        _check(lib().nuc3d_stage(self.h, s, cfl, self.stream), "stage")

    def residual(self):
        """(global, per block) residual of the last step (PDEData3d::setRES)."""
        g = ctypes.c_double()
        per = np.empty(len(self.blocks))
        _check(lib().nuc3d_get_residual(self.h, ctypes.byref(g), _dp(per)), "get_residual")
        return g.value, per

    def run(self, end_step, cfl=0.5, tl=1e30, save_step=0, post_step=0, start_step=0, time=0.0, dir=None,
            verbose=False):
        """singleBlock::solvePDE with the IOController cadence; returns
        dict(step, time, dt, residual)."""
        cfg = _RunConfig(start_step, end_step, save_step, post_step, cfl, tl, time,
                         os.fsencode(dir) if dir else None, 1 if verbose else 0)
        st = _RunStatus()
        _check(lib().nuc3d_run(self.h, ctypes.byref(cfg), ctypes.byref(st), self.stream), "run")
        return dict(step=st.step, time=st.time, dt=st.dt, residual=st.residual)

    def diagnostics(self, b=0):
        """(enstrophy, kinetic, fields5 of block b) of postproc::solveQ."""
        e, k = ctypes.c_double(), ctypes.c_double()
        f = np.empty(self._shape(b, 5))
        _check(lib().nuc3d_diagnostics(self.h, ctypes.byref(e), ctypes.byref(k), b, _dp(f), self.stream),
               "diagnostics")
        return e.value, k.value, f

    def diagnostics_noalloc(self):
        """(enstrophy, kinetic) sums through the entry point that allocates
        nothing (fixed device scratch of the library)."""
        e, k = ctypes.c_double(), ctypes.c_double()
        _check(lib().nuc3d_diagnostics_noalloc(self.h, ctypes.byref(e), ctypes.byref(k), self.stream),
               "diagnostics_noalloc")
        return e.value, k.value

    def write_tecplot(self, b, path, xn, yn, zn, geo=False):
        """Tecplot file of block b (block::outputQ_tecplot, or with geo the
        metric file of block::outputGEO_tecplot) over the node coordinates."""
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (xn, yn, zn)]
        fn = lib().nuc3d_write_tecplot_geo if geo else lib().nuc3d_write_tecplot
        _check(fn(self.h, int(b), os.fsencode(path), _dp(a[0]), _dp(a[1]), _dp(a[2])), "write_tecplot")

    def save_q(self, b, path, time):
        """Restart file of block b in the reference's format (block.cpp:754-778)."""
        _check(lib().nuc3d_save_q(self.h, b, os.fsencode(path), time), "save_q")

    def load_q(self, b, path):
        """Loads a restart file (block.cpp:780-810) into block b; returns its time."""
        t = ctypes.c_double()
        _check(lib().nuc3d_load_q(self.h, b, os.fsencode(path), ctypes.byref(t)), "load_q")
        return t.value

    def upload_mesh(self, b, xn, yn, zn):
        """Metrics of block b computed on the GPU from node coordinates of shape
        (nz+1, ny+1, nx+1); returns the cell centres (3, nz, ny, nx)."""
        xn, yn, zn = (np.ascontiguousarray(a, dtype=np.float64) for a in (xn, yn, zn))
        nz, ny, nx = (d - 1 for d in xn.shape)
        centre = np.empty((3, nz, ny, nx))
        _check(lib().nuc3d_upload_mesh(self.h, b, _dp(xn), _dp(yn), _dp(zn), _dp(centre)), "upload_mesh")
        return centre

    def get_metrics(self, b):
        """(jac, met, dxyz) of block b as held on the GPU."""
        nx, ny, nz = self.blocks[b].dims
        jac, met, dxyz = np.empty((nz, ny, nx)), np.empty((9, nz, ny, nx)), np.empty((9, nz, ny, nx))
        _check(lib().nuc3d_download_metrics(self.h, b, _dp(jac), _dp(met), _dp(dxyz)), "download_metrics")
        return jac, met, dxyz

    def steps(self, n, cfl):
        _check(lib().nuc3d_steps(self.h, n, cfl, self.stream), "steps")

    def steps_graph(self, n, cfl):
        """n steps replayed from a captured two-step HIP graph (bit-identical
        to steps())."""
        _check(lib().nuc3d_step_graph(self.h, n, cfl, self.stream), "step_graph")

    def sync(self):
        _check(lib().nuc3d_sync(self.h), "sync")

    def dt(self):
        v = ctypes.c_double()
        _check(lib().nuc3d_get_dt(self.h, ctypes.byref(v)), "get_dt")
        return v.value

    def set_group(self, b, group, mach=0.0, reynolds=0.0):
        """Block b belongs to independent case `group` (own dt, Mach and Reynolds numbers)."""
        _check(lib().nuc3d_set_block_group(self.h, b, group, float(mach), float(reynolds)), "set_block_group")

    def autotune(self):
        """Times the pointwise passes per block shape and keeps the fastest k
        segment of the column tiles; returns the choice per block."""
        out = (ctypes.c_int * len(self.blocks))()
        _check(lib().nuc3d_autotune(self.h, self.stream, out), "autotune")
        return list(out)

    def dt_group(self, group):
        v = ctypes.c_double()
        _check(lib().nuc3d_get_dt_group(self.h, group, ctypes.byref(v)), "get_dt_group")
        return v.value

    def max_eigen(self, bi=0):
        out = np.empty(3)
        _check(lib().nuc3d_get_max_eigen(self.h, bi, _dp(out)), "get_max_eigen")
        return out

    def fixups(self):
        v = ctypes.c_longlong()
        _check(lib().nuc3d_get_fixups(self.h, ctypes.byref(v)), "get_fixups")
        return v.value

    def sweep_tiles(self, b):
        """Faces per tile line of block b's x, y, z sweeps (16 or 32)."""
        v = (ctypes.c_int * 3)()
        _check(lib().nuc3d_get_sweep_tiles(self.h, int(b), v), "get_sweep_tiles")
        return tuple(int(x) for x in v)

    def set_sweep_tiles(self, b, nf3):
        v = (ctypes.c_int * 3)(*[int(x) for x in nf3])
        _check(lib().nuc3d_set_sweep_tiles(self.h, int(b), v), "set_sweep_tiles")

    def set_division(self, ieee):
        """ieee: the compiler's IEEE division everywhere (same results)."""
        _check(lib().nuc3d_set_division(self.h, 1 if ieee else 0), "set_division")

    def positivity(self):
        """Cells that failed the positivity check since the state was set:
        (non-positive density, negative pressure after the e<0 fix-up)."""
        v = (ctypes.c_longlong * 2)()
        _check(lib().nuc3d_get_positivity(self.h, v), "get_positivity")
        return int(v[0]), int(v[1])

    def timing(self, enable=True):
        _check(lib().nuc3d_timing(self.h, 1 if enable else 0), "timing")

    TIMING_CLASSES = ("x_sweep", "y_sweep", "z_sweep", "primitives", "viscous_flux", "viscous_div_update")

    def set_sensor(self, params, threshold=0.0):
        """The learned sensor of the hybrid directions: SENSOR_NPARAM FP32
        parameters and the logit threshold at and above which a face takes
        WENO5-JS."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == SENSOR_NPARAM
        _check(lib().nuc3d_set_sensor(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), float(threshold)),
               "set_sensor")

    def set_overlap(self, enable):
        """Run the viscous-flux pass on a second stream beside the x and y
        sweeps (default on; results are identical)."""
        _check(lib().nuc3d_set_overlap(self.h, 1 if enable else 0), "set_overlap")

    def set_persistent(self, enable):
        """Launch each direction sweep as one persistent grid that takes
        (block, tile) work items from a device queue (default off; results
        are identical)."""
        _check(lib().nuc3d_set_persistent(self.h, 1 if enable else 0), "set_persistent")

    def set_uniform_mesh(self, enable):
        """Let blocks of one J and one value per metric term take them from
        their descriptor (default on; results are identical)."""
        _check(lib().nuc3d_set_uniform_mesh(self.h, 1 if enable else 0), "set_uniform_mesh")

    def uniform_mesh(self, b):
        """Whether block b takes the uniform-mesh path."""
        v = ctypes.c_int(0)
        _check(lib().nuc3d_get_uniform_mesh(self.h, int(b), ctypes.byref(v)), "get_uniform_mesh")
        return bool(v.value)

    def set_sensor_interval(self, steps):
        """Steps between hybrid decisions (1 or 2)."""
        _check(lib().nuc3d_set_sensor_interval(self.h, int(steps)), "set_sensor_interval")

    def hybrid_flags(self, b, d):
        """Last decisions (of nuc3d_rhs, or of the current step's stage 0) for
        direction d of block b, in
        the face-array layout (n_{d+2}, n_{d+1}, n_d + 1): 0 central, 1 WENO5-JS
        by the network, 2 WENO5-JS by the fallback, 3 physical-face closure."""
        n = self.blocks[b].dims
        shape = (n[(d + 2) % 3], n[(d + 1) % 3], n[d] + 1)
        out = np.empty(shape, dtype=np.uint8)
        _check(lib().nuc3d_download_hybrid_flags(self.h, b, d, out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))),
               "download_hybrid_flags")
        return out

    def hybrid_counts(self):
        """Faces decided so far: (central, WENO5-JS by the network, WENO5-JS by the fallback)."""
        v = (ctypes.c_longlong * 3)()
        _check(lib().nuc3d_get_hybrid_counts(self.h, v), "get_hybrid_counts")
        return tuple(v)

    def timing_read(self):
        """{class: (total_ms, launches)} since the last read."""
        ms = (ctypes.c_double * len(self.TIMING_CLASSES))()
        n = (ctypes.c_int * len(self.TIMING_CLASSES))()
        _check(lib().nuc3d_timing_read(self.h, ms, n), "timing_read")
        return {k: (ms[i], n[i]) for i, k in enumerate(self.TIMING_CLASSES)}


# ---------------------------------------------------------------------------
# operator level (include/nuc3d_ops.h): one Field operation per call on host
# arrays of shape (nz, ny, nx), the operator acting along x


def _field(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def op_interpolate(scheme, part, upwind, w, field, ghost, out, zquirk=1):
    """interpolationInner / BoundaryL / BoundaryR of the reference's scheme
    classes (fieldOperator.h:23-38) into `out` (nz, ny, nx+1), in place."""
    f = _field(field)
    nz, ny, nx = f.shape
    g = None if ghost is None else _field(ghost)
    assert out.dtype == np.float64 and out.flags["C_CONTIGUOUS"] and out.shape == (nz, ny, nx + 1)
    _check(lib().nuc3d_op_interpolate(scheme, part, upwind, w, zquirk, nx, ny, nz, _dp(f),
                                      None if g is None else _dp(g), _dp(out), None), "op_interpolate")
    return out


def op_differentiate(scheme, part, w, field, ghost, out):
    """differentialInner / BoundaryL / BoundaryR (fieldOperator.h:68-79) into
    `out` (nz, ny, nx), in place."""
    f = _field(field)
    nz, ny, nx = f.shape
    g = None if ghost is None else _field(ghost)
    assert out.dtype == np.float64 and out.flags["C_CONTIGUOUS"] and out.shape == f.shape
    _check(lib().nuc3d_op_differentiate(scheme, part, w, nx, ny, nz, _dp(f), None if g is None else _dp(g),
                                        _dp(out), None), "op_differentiate")
    return out


def op_rk3(stage, dt, rhs, u_n, u_i):
    """tvdrk3rd::integrationAll (TVD-RK3.cpp:12-137) over lists of fields;
    updates the arrays of u_i in place."""
    rs = [_field(a) for a in rhs]
    ns = [_field(a) for a in u_n]
    for a in u_i:
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    P = ctypes.POINTER(ctypes.c_double)
    arr = lambda xs: (P * len(xs))(*[_dp(x) for x in xs])
    _check(lib().nuc3d_op_rk3(stage, dt, len(u_i), u_i[0].size, arr(rs), arr(ns), arr(u_i), None), "op_rk3")
    return u_i


# ---------------------------------------------------------------------------
# learned sensor of the hybrid reconstruction (no counterpart in the reference)

class Sensor:
    """The sensor network's parameters, Adam state and a device-resident
    dataset of (30 bf16 features, label) samples; sampling from cases and
    training run on the GPU (include/nuc3d_hip.h)."""

    def __init__(self, capacity):
        h = ctypes.c_void_p()
        _check(lib().nuc3d_sensor_create(int(capacity), ctypes.byref(h)), "sensor_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().nuc3d_sensor_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _fp(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == SENSOR_NPARAM
        _check(lib().nuc3d_sensor_set_params(self.h, self._fp(p)), "sensor_set_params")

    def params(self):
        p = np.empty(SENSOR_NPARAM, dtype=np.float32)
        _check(lib().nuc3d_sensor_get_params(self.h, self._fp(p)), "sensor_get_params")
        return p

    def sample(self, case, tau, stride=1, offset=0):
        n = ctypes.c_longlong()
        _check(lib().nuc3d_sensor_sample(self.h, case.h, float(tau), int(stride), int(offset), ctypes.byref(n)),
               "sensor_sample")
        return n.value

    def count(self):
        n, pos = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().nuc3d_sensor_count(self.h, ctypes.byref(n), ctypes.byref(pos)), "sensor_count")
        return n.value, pos.value

    def clear(self):
        _check(lib().nuc3d_sensor_clear(self.h), "sensor_clear")

    def upload(self, xbits, y):
        x = np.ascontiguousarray(xbits, dtype=np.uint16)
        yy = np.ascontiguousarray(y, dtype=np.float32)
        assert x.shape == (yy.size, 32)
        _check(lib().nuc3d_sensor_upload(self.h, x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), self._fp(yy),
                                         yy.size), "sensor_upload")

    def download(self, first=0, n=None):
        if n is None:
            n = self.count()[0] - first
        x = np.empty((n, 32), dtype=np.uint16)
        y = np.empty(n, dtype=np.float32)
        _check(lib().nuc3d_sensor_download(self.h, first, n, x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                           self._fp(y)), "sensor_download")
        return x, y

    def grad(self, first, n, pos_weight=1.0):
        g = np.empty(SENSOR_NPARAM, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().nuc3d_sensor_grad(self.h, first, n, float(pos_weight), self._fp(g), ctypes.byref(loss)),
               "sensor_grad")
        return loss.value, g

    def infer(self, first=0, n=None):
        if n is None:
            n = self.count()[0] - first
        z = np.empty(n, dtype=np.float32)
        _check(lib().nuc3d_sensor_infer(self.h, first, n, self._fp(z)), "sensor_infer")
        return z

    def train(self, steps, batch=8192, lr=1e-3, pos_weight=1.0, seed=0):
        losses = np.empty(max(steps, 1), dtype=np.float32)
        _check(lib().nuc3d_sensor_train(self.h, int(steps), int(batch), float(lr), float(pos_weight), int(seed),
                                        self._fp(losses)), "sensor_train")
        return losses[:steps]

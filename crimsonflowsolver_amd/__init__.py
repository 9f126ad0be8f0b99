"""MI355X-native stabilized-FEM formation and memLS solve for CRIMSON Flowsolver.

The product is the C-ABI library libcrimson_amd.so (HIP kernels for gfx950,
declared in include/crimson_amd.h).  This package is its host-side mirror:
ctypes bindings whose names and argument meanings follow the reference's
interfaces for this path:

* ``ElmGMRInterior`` -- the interior part of ``ElmGMR`` (elmgmr.f90:1-255):
  flux reconstruction, e3 formation, residual scatter and fillsparseI.
* ``memLS_LS_CREATE`` / ``memLS_LHS_CREATE`` / ``memLS_BC_CREATE`` /
  ``memLS_SOLVE`` -- memLS.f90:1379,1080,272,1947 with the same argument
  order, optional arguments and in/out semantics (``Ri`` is overwritten by the
  solution).

Errors raise ``CrimsonError`` carrying the library's status code, the
counterpart of the reference's PRINT + STOP.  There is no CPU fallback: if the
library is missing or no GPU is visible, the calls fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcrimson_amd.so")

LS_TYPE_CG, LS_TYPE_GMRES, LS_TYPE_NS = 1, 2, 3
BC_TYPE_Dir, BC_TYPE_Neu = 0, 1
FLUX_NOALLOC_MAX_SURFACES = 4096  # CRIMSON_FLUX_NOALLOC_MAX_SURFACES

ERRORS = {1: "EARG", 2: "EUNSUPPORTED", 3: "EGRAPH", 4: "ENODIAG", 5: "EHIP", 6: "ENOMEM",
          7: "ESINGULAR", 8: "ESTATE"}

EXPORTS = (
    "crimson_last_error", "crimson_device_count", "crimson_form_setup", "crimson_form_interior",
    "crimson_form_interior_dev", "crimson_form_interior_to_lhs", "crimson_form_bc3res_dev",
    "crimson_form_timing", "crimson_form_free", "crimson_genadj", "crimson_ls_create", "crimson_ls_free",
    "crimson_lhs_create", "crimson_lhs_free", "crimson_bc_create", "crimson_ls_solve",
    "crimson_ls_solve_dev", "crimson_lhs_set_val", "crimson_lhs_get_val", "crimson_ls_stats",
    "crimson_lhs_spmv_dev", "crimson_itr_predict_dev", "crimson_itr_yalpha_dev", "crimson_itr_correct_dev",
    "crimson_itr_bc_dev", "crimson_elmpvsq_dev", "crimson_form_sync", "crimson_lhs_set_precision",
    "crimson_form_vjp_dev", "crimson_form_set_instances", "crimson_form_graph_create", "crimson_graph_launch",
    "crimson_graph_free", "crimson_form_set_boundary", "crimson_form_autotune", "crimson_form_set_bc",
    "crimson_elmpvsq_coef_dev", "crimson_form_boundary_flux", "crimson_form_boundary_flux_noalloc",
    "crimson_measure_peaks", "crimson_form_vjp_ac_dev", "crimson_elmpvsq_vjp_dev", "crimson_lhs_transpose_dev",
    "crimson_ls_graph_create", "crimson_step_graph_create", "crimson_form_set_wall",
    "crimson_form_set_displacement_dev", "crimson_wall_predict_dev", "crimson_wall_correct_dev",
    "crimson_wall_alpha_dev", "crimson_form_set_wall_state_dev",
)


class CrimsonError(RuntimeError):
    def __init__(self, code, where, text):
        super().__init__(f"{where}: {ERRORS.get(code, code)}: {text}")
        self.code = code


class FormParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("Delt", C.c_double),
                ("alfi", C.c_double), ("almi", C.c_double), ("gami", C.c_double),
                ("dtsfct", C.c_double), ("taucfct", C.c_double), ("flmpl", C.c_double),
                ("flmpr", C.c_double), ("itau", C.c_int), ("iconvflow", C.c_int),
                ("idiff", C.c_int), ("ipord", C.c_int), ("precision", C.c_int),
                ("aspect_fp64", C.c_double), ("ideformwall", C.c_int)]


class WallParams(C.Structure):
    """crimson_wall_params: the vessel wall's values (solver.inp)."""
    _fields_ = [("rhovw", C.c_double), ("thicknessvw", C.c_double), ("evw", C.c_double), ("rnuvw", C.c_double),
                ("rshearconstantvw", C.c_double), ("tissSuppStiffCoeff", C.c_double),
                ("tissSuppDampCoeff", C.c_double), ("tissSuppRingStiffCoeff", C.c_double),
                ("tissSuppRingDampCoeff", C.c_double), ("iwallmassfactor", C.c_int), ("iwallstiffactor", C.c_int),
                ("iwallsupp", C.c_int), ("iwalldamp", C.c_int), ("iringsupp", C.c_int), ("iringdamp", C.c_int)]


def wall_params(rhovw=1.06, thicknessvw=0.3, evw=7.0e5, rnuvw=0.5, rshearconstantvw=0.833333,
                tissSuppStiffCoeff=0.1, tissSuppDampCoeff=0.01, tissSuppRingStiffCoeff=5e5,
                tissSuppRingDampCoeff=0.01, iwallmassfactor=1, iwallstiffactor=1, iwallsupp=1, iwalldamp=1,
                iringsupp=1, iringdamp=1):
    """The wall block of solver.inp ("Deformable Wall: True"); the defaults
    are the reference's basic case's values with every term on."""
    return WallParams(rhovw, thicknessvw, evw, rnuvw, rshearconstantvw, tissSuppStiffCoeff, tissSuppDampCoeff,
                      tissSuppRingStiffCoeff, tissSuppRingDampCoeff, iwallmassfactor, iwallstiffactor, iwallsupp,
                      iwalldamp, iringsupp, iringdamp)


_lib = None


def lib():
    """The loaded library; raises if it was not built.

    PyTorch ships its own HIP runtime with the same SONAME as ROCm's
    (libamdhip64.so.7).  Importing torch first makes this library bind to that
    already-loaded runtime, so a process that also allocates through torch
    holds a single HIP runtime instead of two that contend for the device."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                              "(make -C crimsonflowsolver_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.crimson_last_error.argtypes = [C.c_char_p, C.c_int]
        L.crimson_device_count.argtypes = []
        L.crimson_form_setup.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int, vp,
                                         C.POINTER(FormParams), C.POINTER(C.c_void_p)]
        L.crimson_form_interior.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.c_int]
        L.crimson_form_interior_dev.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.c_int]
        L.crimson_form_interior_to_lhs.argtypes = [vp, vp, vp, vp, vp, vp]
        L.crimson_form_bc3res_dev.argtypes = [vp, vp, C.c_int]
        L.crimson_form_timing.argtypes = [vp, dp]
        L.crimson_form_free.argtypes = [vp]
        L.crimson_genadj.argtypes = [C.c_int, C.c_int, vp, vp, vp, C.c_int, ip]
        L.crimson_ls_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                        vp, vp, vp, C.POINTER(C.c_void_p)]
        L.crimson_ls_free.argtypes = [vp]
        L.crimson_lhs_create.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int,
                                         C.POINTER(C.c_void_p)]
        L.crimson_lhs_free.argtypes = [vp]
        L.crimson_bc_create.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]
        L.crimson_ls_solve.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp]
        L.crimson_ls_solve_dev.argtypes = [vp, vp, C.c_int, vp, vp, vp]
        L.crimson_lhs_set_val.argtypes = [vp, C.c_int, vp]
        L.crimson_lhs_get_val.argtypes = [vp, C.c_int, vp]
        L.crimson_ls_stats.argtypes = [vp, C.c_int, dp]
        L.crimson_lhs_spmv_dev.argtypes = [vp, C.c_int, vp, vp, C.c_int, dp]
        L.crimson_itr_predict_dev.argtypes = [vp, vp, vp, vp, vp, C.c_double]
        L.crimson_itr_yalpha_dev.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_double, C.c_double]
        L.crimson_itr_correct_dev.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, C.c_double]
        L.crimson_itr_bc_dev.argtypes = [vp, vp, vp, C.c_int, vp]
        L.crimson_elmpvsq_dev.argtypes = [vp, C.c_int, vp, vp, C.c_double, vp, vp, vp]
        L.crimson_form_sync.argtypes = [vp]
        L.crimson_form_set_boundary.argtypes = [vp, C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_double, vp, C.c_int]
        L.crimson_form_set_bc.argtypes = [vp, vp, C.c_int]
        L.crimson_form_boundary_flux.argtypes = [vp, C.c_int, vp]
        L.crimson_form_boundary_flux_noalloc.argtypes = [vp, C.c_int, vp]
        L.crimson_measure_peaks.argtypes = [C.c_size_t, dp, dp, dp]
        L.crimson_elmpvsq_coef_dev.argtypes = [vp, C.c_int, vp, vp, C.c_double, C.c_double, vp, vp, vp]
        L.crimson_form_vjp_ac_dev.argtypes = [vp, vp, vp, vp, vp, vp, vp, dp]
        L.crimson_elmpvsq_vjp_dev.argtypes = [vp, C.c_int, vp, vp, C.c_double, vp, vp, vp, vp]
        L.crimson_lhs_transpose_dev.argtypes = [vp, vp]
        L.crimson_ls_graph_create.argtypes = [vp, vp, C.c_int, vp, vp, vp, C.POINTER(C.c_void_p)]
        L.crimson_step_graph_create.argtypes = ([vp] * 13 + [C.c_int, C.c_int] + [vp] * 7 +
                                                [C.c_int, C.POINTER(C.c_void_p)])
        L.crimson_form_set_wall.argtypes = [vp, C.POINTER(WallParams), vp, vp]
        L.crimson_form_set_displacement_dev.argtypes = [vp, vp]
        L.crimson_form_set_wall_state_dev.argtypes = [vp, vp, vp]
        L.crimson_wall_predict_dev.argtypes = [vp, vp, vp, vp, vp]
        L.crimson_wall_correct_dev.argtypes = [vp, vp, vp]
        L.crimson_wall_alpha_dev.argtypes = [vp, vp, vp, vp]
        L.crimson_graph_launch.argtypes = [vp]
        L.crimson_graph_free.argtypes = [vp]
        L.crimson_form_autotune.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.POINTER(C.c_int), dp]
        _lib = L
    return _lib


def measure_peaks(copy_bytes=1 << 31):
    """Measured device ceilings: {copy_gbs, fp64_tflops, fp32_tflops}."""
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    check(lib().crimson_measure_peaks(int(copy_bytes), C.byref(a), C.byref(b), C.byref(c)), "crimson_measure_peaks")
    return {"copy_gbs": a.value, "fp64_tflops": b.value, "fp32_tflops": c.value}


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().crimson_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc, where):
    if rc != 0:
        raise CrimsonError(rc, where, last_error())


def device_count() -> int:
    return lib().crimson_device_count()


def _f(a, dt=np.float64):
    return np.asfortranarray(a, dtype=dt)


def _p(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):  # torch tensor on the device
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


def form_params(rho=1.06, mu=0.04, delt=1e-3, rhoinf=0.5, precision=64, aspect_fp64=20.0, ideformwall=0):
    """Material and generalized-alpha parameters (itrPC.f90:16-37);
    ideformwall 1: deformable walls (set_wall before the first formation)."""
    if rhoinf < 0 or rhoinf > 1:
        almi = alfi = gami = 1.0
    else:
        almi = (3.0 - rhoinf) / (1.0 + rhoinf) / 2.0
        alfi = 1.0 / (1.0 + rhoinf)
        gami = 0.5 + almi - alfi
    return FormParams(rho, mu, delt, alfi, almi, gami, 1.0, 1.0, 0.0, 0.0, 0, 2, 1, 1,
                      precision, aspect_fp64, ideformwall)


def genadj(ien, nshg):
    """Native genadj: (colm, rowp) 1-based."""
    ien = _f(ien, np.int32)
    nnz = C.c_int()
    colm = np.zeros(nshg + 1, np.int32)
    lib().crimson_genadj(nshg, ien.shape[0], _p(ien), _p(colm), None, 0, C.byref(nnz))
    rowp = np.zeros(nnz.value, np.int32)
    check(lib().crimson_genadj(nshg, ien.shape[0], _p(ien), _p(colm), _p(rowp), nnz.value,
                               C.byref(nnz)), "genadj")
    return colm, rowp


class FormGraph:
    """A captured HIP graph (crimson_form_graph_create, crimson_ls_graph_create,
    crimson_step_graph_create); keeps its arrays alive."""

    def __init__(self, h, keep):
        self.h = h
        self._keep = keep

    def launch(self):
        check(lib().crimson_graph_launch(self.h), "crimson_graph_launch")

    def __del__(self):
        try:
            lib().crimson_graph_free(self.h)
        except Exception:
            pass


class ElmGMRInterior:
    """Interior ElmGMR on the GPU (elmgmr.f90:82-255).

    ``ien`` (nel,4) 1-based, ``colm`` (nshg+1) and ``rowp`` (nnz) 1-based as
    genadj builds them, ``iBC`` (nshg) optional."""

    def __init__(self, ien, colm, rowp, nshg, numnp=None, iBC=None, params=None):
        self.ien = _f(ien, np.int32)
        self.colm = _f(colm, np.int32)
        self.rowp = _f(rowp, np.int32)
        self.nshg = nshg
        self.numnp = nshg if numnp is None else numnp
        self.nnz = self.rowp.shape[0]
        self.params = params if params is not None else form_params()
        ibc = None if iBC is None else _f(iBC, np.int32)
        h = C.c_void_p()
        check(lib().crimson_form_setup(nshg, self.numnp, self.ien.shape[0], _p(self.ien), _p(self.colm),
                                       _p(self.rowp), self.nnz, _p(ibc), C.byref(self.params), C.byref(h)),
              "crimson_form_setup")
        self.h = h

    def __call__(self, y, ac, x, lhs=1, want_q=False):
        """Returns res(nshg,4), lhsK(9,nnz), lhsP(4,nnz)[, qres(nshg,9)]."""
        y, ac, x = _f(y), _f(ac), _f(x)
        res = np.zeros((self.nshg, 4), order="F")
        K = np.zeros((9, self.nnz), order="F") if lhs else None
        P = np.zeros((4, self.nnz), order="F") if lhs else None
        q = np.zeros((self.nshg, 9), order="F") if want_q else None
        check(lib().crimson_form_interior(self.h, _p(y), _p(ac), _p(x), _p(res), _p(K), _p(P), _p(q),
                                          1 if lhs else 0), "crimson_form_interior")
        return (res, K, P, q) if want_q else (res, K, P)

    def dev(self, y, ac, x, res, lhsK=None, lhsP=None, qres=None, lhs=1):
        """Device-pointer form on torch tensors laid out as the host arrays."""
        check(lib().crimson_form_interior_dev(self.h, _p(y), _p(ac), _p(x), _p(res), _p(lhsK), _p(lhsP),
                                              _p(qres), 1 if lhs else 0), "crimson_form_interior_dev")

    def to_lhs(self, y, ac, x, res4, lhs_handle):
        """Formation into the memLS matrix of ``lhs_handle``; ``res4`` receives
        Res4(4,nshg) (a (nshg,4) C-ordered torch tensor)."""
        check(lib().crimson_form_interior_to_lhs(self.h, _p(y), _p(ac), _p(x), _p(res4), lhs_handle.h),
              "crimson_form_interior_to_lhs")

    def vjp(self, y, ac, x, w, ybar):
        """Adjoint of the interior residual on device tensors laid out like
        res: ybar <- w^T d res/d y (y = u1,u2,u3,p); returns w^T d res/d mu."""
        mb = C.c_double()
        check(lib().crimson_form_vjp_dev(self.h, _p(y), _p(ac), _p(x), _p(w), _p(ybar), C.byref(mb)),
              "crimson_form_vjp_dev")
        return mb.value

    def vjp_ac(self, y, ac, x, w, ybar, acbar):
        """vjp with the acceleration's cotangent: ybar <- w^T d res/d y,
        acbar <- w^T d res/d ac (fourth column zero); returns w^T d res/d mu."""
        mb = C.c_double()
        check(lib().crimson_form_vjp_ac_dev(self.h, _p(y), _p(ac), _p(x), _p(w), _p(ybar), _p(acbar),
                                            C.byref(mb)), "crimson_form_vjp_ac_dev")
        return mb.value

    def elmpvsq_vjp(self, glob_dev, nabi_dev, c1, y, w4, ybar, cbar):
        """Adjoint of elmpvsq_coef's term for w4 laid out like Res4: ybar
        (like y) gets the face term added; cbar (2 doubles, device) <- the
        cotangents of c1 and c2."""
        check(lib().crimson_elmpvsq_vjp_dev(self.h, int(glob_dev.numel()), _p(glob_dev), _p(nabi_dev), float(c1),
                                            _p(y), _p(w4), _p(ybar), _p(cbar)), "crimson_elmpvsq_vjp_dev")

    def set_instances(self, inst, mu):
        """Ensemble: inst(nel) 0-based instance of each element, mu(ninst);
        an empty mu (inst None) returns to the single viscosity."""
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        inst = None if mu.size == 0 else np.ascontiguousarray(inst, dtype=np.uint16)
        check(lib().crimson_form_set_instances(self.h, _p(inst), int(mu.size), _p(mu)), "crimson_form_set_instances")

    def set_boundary(self, ienb, iBCB, BCB, nqb=3, influx=0.5, permit=None):
        """Boundary elements (the AsBMFG loop): ienb(nelb,4) 1-based with the
        face nodes first, iBCB(nelb,2), BCB(nelb,3,ndBCB); permit(0..nsrf)
        flow-permitted flags by srfID or None.  Every later formation adds
        their residual rows; pass nelb = 0 rows to remove them."""
        ienb = np.asfortranarray(ienb, dtype=np.int32)
        iBCB = np.asfortranarray(iBCB, dtype=np.int32)
        BCB = np.asfortranarray(BCB, dtype=np.float64)
        if BCB.ndim != 3 or BCB.shape[:2] != (ienb.shape[0], 3) or iBCB.shape != (ienb.shape[0], 2):
            raise ValueError("ienb(nelb,4), iBCB(nelb,2) and BCB(nelb,3,ndBCB) disagree")
        perm = None if permit is None else np.ascontiguousarray(permit, dtype=np.int32)
        self._bnd = (ienb, iBCB, BCB, perm)  # kept alive with the handle
        check(lib().crimson_form_set_boundary(self.h, int(ienb.shape[0]), _p(ienb), _p(iBCB), _p(BCB),
                                              int(BCB.shape[2]), int(nqb), float(influx), _p(perm),
                                              -1 if perm is None else int(perm.size - 1)),
              "crimson_form_set_boundary")

    def set_wall(self, params=None, nodetag=None, uref=None):
        """Deformable walls on the boundary's iBCB-bit-4 faces (the handle
        needs ideformwall = 1): params a WallParams (wall_params()), nodetag
        (numnp) node tags (bit 0: ring), uref(nshg,3) the reference
        displacement.  set_displacement gives the displacement each later
        formation uses."""
        params = params if params is not None else wall_params()
        tag = None if nodetag is None else np.ascontiguousarray(nodetag, dtype=np.int32)
        ur = None if uref is None else np.asfortranarray(uref, dtype=np.float64)
        if ur is not None and ur.shape != (self.nshg, 3):
            raise ValueError("uref must be (nshg, 3)")
        check(lib().crimson_form_set_wall(self.h, C.byref(params), _p(tag), _p(ur)), "crimson_form_set_wall")

    def set_displacement(self, u_dev):
        """The device displacement u(nshg,3) (a (3, nshg) C tensor, its
        Fortran layout) of the following formations; kept alive here."""
        self._disp = u_dev
        check(lib().crimson_form_set_displacement_dev(self.h, _p(u_dev)), "crimson_form_set_displacement_dev")

    def set_wall_state(self, uold_dev, u_dev):
        """The displacement at t_n and t_{n+1} a captured timestep
        (step_graph) updates; kept alive here."""
        self._wstate = (uold_dev, u_dev)
        check(lib().crimson_form_set_wall_state_dev(self.h, _p(uold_dev), _p(u_dev)),
              "crimson_form_set_wall_state_dev")

    def wall_predict(self, uold, yold, acold, u):
        check(lib().crimson_wall_predict_dev(self.h, _p(uold), _p(yold), _p(acold), _p(u)),
              "crimson_wall_predict_dev")

    def wall_correct(self, u, res4):
        check(lib().crimson_wall_correct_dev(self.h, _p(u), _p(res4)), "crimson_wall_correct_dev")

    def wall_alpha(self, uold, u, ualpha):
        check(lib().crimson_wall_alpha_dev(self.h, _p(uold), _p(u), _p(ualpha)), "crimson_wall_alpha_dev")

    # row schedule (numbering, Morton, Morton-ordered runs of 64 rows) / FP64 incidence chunk /
    # XCD dealing (one range per XCD, or turns of 64 workgroups)
    AUTOTUNE_CANDIDATES = tuple(f"{s}/{c}/{d}" for d in ("range", "turns64") for c in (16, 8)
                                for s in ("natural", "morton", "tiles"))

    def autotune(self, y, ac, x, res4, lhs_handle, reps=3, only=None):
        """Time each launch configuration of to_lhs on these device arrays
        and keep the fastest (row schedule x assembly chunk); returns
        (chosen name, {name: mean assembly ms}).  ``only`` restricts the
        candidates to these names.  Results do not depend on the choice."""
        mask = 0 if only is None else sum(1 << self.AUTOTUNE_CANDIDATES.index(n) for n in only)
        choice = C.c_int()
        ms = (C.c_double * 12)(*([float("nan")] * 12))
        check(lib().crimson_form_autotune(self.h, _p(y), _p(ac), _p(x), _p(res4), lhs_handle.h, int(reps),
                                          int(mask), C.byref(choice), ms), "crimson_form_autotune")
        times = {n: ms[i] for i, n in enumerate(self.AUTOTUNE_CANDIDATES) if ms[i] == ms[i]}
        return self.AUTOTUNE_CANDIDATES[choice.value], times

    def boundary_flux(self, nflx):
        """flxID(5, 0..nflx) of the last formation: area, outflow and force
        on the fluid by srfID (e3b.f90:216-247)."""
        out = np.zeros((5, nflx + 1), order="F")
        check(lib().crimson_form_boundary_flux(self.h, int(nflx), _p(out)), "crimson_form_boundary_flux")
        return out

    def boundary_flux_noalloc(self, nflx):
        """boundary_flux through the entry point that allocates no device
        memory (|srfID| below FLUX_NOALLOC_MAX_SURFACES)."""
        out = np.zeros((5, nflx + 1), order="F")
        check(lib().crimson_form_boundary_flux_noalloc(self.h, int(nflx), _p(out)),
              "crimson_form_boundary_flux_noalloc")
        return out

    def set_bc(self, BC):
        """BC(nshg, ndofBC) of the handle's iBC (ndofBC >= 6 for 1- and
        2-component velocity codes): bc3lhs on every later tangent, bc3res's
        projections; None removes it."""
        if BC is None:
            self._bc = None
            check(lib().crimson_form_set_bc(self.h, None, 0), "crimson_form_set_bc")
            return
        bc = np.asfortranarray(BC, dtype=np.float64)
        self._bc = bc
        check(lib().crimson_form_set_bc(self.h, _p(bc), int(bc.shape[1])), "crimson_form_set_bc")

    def graph_to_lhs(self, y, ac, x, res4, lhs_handle):
        """Capture to_lhs on these device arrays into a HIP graph; returns a
        FormGraph whose launch() replays it."""
        g = C.c_void_p()
        check(lib().crimson_form_graph_create(self.h, _p(y), _p(ac), _p(x), _p(res4), lhs_handle.h, C.byref(g)),
              "crimson_form_graph_create")
        return FormGraph(g, keep=(y, ac, x, res4, lhs_handle))

    def step_graph(self, lhs_handle, ls, yold, acold, y, ac, ya, aca, x, res4, ibc_dev, bc_dev, surfaces=(),
                   faceres=None, nonlinear=3):
        """Capture one generalized-alpha timestep (crimson_step_graph_create):
        itrPredict, itrBC, then `nonlinear` times itrYAlpha, the formation
        into lhs, bc3Res, ElmpvsQ of each coupled surface, the device-driven
        NS solve, itrCorrect and itrBC, and yold, acold <- y, ac, all on
        these device arrays.  surfaces: [(glob_dev int32 0-based, nabi_dev
        (nf,3), c1, c2)] of the coupled surfaces; faceres: memLS's face
        coefficients (res of memLS_SOLVE).  Each launch() advances one
        timestep; the BC array may be rewritten between launches."""
        import torch
        dev = y.device
        sizes = [int(s[0].numel()) for s in surfaces]
        surf_ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        if surfaces:
            glob = torch.cat([s[0] for s in surfaces]).to(torch.int32).contiguous()
            nabi = torch.cat([s[1] for s in surfaces]).to(torch.float64).contiguous()
        else:
            glob = torch.zeros(1, dtype=torch.int32, device=dev)
            nabi = torch.zeros((1, 3), dtype=torch.float64, device=dev)
        c1 = np.array([float(s[2]) for s in surfaces] or [0.0])
        c2 = np.array([float(s[3]) for s in surfaces] or [0.0])
        q = torch.zeros(max(1, len(surfaces)), dtype=torch.float64, device=dev)
        fr = None if faceres is None else _f(faceres)
        g = C.c_void_p()
        check(lib().crimson_step_graph_create(self.h, lhs_handle.h, ls.h, _p(yold), _p(acold), _p(y), _p(ac), _p(ya),
                                              _p(aca), _p(x), _p(res4), _p(ibc_dev), _p(bc_dev),
                                              int(bc_dev.shape[0]), len(surfaces), _p(surf_ptr), _p(glob),
                                              _p(nabi), _p(c1), _p(c2), _p(q), _p(fr), int(nonlinear), C.byref(g)),
              "crimson_step_graph_create")
        return FormGraph(g, keep=(yold, acold, y, ac, ya, aca, x, res4, ibc_dev, bc_dev, glob, nabi, c1, c2, q, fr,
                                  lhs_handle, ls, surf_ptr))

    def bc3res(self, res, res4_layout=False):
        check(lib().crimson_form_bc3res_dev(self.h, _p(res), 1 if res4_layout else 0),
              "crimson_form_bc3res_dev")

    # steps either side of formation and solve (itrPC.f90, itrbc.f90, ElmpvsQ)
    def predict(self, yold, acold, y, ac):
        check(lib().crimson_itr_predict_dev(self.h, _p(yold), _p(acold), _p(y), _p(ac), self.params.gami),
              "itrPredict")

    def yalpha(self, yold, acold, y, ac, ya, aca):
        check(lib().crimson_itr_yalpha_dev(self.h, _p(yold), _p(acold), _p(y), _p(ac), _p(ya), _p(aca),
                                           self.params.alfi, self.params.almi), "itrYAlpha")

    def correct(self, y, ac, res4):
        p = self.params
        check(lib().crimson_itr_correct_dev(self.h, _p(y), _p(ac), _p(res4), p.gami, p.alfi, p.Delt),
              "itrCorrect")

    def itr_bc(self, ibc_dev, bc_dev, y):
        check(lib().crimson_itr_bc_dev(self.h, _p(ibc_dev), _p(bc_dev), int(bc_dev.shape[0]), _p(y)), "itrBC")

    def elmpvsq(self, glob_dev, nabi_dev, R, y, res4, q_dev):
        check(lib().crimson_elmpvsq_dev(self.h, int(glob_dev.numel()), _p(glob_dev), _p(nabi_dev), R, _p(y),
                                        _p(res4), _p(q_dev)), "ElmpvsQ")

    def elmpvsq_coef(self, glob_dev, nabi_dev, c1, c2, y, res4, q_dev):
        """A coupled surface with its 0D model's implicit coefficients
        (P = c1 Q + c2; crimsonflowsolver_amd.zerod)."""
        check(lib().crimson_elmpvsq_coef_dev(self.h, int(glob_dev.numel()), _p(glob_dev), _p(nabi_dev), float(c1),
                                             float(c2), _p(y), _p(res4), _p(q_dev)), "ElmpvsQ")

    def sync(self):
        check(lib().crimson_form_sync(self.h), "crimson_form_sync")

    def timing(self):
        out = (C.c_double * 4)()
        check(lib().crimson_form_timing(self.h, out), "crimson_form_timing")
        return dict(flux_ms=out[0], e3_ms=out[1], assembly_ms=out[2], total_ms=out[3])

    def close(self):
        if getattr(self, "h", None):
            lib().crimson_form_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class memLS_lsType:
    """Opaque memLS_lsType (memLS.f90:190-202) with the public stats fields."""

    def __init__(self, h, LS_type):
        self.h = h
        self.LS_type = LS_type

    def stats(self, sub="RI"):
        out = (C.c_double * 8)()
        check(lib().crimson_ls_stats(self.h, {"RI": 0, "GM": 1, "CG": 2}[sub], out), "crimson_ls_stats")
        keys = ("iNorm", "fNorm", "itr", "suc", "dB", "callD", "Resm", "Resc")
        return dict(zip(keys, list(out)))

    @property
    def RI(self):
        return self.stats("RI")

    @property
    def GM(self):
        return self.stats("GM")

    @property
    def CG(self):
        return self.stats("CG")

    def __del__(self):
        try:
            lib().crimson_ls_free(self.h)
        except Exception:
            pass


class memLS_lhsType:
    """Opaque memLS_lhsType (memLS.f90:145-160)."""

    def __init__(self, h, nNo, nnz, nFaces):
        self.h = h
        self.nNo, self.nnz, self.nFaces = nNo, nnz, nFaces
        self._keep = []

    def spmv_dev(self, dof, x, y, reps=1):
        ms = C.c_double()
        check(lib().crimson_lhs_spmv_dev(self.h, dof, _p(x), _p(y), reps, C.byref(ms)), "crimson_lhs_spmv_dev")
        return ms.value

    def set_val(self, dof, Val):
        Val = _f(Val)
        check(lib().crimson_lhs_set_val(self.h, dof, _p(Val)), "crimson_lhs_set_val")

    def set_precision(self, bits):
        """32: hold the matrix in FP32 and solve in FP32 (mixed form); 64: back to FP64."""
        check(lib().crimson_lhs_set_precision(self.h, int(bits)), "crimson_lhs_set_precision")

    def transpose_from(self, src):
        """Hold the block transpose of src's matrix (same CSR graph)."""
        check(lib().crimson_lhs_transpose_dev(self.h, src.h), "crimson_lhs_transpose_dev")

    def get_val(self, dof=4):
        out = np.zeros((dof * dof, self.nnz), order="F")
        check(lib().crimson_lhs_get_val(self.h, dof, _p(out)), "crimson_lhs_get_val")
        return out

    def __del__(self):
        try:
            lib().crimson_lhs_free(self.h)
        except Exception:
            pass


def memLS_LS_CREATE(LS_type, relTol=None, absTol=None, maxItr=None, dimKry=None, relTolIn=None,
                    absTolIn=None, maxItrIn=None):
    present = 0
    for bit, v in ((1, relTol), (2, absTol), (4, maxItr), (8, dimKry), (16, relTolIn), (32, absTolIn),
                   (64, maxItrIn)):
        if v is not None:
            present |= bit
    rti = None if relTolIn is None else np.asarray(relTolIn, np.float64)
    ati = None if absTolIn is None else np.asarray(absTolIn, np.float64)
    mti = None if maxItrIn is None else np.asarray(maxItrIn, np.int32)
    h = C.c_void_p()
    check(lib().crimson_ls_create(LS_type, present, relTol or 0.0, absTol or 0.0, maxItr or 0, dimKry or 0,
                                  _p(rti), _p(ati), _p(mti), C.byref(h)), "memLS_LS_CREATE")
    return memLS_lsType(h, LS_type)


def memLS_LHS_CREATE(gnNo, nNo, nnz, gNodes, rowPtr, colPtr, nFaces):
    rowPtr, colPtr = _f(rowPtr, np.int32), _f(colPtr, np.int32)
    g = None if gNodes is None else _f(gNodes, np.int32)
    h = C.c_void_p()
    check(lib().crimson_lhs_create(gnNo, nNo, nnz, _p(g), _p(rowPtr), _p(colPtr), nFaces, C.byref(h)),
          "memLS_LHS_CREATE")
    return memLS_lhsType(h, nNo, nnz, nFaces)


def memLS_BC_CREATE(lhs, faIn, nNo, dof, BC_type, gNodes, Val=None):
    g = _f(gNodes, np.int32)
    v = None if Val is None else _f(Val)
    check(lib().crimson_bc_create(lhs.h, faIn, nNo, dof, BC_type, _p(g), _p(v)), "memLS_BC_CREATE")


def memLS_SOLVE(lhs, ls, dof, Ri, Val, incL=None, res=None):
    """Ri (dof,nNo) Fortran-ordered float64 is overwritten with the solution."""
    if not (isinstance(Ri, np.ndarray) and Ri.dtype == np.float64 and Ri.flags.f_contiguous):
        raise CrimsonError(1, "memLS_SOLVE", "Ri must be a Fortran-ordered float64 array (in/out)")
    Val = _f(Val)
    il = None if incL is None else _f(incL, np.int32)
    rs = None if res is None else _f(res)
    check(lib().crimson_ls_solve(lhs.h, ls.h, dof, _p(Ri), _p(Val), _p(il), _p(rs)), "memLS_SOLVE")


def memLS_SOLVE_graph(lhs, ls, dof, Ri_dev, incL=None, res=None):
    """The NS solve of ``lhs``'s matrix on the device array Ri_dev captured
    device-driven into a HIP graph (crimson_ls_graph_create); launch() solves
    in place and leaves the statistics in ``ls``."""
    il = None if incL is None else _f(incL, np.int32)
    rs = None if res is None else _f(res)
    g = C.c_void_p()
    check(lib().crimson_ls_graph_create(lhs.h, ls.h, dof, _p(Ri_dev), _p(il), _p(rs), C.byref(g)),
          "crimson_ls_graph_create")
    return FormGraph(g, keep=(lhs, ls, Ri_dev, il, rs))


def memLS_SOLVE_dev(lhs, ls, dof, Ri_dev, incL=None, res=None):
    """Device form: the matrix is the one held by ``lhs``; Ri_dev is a torch
    tensor (nNo, dof) C-ordered = (dof, nNo) Fortran-ordered, in/out."""
    il = None if incL is None else _f(incL, np.int32)
    rs = None if res is None else _f(res)
    check(lib().crimson_ls_solve_dev(lhs.h, ls.h, dof, _p(Ri_dev), _p(il), _p(rs)), "memLS_SOLVE_dev")

"""ctypes binding of include/homo3d.h and the reference-named Python façade.

In a process that also uses PyTorch-ROCm, import torch first: torch ships its own HIP runtime, and
the one loaded first serves the process; with libhomo3d.so's loaded first, homo3d_create fails once
torch has initialised the GPU."""
import ctypes
import os
import time

import numpy as np

HOMO3D_OK, HOMO3D_NOT_CONVERGED, HOMO3D_DIVERGED, HOMO3D_BAD_ARG, HOMO3D_HIP_ERROR, HOMO3D_NO_SUPPORT = range(6)
SOLVER_MG, SOLVER_PCG, SOLVER_NONE = 0, 1, 2
FIELD_U, FIELD_F, FIELD_R = 0, 1, 2
SYM = {"reflect3": 0, "reflect6": 1, "rotate3": 2, "none": 3}

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "libhomo3d.so")
_lib = None


class Homo3dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("homo3d status %d: %s" % (code, msg))
        self.code = code


class _Config(ctypes.Structure):
    _fields_ = [("reso", ctypes.c_int * 3), ("youngs_modulus", ctypes.c_double), ("poisson_ratio", ctypes.c_double),
                ("emin", ctypes.c_double), ("fem_rel_thres", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("solver", ctypes.c_int), ("device", ctypes.c_int), ("sor_omega", ctypes.c_double),
                ("diag_clamp", ctypes.c_double), ("precond_half", ctypes.c_int), ("coarse_half", ctypes.c_int),
                ("fine_mfma", ctypes.c_int)]


class _NetConfig(ctypes.Structure):
    _fields_ = [("grid", ctypes.c_int), ("levels", ctypes.c_int), ("base_channels", ctypes.c_int),
                ("batch", ctypes.c_int), ("target_scale", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("device", ctypes.c_int)]


class _TuneReport(ctypes.Structure):
    _fields_ = [("row_ms", ctypes.c_double), ("mfma_ms", ctypes.c_double), ("chosen_mfma", ctypes.c_int)]


class _Stats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iters", ctypes.c_int * 6), ("relres", ctypes.c_double * 6),
                ("seconds", ctypes.c_double)]


EXPORTS = {
    "homo3d_default_config": (None, [ctypes.POINTER(_Config)]),
    "homo3d_create": (ctypes.c_int, [ctypes.POINTER(_Config), ctypes.POINTER(ctypes.c_void_p)]),
    "homo3d_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "homo3d_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "homo3d_levels": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
    "homo3d_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "homo3d_update_density": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "homo3d_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.POINTER(_Stats)]),
    "homo3d_elastic_matrix": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "homo3d_residual_norms": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "homo3d_stage_tiles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "homo3d_density_pow": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_int]),
    "homo3d_density_pow_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    "homo3d_field_sum": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "homo3d_field_fill": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int]),
    "homo3d_device_alloc": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]),
    "homo3d_device_free": (ctypes.c_int, [ctypes.c_void_p]),
    "homo3d_device_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]),
    "homo3d_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_TuneReport)]),
    "homo3d_sensitivity": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                          ctypes.c_int, ctypes.c_int]),
    "homo3d_sensitivity_order": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "homo3d_field_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_int]),
    "homo3d_solution_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "homo3d_macro_rhs": (ctypes.c_int, [ctypes.c_void_p]),
    "homo3d_residual": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "homo3d_apply_K": (ctypes.c_int, [ctypes.c_void_p]),
    "homo3d_apply_K16": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "homo3d_smooth16": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]),
    "homo3d_presmooth16": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "homo3d_presmooth": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]),
    "homo3d_smooth": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    "homo3d_restrict": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "homo3d_prolong": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "homo3d_coarse_solve": (ctypes.c_int, [ctypes.c_void_p]),
    "homo3d_vcycle": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "homo3d_vcycle_sweeps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    "homo3d_get_stencil": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "homo3d_symmetrize": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int]),
    "homo3d_oc_filter": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    "homo3d_oc_update": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int]),
    "homo3d_density_filter": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "homo3d_projection": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    "homo3d_oc_update_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_void_p]),
    "homo3d_mma_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "homo3d_mma_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "homo3d_mma_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "homo3d_mma_set_bound": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "homo3d_mma_set_bounds": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "homo3d_mma_update": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    "homo3d_mma_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_int)]),
    "homo3d_randc_density": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_void_p]),
    "homo3d_net_default_config": (None, [ctypes.POINTER(_NetConfig)]),
    "homo3d_net_create": (ctypes.c_int, [ctypes.POINTER(_NetConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "homo3d_net_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "homo3d_net_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "homo3d_net_num_params": (ctypes.c_long, [ctypes.c_void_p]),
    "homo3d_net_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "homo3d_net_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "homo3d_net_loss_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]),
    "homo3d_net_train_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_double, ctypes.POINTER(ctypes.c_double)]),
    "homo3d_net_dataset_reserve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "homo3d_net_dataset_add": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "homo3d_net_dataset_size": (ctypes.c_int, [ctypes.c_void_p]),
    "homo3d_net_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_uint64, ctypes.c_void_p]),
    "homo3d_net_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "homo3d_net_warm_start": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "homo3d_batch_create": (ctypes.c_int, [ctypes.POINTER(_Config), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                           ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "homo3d_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "homo3d_batch_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "homo3d_batch_handle": (ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_int]),
    "homo3d_batch_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_double,
                                          ctypes.c_int, ctypes.POINTER(_Stats), ctypes.POINTER(ctypes.c_double)]),
    "homo3d_profile_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "homo3d_profile_only": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "homo3d_profile_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long),
                                          ctypes.POINTER(ctypes.c_double)]),
}


def lib():
    """Load libhomo3d.so (raises when it has not been built: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(lib_path):
            raise ImportError("libhomo3d.so not built: run `python -m homo3d_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(lib_path)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class DeviceBuffer:
    """A device allocation made through the HIP runtime that libhomo3d.so itself links (no PyTorch):
    for callers that keep densities or sensitivities resident in HBM and pass on_device=True."""

    _hip = None

    @classmethod
    def _rt(cls):
        if cls._hip is None:
            lib()  # loads libamdhip64 as a dependency of libhomo3d.so
            h = ctypes.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipDeviceSynchronize.argtypes = []
            for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
                f.restype = ctypes.c_int
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        rc = self._rt().hipMalloc(ctypes.byref(self.ptr), self.nbytes)
        if rc:
            raise Homo3dError(HOMO3D_HIP_ERROR, "hipMalloc(%d) failed (%d)" % (self.nbytes, rc))

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        buf.upload(a)
        return buf

    def upload(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        rc = self._rt().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1)  # host to device
        if rc:
            raise Homo3dError(HOMO3D_HIP_ERROR, "hipMemcpy H2D failed (%d)" % rc)

    def download(self, dtype=np.float32, count=None):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize if count is None else count, dtype=dtype)
        rc = self._rt().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, 2)  # device to host
        if rc:
            raise Homo3dError(HOMO3D_HIP_ERROR, "hipMemcpy D2H failed (%d)" % rc)
        return out

    @classmethod
    def synchronize(cls):
        rc = cls._rt().hipDeviceSynchronize()
        if rc:
            raise Homo3dError(HOMO3D_HIP_ERROR, "hipDeviceSynchronize failed (%d)" % rc)

    @property
    def value(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            self._rt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def randc_density(n, seed, vol, sym="reflect6"):
    out = np.zeros(int(n) ** 3, dtype=np.float32)
    rc = lib().homo3d_randc_density(int(n), int(seed), float(vol), SYM[sym] if isinstance(sym, str) else int(sym),
                                    _ptr(out))
    if rc:
        raise Homo3dError(rc, "randc_density")
    return out


class HomoConfig:
    """cfg::HomoConfig (cmdline.h:39-60) fields used by the solve path, with the reference defaults
    (cmdline.cpp:4-27: E = 10, nu = 0.3, relthres = 1e-2, reso 64)."""

    def __init__(self, reso=64, youngsModulu=10.0, poissonRatio=0.3, femRelThres=1e-2, max_iter=200,
                 solver="mg", emin=1e-9, device=0, sor_omega=1.0, diag_clamp=1e-4, precond_half=True,
                 coarse_half=True, fine_mfma=False, **kw):
        self.reso = [reso] * 3 if np.isscalar(reso) else list(reso)
        self.youngsModulu = youngsModulu
        self.poissonRatio = poissonRatio
        self.femRelThres = femRelThres
        self.max_iter = max_iter
        self.solver = solver
        self.emin = emin
        self.device = device
        self.sor_omega = sor_omega
        self.diag_clamp = diag_clamp
        self.precond_half = precond_half
        self.coarse_half = coarse_half
        self.fine_mfma = fine_mfma
        # optimisation-side fields of cfg::HomoConfig (kept for drivers)
        self.volRatio = kw.get("volRatio", 0.3)
        self.filterRadius = kw.get("filterRadius", 2.0)
        self.designStep = kw.get("designStep", 0.05)
        self.dampRatio = kw.get("dampRatio", 0.5)
        self.finthres = kw.get("finthres", 5e-4)
        self.sym = kw.get("sym", "reflect6")

    def _c(self):
        c = _Config()
        lib().homo3d_default_config(ctypes.byref(c))
        for i in range(3):
            c.reso[i] = int(self.reso[i])
        c.youngs_modulus = self.youngsModulu
        c.poisson_ratio = self.poissonRatio
        c.emin = self.emin
        c.fem_rel_thres = self.femRelThres
        c.max_iter = self.max_iter
        c.solver = {"pcg": SOLVER_PCG, "none": SOLVER_NONE}.get(self.solver, SOLVER_MG)
        c.device = self.device
        c.sor_omega = self.sor_omega
        c.diag_clamp = self.diag_clamp
        c.precond_half = 1 if self.precond_half else 0
        c.coarse_half = 1 if self.coarse_half else 0
        c.fine_mfma = 1 if self.fine_mfma else 0
        return c


class Homogenization:
    """homo::Homogenization (homogenization/homogenization.h:17-61) over the C ABI."""

    def __init__(self, config):
        self.config = config
        self._h = ctypes.c_void_p()
        c = config._c()
        rc = lib().homo3d_create(ctypes.byref(c), ctypes.byref(self._h))
        if rc:
            raise Homo3dError(rc, "homo3d_create(reso=%r)" % (config.reso,))
        buf = (ctypes.c_int * 16)()
        nl = lib().homo3d_levels(self._h, buf, 16)
        self.sizes = [buf[i] for i in range(nl)]
        self.n = self.sizes[0]
        self.N = self.n ** 3
        self.last_stats = None

    def close(self):
        if self._h:
            lib().homo3d_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what, allow=()):
        if rc and rc not in allow:
            raise Homo3dError(rc, "%s: %s" % (what, lib().homo3d_last_error(self._h).decode()))
        return rc

    @property
    def handle(self):
        return self._h

    @property
    def stream(self):
        return lib().homo3d_stream(self._h)

    # --- Homogenization API ---
    def update(self, rho, pitchT=-1, on_device=False):
        """Homogenization::update(float* rho, int pitchT): rho holds x = rho^p per cell."""
        if on_device:
            ptr, pitch = rho, pitchT
        else:
            a = np.ascontiguousarray(rho, dtype=np.float32).ravel()
            pitch = self.n if pitchT < 0 else pitchT
            if a.size < pitch * self.n * self.n:
                raise ValueError("density array too small for pitch %d" % pitch)
            self._rho = a
            ptr = _ptr(a)
        self._chk(lib().homo3d_update_density(self._h, ptr, int(pitch), int(bool(on_device))), "update")

    def solve(self, tol=None, warm=True):
        st = _Stats()
        rc = lib().homo3d_solve(self._h, float(tol if tol is not None else self.config.femRelThres), int(warm),
                                ctypes.byref(st))
        self._chk(rc, "solve", allow=(HOMO3D_NOT_CONVERGED,))
        self.last_stats = {"status": st.status, "iters": list(st.iters), "relres": list(st.relres),
                           "seconds": st.seconds}
        return self.last_stats

    def elasticMatrix(self, solve=True, tol=None, warm=True):
        """Homogenization::elasticMatrix(double C[6][6]): solves the 6 load cases, returns C^H."""
        if solve:
            self.solve(tol, warm)
        C = (ctypes.c_double * 36)()
        self._chk(lib().homo3d_elastic_matrix(self._h, C), "elasticMatrix")
        return np.array(C[:]).reshape(6, 6)

    def autotune(self):
        """Time the finest-level operator's tile forms on this grid and density; keep the faster."""
        r = _TuneReport()
        self._chk(lib().homo3d_autotune(self._h, ctypes.byref(r)), "autotune")
        return {"row_ms": r.row_ms, "mfma_ms": r.mfma_ms, "chosen": "mfma" if r.chosen_mfma else "row"}

    def stage_tiles(self, on=True):
        """CG pass: stage each workgroup's z node box in LDS (True, the default) or gather from memory."""
        self._chk(lib().homo3d_stage_tiles(self._h, int(bool(on))), "stage_tiles")

    def residual_norms(self):
        """FP64 relres per load case of the current FP64 iterate (no solve)."""
        r = (ctypes.c_double * 6)()
        self._chk(lib().homo3d_residual_norms(self._h, r), "residual_norms")
        return list(r)

    def Sensitivity(self, dC, pitchT=-1, lexiOrder=True):
        """Homogenization::Sensitivity(dC, sens, pitchT, lexiOrder): returns dJ/dx per cell, lexicographic
        and pitched, or (lexiOrder=False) in the library's parity-blocked cell order."""
        d = np.ascontiguousarray(dC, dtype=np.float32).ravel()
        pitch = self.n if pitchT < 0 else pitchT
        out = np.zeros((pitch if lexiOrder else self.n) * self.n * self.n, dtype=np.float32)
        self._chk(lib().homo3d_sensitivity_order(self._h, d.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                 _ptr(out), int(pitch), int(bool(lexiOrder)), 0), "Sensitivity")
        return out

    # --- Grid / MG level operations (all 6 load cases) ---
    def level_shape(self, level):
        return (6, 3, self.sizes[level] ** 3)

    def get_field(self, level, field):
        out = np.zeros(self.level_shape(level), dtype=np.float32)
        self._chk(lib().homo3d_field_copy(self._h, level, field, _ptr(out), 0), "field_copy")
        return out

    def set_field(self, level, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(self.level_shape(level))
        self._chk(lib().homo3d_field_copy(self._h, level, field, _ptr(a), 1), "field_copy")

    def get_solution64(self):
        out = np.zeros(self.level_shape(0), dtype=np.float64)
        self._chk(lib().homo3d_solution_copy(self._h, _ptr(out), 0), "solution_copy")
        return out

    def set_solution64(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.level_shape(0))
        self._chk(lib().homo3d_solution_copy(self._h, _ptr(a), 1), "solution_copy")

    def macro_rhs(self):
        self._chk(lib().homo3d_macro_rhs(self._h), "macro_rhs")

    def residual(self, level=0):
        self._chk(lib().homo3d_residual(self._h, level), "residual")

    def apply_K(self):
        self._chk(lib().homo3d_apply_K(self._h), "apply_K")

    def apply_K16(self, residual=False):
        """r[0] = K fp16(u[0]) (or f - K fp16(u[0])) with the packed-FP16 finest-level operator."""
        self._chk(lib().homo3d_apply_K16(self._h, int(residual)), "apply_K16")

    def smooth16(self, omega=1.0, reverse=False):
        """One finest-level sweep on fp16(u[0]) with FP16 storage; u[0] receives the result."""
        self._chk(lib().homo3d_smooth16(self._h, float(omega), int(reverse)), "smooth16")

    def presmooth16(self, omega=1.0):
        """The preconditioner's level-0 pre-smoothing: a forward sweep from fp16 u = 0 with f[0]; u[0]
        receives the result and r[0] the residual f - K u, formed from the sweep's own-colour residuals
        and the later-colour terms."""
        self._chk(lib().homo3d_presmooth16(self._h, float(omega)), "presmooth16")

    def presmooth(self, level, omega=1.0):
        """A coarse level's pre-smoothing: a forward sweep from u = 0 with f[level]; u[level] receives the
        result and r[level] the residual, formed from the own-colour residuals and the later-colour terms."""
        self._chk(lib().homo3d_presmooth(self._h, int(level), float(omega)), "presmooth")

    def smooth(self, level=0, omega=1.0, reverse=False):
        self._chk(lib().homo3d_smooth(self._h, level, float(omega), int(reverse)), "smooth")

    def restrict(self, level):
        self._chk(lib().homo3d_restrict(self._h, level), "restrict")

    def prolong(self, level):
        self._chk(lib().homo3d_prolong(self._h, level), "prolong")

    def coarse_solve(self):
        self._chk(lib().homo3d_coarse_solve(self._h), "coarse_solve")

    def vcycle(self, symmetric=False):
        self._chk(lib().homo3d_vcycle(self._h, int(symmetric)), "vcycle")

    def v_cycle(self, w_SOR=1.0, pre=1, post=1):
        """MG::v_cycle(w_SOR, pre, post) on the FP32 level-0 fields."""
        self._chk(lib().homo3d_vcycle_sweeps(self._h, float(w_SOR), int(pre), int(post)), "v_cycle")

    def stencil(self, level):
        out = np.zeros((self.sizes[level] ** 3, 27, 3, 3), dtype=np.float32)
        self._chk(lib().homo3d_get_stencil(self._h, level, _ptr(out)), "get_stencil")
        return out

    # --- density utilities on this grid ---
    def symmetrize(self, field, sym):
        a = np.ascontiguousarray(field, dtype=np.float32).ravel().copy()
        self._chk(lib().homo3d_symmetrize(self._h, _ptr(a), self.n, SYM[sym] if isinstance(sym, str) else int(sym),
                                          0), "symmetrize")
        return a

    def density_filter(self, field, radius=2.0, kernel="spline4", periodic=False, backward=False):
        """conv with radial_convker_t<Spline4|Linear>(radius) on an n^3 field (or its adjoint)."""
        a = np.ascontiguousarray(field, dtype=np.float32).ravel()
        out = np.zeros_like(a)
        self._chk(lib().homo3d_density_filter(self._h, _ptr(a), _ptr(out), self.n, float(radius),
                                              0 if kernel == "spline4" else 1, int(periodic), int(backward), 0),
                  "density_filter")
        return out

    def projection(self, field, k=30.0, c=0.5, y=None):
        """sgm(k, c) of an n^3 field; with y (the forward output) the backward pass of dJ/dy = field."""
        a = np.ascontiguousarray(field, dtype=np.float32).ravel()
        out = np.zeros_like(a)
        yy = None if y is None else np.ascontiguousarray(y, dtype=np.float32).ravel()
        self._chk(lib().homo3d_projection(self._h, _ptr(a), None if yy is None else _ptr(yy), _ptr(out), self.n,
                                          float(k), float(c), int(y is not None), 0), "projection")
        return out

    def profile_enable(self, on=True, only=None):
        """Event timing of the finest-level kernels; `only` restricts it to one kind (profile_get names)."""
        self._chk(lib().homo3d_profile_only(self._h, (only or "").encode()), "profile_only")
        self._chk(lib().homo3d_profile_enable(self._h, int(on)), "profile_enable")

    def profile_get(self, kernel):
        cnt = ctypes.c_long()
        ms = ctypes.c_double()
        self._chk(lib().homo3d_profile_get(self._h, kernel.encode(), ctypes.byref(cnt), ctypes.byref(ms)),
                  "profile_get")
        return cnt.value, ms.value


class OCOptimizer:
    """homo::OCOptimizer (optimization/OCOptimizer.h:7-20): filterSens and update on the GPU."""

    def __init__(self, hom, min_density=0.001, stepLimit=0.05, dampExponent=0.5):
        self.hom = hom
        self.minRho = min_density
        self.step_limit = stepLimit
        self.damp = dampExponent

    def filterSens(self, sens, rho, radius=2.0, periodic=False):
        s = np.ascontiguousarray(sens, dtype=np.float32).ravel().copy()
        r = np.ascontiguousarray(rho, dtype=np.float32).ravel()
        self.hom._chk(lib().homo3d_oc_filter(self.hom.handle, _ptr(s), _ptr(r), self.hom.n, float(radius),
                                             int(periodic), 0), "oc_filter")
        return s

    def update(self, sens, rho, volratio):
        s = np.ascontiguousarray(sens, dtype=np.float32).ravel()
        r = np.ascontiguousarray(rho, dtype=np.float32).ravel().copy()
        self.hom._chk(lib().homo3d_oc_update(self.hom.handle, _ptr(s), _ptr(r), self.hom.n, float(volratio),
                                             float(self.minRho), float(self.step_limit), float(self.damp), 0),
                      "oc_update")
        return r


class MMAOptimizer:
    """homo::MMAOptimizer(int m, int n, double a0, double a, double c, double d)
    (optimization/mmaOptimizer.h:12-34): Svanberg's MMA step on the GPU in FP64."""

    def __init__(self, m, n, a0=1.0, a=0.0, c=1000.0, d=1.0, device=0):
        self.m, self.n = int(m), int(n)
        self._h = ctypes.c_void_p()
        rc = lib().homo3d_mma_create(self.m, self.n, float(a0), float(a), float(c), float(d), int(device),
                                     ctypes.byref(self._h))
        if rc:
            raise Homo3dError(rc, "homo3d_mma_create")

    def close(self):
        if self._h:
            lib().homo3d_mma_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            raise Homo3dError(rc, "%s: %s" % (what, lib().homo3d_mma_last_error(self._h).decode()))

    def setBound(self, xmin, xmax):
        """setBound(float xmin, float xmax) or, with arrays, setBound(float* xmin, float* xmax)."""
        if np.isscalar(xmin) and np.isscalar(xmax):
            self._chk(lib().homo3d_mma_set_bound(self._h, float(xmin), float(xmax)), "setBound")
            return
        lo = np.ascontiguousarray(np.broadcast_to(xmin, (self.n,)), dtype=np.float32)
        hi = np.ascontiguousarray(np.broadcast_to(xmax, (self.n,)), dtype=np.float32)
        self._chk(lib().homo3d_mma_set_bounds(self._h, _ptr(lo), _ptr(hi), 0), "setBound")

    def update(self, itn, x, dfdx, gval, dgdx):
        """One MMA step; returns the new design (float32).  dgdx: sequence of m gradient arrays."""
        xv = np.ascontiguousarray(x, dtype=np.float32).ravel().copy()
        df = np.ascontiguousarray(dfdx, dtype=np.float32).ravel()
        gv = np.ascontiguousarray(gval, dtype=np.float32).ravel()
        dg = [np.ascontiguousarray(g, dtype=np.float32).ravel() for g in dgdx]
        ptrs = (ctypes.c_void_p * self.m)(*[g.ctypes.data for g in dg])
        self._chk(lib().homo3d_mma_update(self._h, int(itn), _ptr(xv), _ptr(df), _ptr(gv), ptrs, 0), "update")
        return xv

    def state(self):
        x = np.zeros(self.n)
        lo = np.zeros(self.n)
        up = np.zeros(self.n)
        it = ctypes.c_int()
        self._chk(lib().homo3d_mma_state(self._h, _ptr(x), _ptr(lo), _ptr(up), ctypes.byref(it)), "state")
        return x, lo, up, it.value


class WarmStartNet:
    """The learned warm start (homo3d_net_*): a periodic 3D U-Net trained on the GPU on (density,
    solution) pairs the solver generates; its prediction seeds the refined MG-PCG solve."""

    def __init__(self, grid=64, levels=4, base_channels=16, batch=4, target_scale=8.0, seed=0, device=0):
        c = _NetConfig()
        lib().homo3d_net_default_config(ctypes.byref(c))
        c.grid, c.levels, c.base_channels, c.batch = int(grid), int(levels), int(base_channels), int(batch)
        c.target_scale, c.seed, c.device = float(target_scale), int(seed), int(device)
        self._h = ctypes.c_void_p()
        rc = lib().homo3d_net_create(ctypes.byref(c), ctypes.byref(self._h))
        if rc:
            raise Homo3dError(rc, "homo3d_net_create(grid=%d, levels=%d)" % (grid, levels))
        self.grid, self.levels, self.base_channels, self.batch = int(grid), int(levels), int(base_channels), int(batch)
        self.target_scale = float(target_scale)
        self.num_params = int(lib().homo3d_net_num_params(self._h))

    def close(self):
        if self._h:
            lib().homo3d_net_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            raise Homo3dError(rc, "%s: %s" % (what, lib().homo3d_net_last_error(self._h).decode()))

    @property
    def handle(self):
        return self._h

    def get_params(self):
        out = np.zeros(self.num_params, np.float32)
        self._chk(lib().homo3d_net_params(self._h, _ptr(out), 0), "params")
        return out

    def set_params(self, theta):
        a = np.ascontiguousarray(theta, dtype=np.float32).ravel()
        if a.size != self.num_params:
            raise ValueError("expected %d parameters, got %d" % (self.num_params, a.size))
        self._chk(lib().homo3d_net_params(self._h, _ptr(a), 1), "params")

    def forward(self, rho):
        r = np.ascontiguousarray(rho, dtype=np.float32).reshape(-1, self.grid ** 3)
        out = np.zeros((r.shape[0], self.grid ** 3, 18), np.float32)
        self._chk(lib().homo3d_net_forward(self._h, _ptr(r), r.shape[0], _ptr(out)), "forward")
        return out

    def loss_grad(self, rho, target):
        r = np.ascontiguousarray(rho, dtype=np.float32).reshape(-1, self.grid ** 3)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(r.shape[0], self.grid ** 3, 18)
        g = np.zeros(self.num_params, np.float32)
        loss = ctypes.c_double()
        self._chk(lib().homo3d_net_loss_grad(self._h, _ptr(r), _ptr(t), r.shape[0], ctypes.byref(loss), _ptr(g)),
                  "loss_grad")
        return loss.value, g

    def train_step(self, rho, target, lr):
        r = np.ascontiguousarray(rho, dtype=np.float32).reshape(-1, self.grid ** 3)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(r.shape[0], self.grid ** 3, 18)
        loss = ctypes.c_double()
        self._chk(lib().homo3d_net_train_step(self._h, _ptr(r), _ptr(t), r.shape[0], float(lr), ctypes.byref(loss)),
                  "train_step")
        return loss.value

    def dataset_reserve(self, count):
        self._chk(lib().homo3d_net_dataset_reserve(self._h, int(count)), "dataset_reserve")

    def dataset_add(self, hom):
        """Append the (density, solution) pair of a solved Homogenization."""
        self._chk(lib().homo3d_net_dataset_add(self._h, hom.handle), "dataset_add")

    def dataset_size(self):
        return int(lib().homo3d_net_dataset_size(self._h))

    def train(self, epochs, lr, seed=0, count=0):
        """Adam epochs over the first `count` training pairs (all when 0); returns the epoch losses."""
        losses = np.zeros(max(int(epochs), 1))
        self._chk(lib().homo3d_net_train(self._h, int(epochs), int(count), float(lr), int(seed), _ptr(losses)),
                  "train")
        return losses[:int(epochs)]

    def eval(self, first, count):
        loss = ctypes.c_double()
        self._chk(lib().homo3d_net_eval(self._h, int(first), int(count), ctypes.byref(loss)), "eval")
        return loss.value

    def warm_start(self, hom):
        """Predict hom's solution from its current density; the next hom.solve(warm=True) starts there."""
        rc = lib().homo3d_net_warm_start(hom.handle, self._h)
        if rc:
            raise Homo3dError(rc, "net_warm_start: %s" % lib().homo3d_last_error(hom.handle).decode())

    def save(self, path):
        """Parameters and configuration as an .npz (no pickled objects)."""
        np.savez(path, params=self.get_params(), grid=self.grid, levels=self.levels,
                 base_channels=self.base_channels, target_scale=self.target_scale)

    @classmethod
    def load(cls, path, batch=1, device=0):
        d = np.load(path, allow_pickle=False)
        net = cls(grid=int(d["grid"]), levels=int(d["levels"]), base_channels=int(d["base_channels"]), batch=batch,
                  target_scale=float(d["target_scale"]), device=device)
        net.set_params(d["params"])
        return net


class Batch:
    """Batched uneven designs (homo3d_batch_*): one handle per design, solved concurrently."""

    def __init__(self, sizes, config=None, max_inflight=8):
        cfg = (config or HomoConfig(solver="pcg"))._c()
        self.sizes = [int(n) for n in sizes]
        arr = (ctypes.c_int * len(self.sizes))(*self.sizes)
        self._h = ctypes.c_void_p()
        rc = lib().homo3d_batch_create(ctypes.byref(cfg), len(self.sizes), arr, int(max_inflight), ctypes.byref(self._h))
        if rc:
            raise Homo3dError(rc, "homo3d_batch_create")

    def close(self):
        if self._h:
            lib().homo3d_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, xs, tol=1e-5, autotune=False):
        """xs: per-design arrays x = rho^3 (n_i^3 floats).  Returns (C [count, 6, 6], list of stats dicts)."""
        arrs = [np.ascontiguousarray(x, dtype=np.float32).ravel() for x in xs]
        if len(arrs) != len(self.sizes) or any(a.size != n ** 3 for a, n in zip(arrs, self.sizes)):
            raise ValueError("one n^3 density per design expected")
        ptrs = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        st = (_Stats * len(arrs))()
        C = np.zeros((len(arrs), 36))
        rc = lib().homo3d_batch_solve(self._h, ptrs, float(tol), int(bool(autotune)), st,
                                      C.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc and rc != HOMO3D_NOT_CONVERGED:
            raise Homo3dError(rc, "batch_solve: %s" % lib().homo3d_batch_last_error(self._h).decode())
        stats = [{"status": s.status, "iters": list(s.iters), "relres": list(s.relres), "seconds": s.seconds}
                 for s in st]
        return C.reshape(-1, 6, 6), stats

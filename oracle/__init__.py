"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

The oracle restates lavenklau/homo3d's periodic homogenization solve path in FP64 C++
(homo_oracle*.cpp, each function citing the reference file:line it follows).  Only
tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import this
module; the product package ``homo3d_amd`` never does.

Vectors of one load case are numpy float64 arrays of shape (3, N) (component-major,
N = n^3 vertices in lexicographic order x + n*(y + n*z)).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip3 = ctypes.c_int * 3


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_double, c_long, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_long, ctypes.c_void_p
        sig = {
            "orc_ke": (None, [c_double, c_double, _dp]),
            "orc_ke_lame": (None, [c_double, c_double, _dp]),
            "orc_level_sizes": (c_int, [c_int, ctypes.POINTER(c_int), c_int]),
            "orc_macro_rhs": (None, [_ip3, _dp, _dp, c_int, _dp]),
            "orc_apply_K": (None, [_ip3, _dp, _dp, _dp, _dp, c_int]),
            "orc_residual": (None, [_ip3, _dp, _dp, _dp, _dp, _dp]),
            "orc_gs_fine": (None, [_ip3, _dp, _dp, _dp, _dp, c_double, c_int]),
            "orc_stencil_apply": (None, [_ip3, _dp, _dp, _dp]),
            "orc_stencil_residual": (None, [_ip3, _dp, _dp, _dp, _dp]),
            "orc_stencil_gs": (None, [_ip3, _dp, _dp, _dp, c_double, c_int]),
            "orc_restrict": (None, [_ip3, _dp, _dp]),
            "orc_prolong_add": (None, [_ip3, _dp, _dp, c_int]),
            "orc_galerkin_fine": (None, [_ip3, _dp, _dp, _dp, c_double]),
            "orc_galerkin_coarse": (None, [_ip3, _dp, _dp, c_double]),
            "orc_dense_from_stencil": (None, [_ip3, _dp, _dp]),
            "orc_dense_fine": (None, [_ip3, _dp, _dp, _dp]),
            "orc_cholesky": (c_int, [c_int, _dp]),
            "orc_chol_solve": (None, [c_int, _dp, _dp]),
            "orc_elastic_matrix": (None, [_ip3, _dp, _dp, _dp, _dp]),
            "orc_sensitivity": (None, [_ip3, _dp, _dp, _dp, _dp]),
            "orc_mg_create": (c_void_p, [_ip3, c_double, c_double, c_double]),
            "orc_mg_destroy": (None, [c_void_p]),
            "orc_mg_levels": (c_int, [c_void_p, ctypes.POINTER(c_int)]),
            "orc_mg_set_density": (c_int, [c_void_p, _dp]),
            "orc_mg_get_stencil": (None, [c_void_p, c_int, _dp]),
            "orc_mg_vcycle": (c_double, [c_void_p, _dp, _dp, c_double, c_int]),
            "orc_mg_solve_mg": (c_int, [c_void_p, _dp, _dp, c_double, c_int, c_double,
                                        ctypes.POINTER(c_double), ctypes.POINTER(c_int)]),
            "orc_mg_solve_pcg": (c_int, [c_void_p, _dp, _dp, c_double, c_int, c_double,
                                         ctypes.POINTER(c_double), ctypes.POINTER(c_int)]),
            "orc_mg_homogenize": (c_int, [c_void_p, c_double, c_int, c_int, c_int, _dp, _dp,
                                          ctypes.POINTER(c_int)]),
            "orc_philox4x32_10": (None, [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                         ctypes.POINTER(ctypes.c_uint32)]),
            "orc_symmetrize_f32": (None, [c_int, c_int, _fp, _fp]),
            "orc_symmetrize_f64": (None, [c_int, c_int, _dp, _dp]),
            "orc_oc_filter": (None, [_ip3, _dp, _dp, c_double, c_int, _dp]),
            "orc_oc_update": (c_int, [c_long, _fp, _fp, c_double, c_double, c_double, c_double]),
            "orc_mma_update": (c_int, [c_int, c_long, c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                       _dp, _dp, c_double, _dp, _dp, _dp, c_double]),
            "orc_randc_density": (None, [c_int, ctypes.c_uint64, c_double, c_int, _fp]),
            "orc_randc_coeffs": (None, [ctypes.c_uint64, _dp]),
            "orc_randc_field": (None, [c_int, _dp, c_double, c_int, _fp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _n3(n):
    if np.isscalar(n):
        n = (n, n, n)
    return _ip3(*[int(v) for v in n]), int(n[0]) * int(n[1]) * int(n[2])


def ke(E=10.0, nu=0.3):
    out = np.zeros(576)
    lib().orc_ke(E, nu, out)
    return out.reshape(24, 24)


def ke_lame(lam, mu):
    out = np.zeros(576)
    lib().orc_ke_lame(lam, mu, out)
    return out.reshape(24, 24)


def level_sizes(n):
    buf = (ctypes.c_int * 16)()
    nl = lib().orc_level_sizes(int(n), buf, 16)
    if nl < 0:
        raise ValueError("unsupported resolution %r" % n)
    return [buf[i] for i in range(nl)]


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def macro_rhs(n, KE, w, s):
    n3, N = _n3(n)
    f = np.zeros((3, N))
    lib().orc_macro_rhs(n3, _c(KE).ravel(), _c(w).ravel(), int(s), f)
    return f


def apply_K(n, KE, w, u, pin=True):
    n3, N = _n3(n)
    y = np.zeros((3, N))
    lib().orc_apply_K(n3, _c(KE).ravel(), _c(w).ravel(), _c(u), y, int(pin))
    return y


def residual(n, KE, w, u, f):
    n3, N = _n3(n)
    r = np.zeros((3, N))
    lib().orc_residual(n3, _c(KE).ravel(), _c(w).ravel(), _c(u), _c(f), r)
    return r


def gs_fine(n, KE, w, u, f, omega=1.0, reverse=False):
    n3, N = _n3(n)
    u = _c(u).copy()
    lib().orc_gs_fine(n3, _c(KE).ravel(), _c(w).ravel(), u, _c(f), float(omega), int(reverse))
    return u


def stencil_apply(n, st, u):
    n3, N = _n3(n)
    y = np.zeros((3, N))
    lib().orc_stencil_apply(n3, _c(st).ravel(), _c(u), y)
    return y


def stencil_residual(n, st, u, f):
    n3, N = _n3(n)
    r = np.zeros((3, N))
    lib().orc_stencil_residual(n3, _c(st).ravel(), _c(u), _c(f), r)
    return r


def stencil_gs(n, st, u, f, omega=1.0, reverse=False):
    n3, N = _n3(n)
    u = _c(u).copy()
    lib().orc_stencil_gs(n3, _c(st).ravel(), u, _c(f), float(omega), int(reverse))
    return u


def restrict(nf, rf):
    n3, N = _n3(nf)
    fc = np.zeros((3, N // 8))
    lib().orc_restrict(n3, _c(rf), fc)
    return fc


def prolong_add(nf, uc, uf, zero_pin=True):
    n3, N = _n3(nf)
    uf = _c(uf).copy()
    lib().orc_prolong_add(n3, _c(uc), uf, int(zero_pin))
    return uf


def galerkin_fine(nf, KE, w, clamp=1e-4):
    n3, N = _n3(nf)
    st = np.zeros((N // 8, 27, 3, 3))
    lib().orc_galerkin_fine(n3, _c(KE).ravel(), _c(w).ravel(), st.reshape(-1), float(clamp))
    return st


def galerkin_coarse(nf, stf, clamp=1e-4):
    n3, N = _n3(nf)
    st = np.zeros((N // 8, 27, 3, 3))
    lib().orc_galerkin_coarse(n3, _c(stf).ravel(), st.reshape(-1), float(clamp))
    return st


def dense_from_stencil(n, st):
    n3, N = _n3(n)
    A = np.zeros((3 * N, 3 * N))
    lib().orc_dense_from_stencil(n3, _c(st).ravel(), A.reshape(-1))
    return A


def dense_fine(n, KE, w):
    n3, N = _n3(n)
    A = np.zeros((3 * N, 3 * N))
    lib().orc_dense_fine(n3, _c(KE).ravel(), _c(w).ravel(), A.reshape(-1))
    return A


def elastic_matrix(n, KE, w, chi):
    n3, N = _n3(n)
    C = np.zeros(36)
    lib().orc_elastic_matrix(n3, _c(KE).ravel(), _c(w).ravel(), _c(chi).ravel(), C)
    return C.reshape(6, 6)


def sensitivity(n, KE, chi, dC):
    n3, N = _n3(n)
    s = np.zeros(N)
    lib().orc_sensitivity(n3, _c(KE).ravel(), _c(chi).ravel(), _c(dC).ravel(), s)
    return s


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [o[i] for i in range(4)]


SYM = {"reflect3": 0, "reflect6": 1, "rotate3": 2, "none": 3}


def symmetrize(n, field, sym):
    s = SYM[sym] if isinstance(sym, str) else int(sym)
    if field.dtype == np.float32:
        inp = np.ascontiguousarray(field).ravel()
        out = np.zeros_like(inp)
        lib().orc_symmetrize_f32(int(n), s, inp, out)
    else:
        inp = _c(field).ravel()
        out = np.zeros_like(inp)
        lib().orc_symmetrize_f64(int(n), s, inp, out)
    return out.reshape(field.shape)


def oc_filter(n, sens, rho, r=2.0, periodic=False):
    n3, N = _n3(n)
    out = np.zeros(N)
    lib().orc_oc_filter(n3, _c(sens).ravel(), _c(rho).ravel(), float(r), int(periodic), out)
    return out


def oc_update(sens, rho, vol, min_rho=1e-3, step=0.05, damp=0.5):
    s = np.ascontiguousarray(sens, dtype=np.float32).ravel()
    r = np.ascontiguousarray(rho, dtype=np.float32).ravel().copy()
    it = lib().orc_oc_update(s.size, s, r, float(vol), float(min_rho), float(step), float(damp))
    return r, it


def mma_update(state, iter_, df0dx, fval, dfdx, a0=1.0, a=None, c=None, d=None, move=1.0):
    """state: dict with x, xmin, xmax, xold1, xold2, low, upp (float64 arrays, updated in place)."""
    dfdx = _c(dfdx)
    m = dfdx.shape[0] if dfdx.ndim == 2 else 1
    n = state["x"].size
    a = _c(np.zeros(m) if a is None else a)
    c = _c(np.full(m, 1000.0) if c is None else c)
    d = _c(np.ones(m) if d is None else d)
    return lib().orc_mma_update(m, n, int(iter_), state["x"], _c(state["xmin"]), _c(state["xmax"]),
                                state["xold1"], state["xold2"], _c(df0dx), _c(fval), dfdx.ravel(),
                                state["low"], state["upp"], float(a0), a, c, d, float(move))


def randc_density(n, seed, vol, sym="reflect6"):
    out = np.zeros(int(n) ** 3, dtype=np.float32)
    lib().orc_randc_density(int(n), int(seed), float(vol), SYM[sym] if isinstance(sym, str) else int(sym), out)
    return out


def randc_coeffs(seed):
    alpha = np.zeros(512)
    lib().orc_randc_coeffs(int(seed), alpha)
    return alpha


def randc_field(n, alpha, vol, sym="none"):
    out = np.zeros(int(n) ** 3, dtype=np.float32)
    lib().orc_randc_field(int(n), _c(alpha), float(vol), SYM[sym] if isinstance(sym, str) else int(sym), out)
    return out


class MG:
    """FP64 multigrid hierarchy of one cubic design (orc_mg_*)."""

    def __init__(self, n, E=10.0, nu=0.3, emin=1e-9):
        n3, N = _n3(n)
        self.h = lib().orc_mg_create(n3, float(E), float(nu), float(emin))
        if not self.h:
            raise ValueError("unsupported resolution %r" % (n,))
        self.n = int(n)
        self.N = N
        buf = (ctypes.c_int * 16)()
        self.nl = lib().orc_mg_levels(self.h, buf)
        self.sizes = [buf[i] for i in range(self.nl)]
        self.KE = ke(E, nu)
        self.emin = emin

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mg_destroy(self.h)
            self.h = None

    def set_density(self, x):
        rc = lib().orc_mg_set_density(self.h, _c(x).ravel())
        if rc != 0:
            raise RuntimeError("coarsest Cholesky failed")
        self.w = self.emin + _c(x).ravel()

    def stencil(self, level):
        m = self.sizes[level] ** 3
        out = np.zeros((m, 27, 3, 3))
        lib().orc_mg_get_stencil(self.h, int(level), out.reshape(-1))
        return out

    def vcycle(self, u, f, omega=1.0, symmetric=False):
        u = _c(u).copy()
        rn = lib().orc_mg_vcycle(self.h, u, _c(f), float(omega), int(symmetric))
        return u, rn

    def solve(self, f, u0=None, tol=1e-2, maxit=200, mode="pcg", omega=1.0):
        u = np.zeros((3, self.N)) if u0 is None else _c(u0).copy()
        rel = ctypes.c_double()
        it = ctypes.c_int()
        fn = lib().orc_mg_solve_pcg if mode == "pcg" else lib().orc_mg_solve_mg
        code = fn(self.h, u, _c(f), float(tol), int(maxit), float(omega), ctypes.byref(rel), ctypes.byref(it))
        return u, code, rel.value, it.value

    def homogenize(self, tol=1e-2, maxit=200, mode="pcg", chi=None):
        warm = chi is not None
        chi = np.zeros((6, 3, self.N)) if chi is None else _c(chi).copy()
        C = np.zeros(36)
        its = (ctypes.c_int * 6)()
        code = lib().orc_mg_homogenize(self.h, float(tol), int(maxit), 1 if mode == "pcg" else 0, int(warm),
                                       chi, C, its)
        return C.reshape(6, 6), chi, code, [its[i] for i in range(6)]
